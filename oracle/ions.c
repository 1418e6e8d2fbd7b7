/*
 * ions.c -- oracle restatement of the ion cloud at one interaction point
 * and of the two-way Bassetti-Erskine interaction of one bunch passage.
 * Test infrastructure only (see cts_oracle.h).
 */
#include "cts_oracle.h"

#include <math.h>
#include <stdlib.h>

/* New ions for one passage: pairs of standard normals from the project
 * stream (Philox counter {ion, passage | species << 24, turn, 2^31 | call}),
 * scaled as std::normal_distribution does (n*sigma + mean) and rejected
 * outside the 2-sigma ellipse as in src/LatticeInterActionPoint.cpp:1000-1017.
 * Each ion retries on its own counter, at most 64 attempts (the chance of
 * 64 rejections is 0.135^64). */
void orc_ion_generate(uint64_t seed, uint32_t turn, uint32_t passage, int species,
                      int n_new, double xmean, double ymean, double rms_x, double rms_y,
                      double *gx, double *gy)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int j = 0; j < n_new; j++) {
        double tx = xmean, ty = ymean;
        for (uint32_t t = 0; t < 64u; t++) {
            uint32_t ctr[4] = {(uint32_t)j, passage | ((uint32_t)species << 24), turn,
                               0x80000000u | (t >> 1)};
            uint32_t o[4];
            orc_philox4x32_10(ctr, key, o);
            float n0, n1;
            if ((t & 1u) == 0) orc_box_muller(o[0], o[1], &n0, &n1);
            else orc_box_muller(o[2], o[3], &n0, &n1);
            tx = (double)n0 * rms_x + xmean;
            ty = (double)n1 * rms_y + ymean;
            if (pow((tx - xmean) / rms_x, 2) + pow((ty - ymean) / rms_y, 2) > 4.E0) continue;
            break;
        }
        gx[j] = tx;
        gy[j] = ty;
    }
}

/* IonRMSCal(k, p), src/LatticeInterActionPoint.cpp:1113-1155.  With one
 * ion the mean is left as it was and the rms is uncentred, as there. */
void orc_ion_rms(orc_ion_cloud *c, int p)
{
    const int64_t n = c->n[p];
    double x2 = 0, y2 = 0;
    if (n == 0) {
        x2 = 0; y2 = 0;
    } else if (n == 1) {
        x2 = pow(c->x[p][0], 2);
        y2 = pow(c->y[p][0], 2);
    } else {
        double sx = 0.0, sy = 0.0;
        for (int64_t i = 0; i < n; i++) sx += c->x[p][i];
        for (int64_t i = 0; i < n; i++) sy += c->y[p][i];
        c->mean_x[p] = sx / n;
        c->mean_y[p] = sy / n;
        for (int64_t i = 0; i < n; i++) {
            x2 += pow(c->x[p][i] - c->mean_x[p], 2);
            y2 += pow(c->y[p][i] - c->mean_y[p], 2);
        }
    }
    if (n != 0) {
        x2 = x2 / n;
        y2 = y2 / n;
        c->rms_x[p] = sqrt(x2);
        c->rms_y[p] = sqrt(y2);
    } else {
        c->rms_x[p] = 0;
        c->rms_y[p] = 0;
    }
}

/* MPBeam::SSBeamIonEffectOneInteractionPoint for one bunch
 * (src/MPBeam.cpp:2592-2608) with the callees of
 * src/LatticeInterActionPoint.cpp:974-1210 and src/MPBunch.cpp:1094-1210. */
int orc_ion_passage(orc_ion_cloud *c, orc_bunch *b, const orc_bunch_view *v,
                    double size_x_max, double size_y_max, uint64_t seed,
                    uint32_t turn, uint32_t passage, double *cloud_out)
{
    /* GetIonNumberPerInterAction (974-981) and IonGenerator (985-1019) */
    for (int p = 0; p < c->nspec; p++) {
        const double line_density = c->cross_section[p] * c->pressure[p] / c->temperature /
                                    ORC_KBOLTZ * v->n_e;
        const double ion_number = line_density * c->length;
        c->macro_charge[p] = ion_number / c->n_new;
        /* IonsUpdate (1021-1050): append, then the capacity check */
        const int64_t n0 = c->n[p];
        if (n0 + c->n_new > c->cap) {
            c->n[p] = n0 + c->n_new;
            return 1;
        }
        orc_ion_generate(seed, turn, passage, p, c->n_new, v->xmean, v->ymean, v->rms_x,
                         v->rms_y, c->x[p] + n0, c->y[p] + n0);
        for (int j = 0; j < c->n_new; j++) {
            c->vx[p][n0 + j] = 0.0;
            c->vy[p][n0 + j] = 0.0;
            c->fx[p][n0 + j] = 0.0;
            c->fy[p][n0 + j] = 0.0;
        }
        c->n[p] = n0 + c->n_new;
    }
    /* IonRMSCal(k) (1053-1110): only the per-species part feeds the kicks */
    for (int p = 0; p < c->nspec; p++) orc_ion_rms(c, p);

    /* SSIonBunchInteraction (1) ions kicked by the bunch (1113-1160) */
    for (int p = 0; p < c->nspec; p++) {
        const double coeffI = 2.0 * v->n_e * ORC_RE * ORC_EMASS_EV / ORC_IONMASS_EV /
                              c->mass_number[p] * ORC_CLIGHT;
        for (int64_t j = 0; j < c->n[p]; j++) {
            double fx, fy;
            const double posx = c->x[p][j] - v->xmean;
            const double posy = c->y[p][j] - v->ymean;
            orc_ion_field(posx, posy, v->rms_x, v->rms_y, &fx, &fy);
            c->fx[p][j] = coeffI * fx;
            c->fy[p][j] = coeffI * fy;
        }
    }
    if (cloud_out != NULL) {
        for (int p = 0; p < c->nspec; p++) {
            cloud_out[5 * p + 0] = c->macro_charge[p] * c->n[p];
            cloud_out[5 * p + 1] = c->mean_x[p];
            cloud_out[5 * p + 2] = c->mean_y[p];
            cloud_out[5 * p + 3] = c->rms_x[p];
            cloud_out[5 * p + 4] = c->rms_y[p];
        }
    }
    /* (2) electrons kicked by the cloud (1165-1206) */
    for (int64_t i = 0; i < b->n; i++) {
        if (b->lost[i] != 0) continue;
        b->fx_ion[i] = 0.E0;
        b->fy_ion[i] = 0.E0;
        for (int p = 0; p < c->nspec; p++) {
            const double nI0 = c->macro_charge[p] * c->n[p];
            const double coeffE = 2.0 * nI0 * ORC_RE / v->gamma;
            double fx, fy;
            const double posx = b->x[i] - c->mean_x[p];
            const double posy = b->y[i] - c->mean_y[p];
            orc_ion_field(posx, posy, c->rms_x[p], c->rms_y[p], &fx, &fy);
            b->fx_ion[i] += coeffE * fx;
            b->fy_ion[i] += coeffE * fy;
        }
    }
    /* BunchTransferDueToIon (src/Bunch.cpp:213-220): every particle */
    for (int64_t i = 0; i < b->n; i++) {
        b->px[i] += b->fx_ion[i];
        b->py[i] += b->fy_ion[i];
    }
    /* IonTransferDueToBunch (1158-1210) */
    for (int p = 0; p < c->nspec; p++) {
        for (int64_t i = 0; i < c->n[p]; i++) {
            c->vx[p][i] += c->fx[p][i];
            c->vy[p][i] += c->fy[p][i];
            c->x[p][i] += c->vx[p][i] * c->circ / c->harm * v->gap / ORC_CLIGHT;
            c->y[p][i] += c->vy[p][i] * c->circ / c->harm * v->gap / ORC_CLIGHT;
        }
    }
    const double bx = c->ap_x > c->loss_boundary * size_x_max ? c->loss_boundary * size_x_max : c->ap_x;
    const double by = c->ap_y > c->loss_boundary * size_y_max ? c->loss_boundary * size_y_max : c->ap_y;
    for (int p = 0; p < c->nspec; p++) {
        int64_t w = 0;
        for (int64_t i = 0; i < c->n[p]; i++) {
            if (fabs(c->x[p][i]) > bx || fabs(c->y[p][i]) > by) continue;
            c->x[p][w] = c->x[p][i]; c->y[p][w] = c->y[p][i];
            c->vx[p][w] = c->vx[p][i]; c->vy[p][w] = c->vy[p][i];
            c->fx[p][w] = c->fx[p][i]; c->fy[p][w] = c->fy[p][i];
            w++;
        }
        c->n[p] = w;
    }
    return 0;
}
