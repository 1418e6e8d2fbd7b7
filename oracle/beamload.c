/* beamload.c -- CPU oracle (test infrastructure only) of the beam-loaded
 * (rfMode 1) cavities in the reference's stable variant:
 * CavityResonator::Initial (src/CavityResonator.cpp:36-176),
 * MPBeam::InitialcavityResonator's initial beam-induced and generator
 * voltages (src/MPBeam.cpp:236-465), Resonator::GetInitialResonatorGenIg
 * (src/Resonator.cpp:38-57), the per-bunch loop of
 * MPBeam::BeamMomtumUpdateDueToRFTest (src/MPBeam.cpp:2640-2690) with
 * Resonator::ResonatorDynamics and GetBeamInducedVol
 * (src/Resonator.cpp:205-261), and the particle kick of
 * MPBunch::BunchMomentumUpdateDueToRFModeStable (src/MPBunch.cpp:1323-1417),
 * written with C99 complex arithmetic in the reference's operation order.
 * Pinned in tests/test_beamload.py: no beam (the kick is the ideal cavity's
 * at the bin centres), the one-cavity phase that restores U0, the
 * generator's fixed point, and the periodic steady state of a passive cavity
 * driven by one bunch and by an uneven train.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "cts_oracle.h"

#define PI 3.14159265358979323846

typedef struct {
    int mode, type, harm, cold;
    double rs, q0, beta, detune, ql, fre, psi, tf;
    double complex req, vb, gen, ig;
} orc_cav;

/* Tables for turns 0 .. nturns-1: dev[n][b][q][3] = Re V, Im V, Re V_b0
 * seen by bunch b from the q-th rfMode-1 cavity; info[n][b][i][8] =
 * cavity, induced, generator and self-loss voltages (re, im) from cavity i.
 * Cavity arrays per cavity: mode, type, harm, cold (int[4][nres]) and rs,
 * q0, coupling, detune, vol, phase (double[6][nres]). */
int orc_bl_tables(int nres, const int *ip, const double *dp, int harmonic, double f0, double t0, double u0,
                  double rbeta, int nb, const double *ne, const int *gap, int nturns, double *dev, double *info)
{
    orc_cav *cv = (orc_cav *)calloc((size_t)nres, sizeof(orc_cav));
    const int H = harmonic;
    for (int i = 0; i < nres; i++) {
        orc_cav *r = cv + i;
        r->mode = ip[i]; r->type = ip[nres + i]; r->harm = ip[2 * nres + i]; r->cold = ip[3 * nres + i];
        r->rs = dp[i]; r->q0 = dp[nres + i]; r->beta = dp[2 * nres + i]; r->detune = dp[3 * nres + i];
        r->ql = r->q0 / (1 + r->beta);
        r->fre = r->harm * H * f0 + r->detune;
        const double tan_psi = r->ql * (r->fre / (r->harm * H * f0) - r->harm * H * f0 / r->fre);
        r->psi = atan(tan_psi);
        r->tf = 2 * r->ql / (2 * PI * r->fre);
    }
    for (int i = 0; i < nres; i++) {
        double ph = dp[5 * nres + i];
        const double vol = dp[4 * nres + i];
        if (nres == 1) ph = PI - asin(u0 / dp[4 * nres]) - PI / 2.;
        cv[i].req = vol * cos(ph) + I * (vol * sin(ph));
    }
    int *harm = (int *)calloc((size_t)nb, sizeof(int));
    for (int b = 1; b < nb; b++) harm[b] = harm[b - 1] + gap[b - 1];
    double complex *kick_aver = (double complex *)calloc((size_t)nres, sizeof(double complex));
    for (int i = 0; i < nres; i++) {
        orc_cav *r = cv + i;
        const int nt = (int)ceil(100 * r->tf * f0);
        double complex vbAccum = 0, aver2 = 0, kickAccum = 0;
        for (int n = 0; n < nt; n++) {
            int k = 0;
            for (int j = 0; j < H; j++) {
                double complex vb0 = 0;
                if (k < nb && harm[k] == j) {
                    vb0 = (-1 * r->fre * 2 * PI * r->rs / r->q0 + 0.0 * I) * ne[k] * ORC_ECHARGE;
                    k++;
                    if (n == nt - 1) kickAccum += vb0 / 2.0;
                }
                vbAccum += vb0;
                const double tB = 1.0 / f0 / H;
                const double deltaL = tB / r->tf;
                const double cPsi = deltaL * tan(r->psi);
                vbAccum = vbAccum * exp(-deltaL) * cexp(I * cPsi);
                if (n == nt - 1) aver2 += vbAccum;
            }
        }
        r->vb = aver2 / (double)H;
        kick_aver[i] = kickAccum / (double)nb;
    }
    for (int i = 0; i < nres; i++) {
        orc_cav *r = cv + i;
        r->gen = r->type == 0 ? 0 : r->req - r->vb - kick_aver[i];
        if (r->cold || r->mode == 0) r->vb = 0;
        if (r->type == 0) {
            r->ig = 0;
        } else {
            const double argVgr = carg(r->gen) - r->psi;
            const double absVgr = cabs(r->gen) / fabs(cos(r->psi));
            const double complex vgr = absVgr * cexp(I * argVgr);
            r->ig = vgr * (1.0 + r->beta) / r->rs;
        }
    }
    int nbl = 0;
    for (int i = 0; i < nres; i++) nbl += cv[i].mode == 1;
    const double trf = t0 / H;
    for (int n = 0; n < nturns; n++) {
        int q = 0;
        for (int i = 0; i < nres; i++) {
            orc_cav *r = cv + i;
            if (r->mode != 1) continue;
            for (int b = 0; b < nb; b++) {
                const double complex vb0 = (-1 * 2 * PI * r->fre * r->rs / r->q0 + 0.0 * I) * ne[b] * ORC_ECHARGE;
                const double complex gen = r->gen, cav = r->vb + gen;
                double *d = dev + (((size_t)n * nb + b) * nbl + q) * 3;
                d[0] = creal(cav); d[1] = cimag(cav); d[2] = creal(vb0);
                double *f = info + (((size_t)n * nb + b) * nres + i) * 8;
                const double complex ind = r->vb + vb0, self = vb0 / 2.0;
                f[0] = creal(cav); f[1] = cimag(cav); f[2] = creal(ind); f[3] = cimag(ind);
                f[4] = creal(gen); f[5] = cimag(gen); f[6] = creal(self); f[7] = cimag(self);
                r->vb += vb0;
                for (int k = 0; k < gap[b]; k++) {
                    /* Resonator::ResonatorDynamics(tRF) */
                    const double tB = trf;
                    const double sigma = 2.0 * PI * r->fre / (2.0 * r->ql);
                    const double deltaOmega = 2.0 * PI * r->detune;
                    r->gen *= exp(-sigma * tB) * cexp(I * deltaOmega * tB);
                    const double alpha = deltaOmega * exp(-sigma * tB) * sin(deltaOmega * tB) -
                                         sigma * exp(-sigma * tB) * cos(deltaOmega * tB) + sigma;
                    const double beta = sigma * exp(-sigma * tB) * sin(deltaOmega * tB) +
                                        deltaOmega * exp(-sigma * tB) * cos(deltaOmega * tB) - deltaOmega;
                    const double coefB = 2 * PI * r->fre / 2.0 * r->rs / r->q0 / (pow(sigma, 2) + pow(deltaOmega, 2));
                    r->gen += coefB * (alpha * creal(r->ig) + beta * cimag(r->ig)) +
                              I * (coefB * (-beta * creal(r->ig) + alpha * cimag(r->ig)));
                }
                /* Resonator::GetBeamInducedVol(bunchGap * tRF) */
                const double T = gap[b] * trf;
                r->vb *= exp(-T / r->tf) * cexp(I * (2.0 * PI * r->fre * T));
            }
            q++;
        }
    }
    free(cv);
    free(harm);
    free(kick_aver);
    return nbl;
}

/* The kick of BunchMomentumUpdateDueToRFModeStable on the survivors of a
 * bunch: bins over [zmin, zmax] (GetZMinMax's padded range), head first;
 * delta += Re(V exp(-i z_bin / c / beta 2 pi h f_rf)) / E / beta^2 and
 * += Re(V_b0) / 2 / E / beta^2. */
void orc_bl_kick(orc_bunch *b, double vr, double vi, double vb0, double zmin, double zmax, int nbins, int res_harm,
                 double f_rf, double rbeta, double energy)
{
    const double dz = (zmax - zmin) / nbins;
    const double complex v = vr + I * vi;
    for (int64_t i = 0; i < b->n; i++) {
        if (b->lost[i] != 0) continue;
        const int k = (int)floor((zmax - b->z[i]) / dz);
        const double zc = zmax - k * dz - dz / 2.0;
        const double complex e = cexp(-I * zc / ORC_CLIGHT / rbeta * 2. * PI * (double)res_harm * f_rf);
        b->dp[i] += creal(v * e) / energy / pow(rbeta, 2);
        b->dp[i] += vb0 / 2.0 / energy / pow(rbeta, 2);
    }
}
