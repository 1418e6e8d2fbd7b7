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
 * orc_bl_state / orc_bl_walk restate the same turn for the variants whose
 * cavity state depends on the particles: the bin-by-bin beam-induced
 * voltage of MPBunch::BunchMomentumUpdateDueToRFMode
 * (src/MPBunch.cpp:1419-1514, rfResExciteIntability), the time to the next
 * bunch from the bunches' z extents (MPBeam::GetTimeDisToNextBunch,
 * src/MPBeam.cpp:2716-2752), the cavity samples at the buckets
 * (Resonator::GetResonatorInfoAtNTrf, src/Resonator.cpp:263-275) and the
 * direct-feedback kick (Bunch::GetLongiKickDueToCavFB, src/Bunch.cpp:431-449).
 * Pinned in tests/test_beamload.py: no beam (the kick is the ideal cavity's
 * at the bin centres), the one-cavity phase that restores U0, the
 * generator's fixed point, the periodic steady state of a passive cavity
 * driven by one bunch and by an uneven train; for the walk, the stable
 * variant against the tables, the bin-by-bin voltage of a passive cavity
 * against its closed-form sum over the earlier bins, and the feedback kick
 * against the sampled cavity voltage.
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

/* ---------------------------------------------------------------------
 * The cavity state carried from turn to turn (ORC_BLS doubles per cavity):
 * mode, type, harm, cold, exciteInstability, dirFB, rs, q0, coupling,
 * detune, ql, fre, psi, tF, Re/Im of the required voltage, of vbAccum, of
 * the generator voltage and of the generator current, and the direct
 * feedback's gain, phase shift and delay. */
enum { S_MODE, S_TYPE, S_HARM, S_COLD, S_EXC, S_DFB, S_RS, S_Q0, S_BETA, S_DET, S_QL, S_FRE, S_PSI, S_TF,
       S_REQ, S_VB = S_REQ + 2, S_GEN = S_VB + 2, S_IG = S_GEN + 2, S_FBG = S_IG + 2, S_FBP, S_FBD, S_N };

int orc_bl_state_size(void) { return S_N; }

/* The initial state, as orc_bl_tables sets it up: ip[6][nres] = mode,
 * type, harm, cold, excite, dirFB; dp[9][nres] = rs, q0, coupling, detune,
 * vol, phase, feedback gain, phase shift, delay. */
void orc_bl_state(int nres, const int *ip, const double *dp, int harmonic, double f0, double u0, int nb,
                  const double *ne, const int *gap, double *state)
{
    const int H = harmonic;
    orc_cav *cv = (orc_cav *)calloc((size_t)nres, sizeof(orc_cav));
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
        const double complex kick_aver = kickAccum / (double)nb;
        r->gen = r->type == 0 ? 0 : r->req - r->vb - kick_aver;
        if (r->cold || r->mode == 0) r->vb = 0;
        if (r->type == 0) {
            r->ig = 0;
        } else {
            const double argVgr = carg(r->gen) - r->psi;
            const double absVgr = cabs(r->gen) / fabs(cos(r->psi));
            r->ig = absVgr * cexp(I * argVgr) * (1.0 + r->beta) / r->rs;
        }
        double *st = state + (size_t)i * S_N;
        st[S_MODE] = r->mode; st[S_TYPE] = r->type; st[S_HARM] = r->harm; st[S_COLD] = r->cold;
        st[S_EXC] = ip[4 * nres + i]; st[S_DFB] = ip[5 * nres + i];
        st[S_RS] = r->rs; st[S_Q0] = r->q0; st[S_BETA] = r->beta; st[S_DET] = r->detune;
        st[S_QL] = r->ql; st[S_FRE] = r->fre; st[S_PSI] = r->psi; st[S_TF] = r->tf;
        st[S_REQ] = creal(r->req); st[S_REQ + 1] = cimag(r->req);
        st[S_VB] = creal(r->vb); st[S_VB + 1] = cimag(r->vb);
        st[S_GEN] = creal(r->gen); st[S_GEN + 1] = cimag(r->gen);
        st[S_IG] = creal(r->ig); st[S_IG + 1] = cimag(r->ig);
        st[S_FBG] = dp[6 * nres + i]; st[S_FBP] = dp[7 * nres + i]; st[S_FBD] = dp[8 * nres + i];
    }
    free(cv);
    free(harm);
}

/* One turn of MPBeam::BeamMomtumUpdateDueToRFTest (src/MPBeam.cpp:2624-2706)
 * on the whole train after the map and the skew quadrupole: every bunch's z
 * extent (GetZMinMax, padded by 1e-6) and its time to the next bunch, then,
 * cavity by cavity and bunch by bunch, the kick of the stable or of the
 * bin-by-bin variant, the samples at the bunch's buckets with the
 * generator's dynamics bucket by bucket, and the decay of the beam-induced
 * voltage to the next bunch (bunchGap tRF in the stable variant, the time to
 * the next bunch in the other); then the direct-feedback kick of every
 * particle, lost ones included.  bharm[b]: the bunch's harmonic number;
 * macro_charge[b]: electrons per macroparticle; ne[b]: electrons per bunch.
 * info[nb][nres][8] as orc_bl_tables; fb[nb][nres] the feedback kicks
 * [V] (0 without feedback). */
void orc_bl_walk(double *state, int nres, orc_bunch *bs, int nb, const int *gap, const double *ne,
                 const double *macro_charge, int nbins, int harmonic, double f0, double t0, double rbeta,
                 double energy, double *info, double *fb)
{
    const int H = harmonic;
    const double tRF = t0 / H, fRF = f0 * H;
    double *zmin = (double *)malloc(sizeof(double) * (size_t)nb), *zmax = (double *)malloc(sizeof(double) * (size_t)nb);
    double *tnext = (double *)malloc(sizeof(double) * (size_t)nb);
    double complex *vfb = (double complex *)calloc((size_t)nb * nres, sizeof(double complex));
    int *cnt = (int *)malloc(sizeof(int) * (size_t)nbins);
    for (int b = 0; b < nb; b++) orc_zminmax(bs + b, zmin + b, zmax + b);
    for (int b = 0; b < nb; b++) {
        const int nx = b < nb - 1 ? b + 1 : 0;
        tnext[b] = gap[b] * tRF + (zmin[b] - zmax[nx]) / ORC_CLIGHT / rbeta;
    }
    for (int q = 0; q < nb * nres * 8; q++) info[q] = 0;
    for (int q = 0; q < nb * nres; q++) fb[q] = 0;
    for (int j = 0; j < nres; j++) {
        double *st = state + (size_t)j * S_N;
        if ((int)st[S_MODE] != 1) continue;
        const int exc = (int)st[S_EXC] != 0;
        const double resHarm = st[S_HARM], resFre = st[S_FRE], tF = st[S_TF];
        const double rs = st[S_RS], q0 = st[S_Q0], ql = st[S_QL], det = st[S_DET];
        double complex vb = st[S_VB] + I * st[S_VB + 1], gen = st[S_GEN] + I * st[S_GEN + 1];
        const double complex ig = st[S_IG] + I * st[S_IG + 1], req = st[S_REQ] + I * st[S_REQ + 1];
        for (int b = 0; b < nb; b++) {
            orc_bunch *B = bs + b;
            const double dzBin = (zmax[b] - zmin[b]) / nbins;
            const double dtBin = dzBin / ORC_CLIGHT / rbeta;
            for (int k = 0; k < nbins; k++) cnt[k] = 0;
            for (int64_t i = 0; i < B->n; i++) {
                if (B->lost[i] != 0) continue;
                cnt[(int)floor((zmax[b] - B->z[i]) / dzBin)]++;
            }
            double cavR = 0, cavI = 0, genR = 0, genI = 0, indR = 0, indI = 0;
            double complex self = 0;
            int nin = 0;
            if (!exc) {
                /* BunchMomentumUpdateDueToRFModeStable */
                const double complex vb0 = (-1 * 2 * PI * resFre * rs / q0 + 0.0 * I) * ne[b] * ORC_ECHARGE;
                const double complex cav = vb + gen;
                for (int k = 0; k < nbins; k++) {
                    const double zc = zmax[b] - k * dzBin - dzBin / 2.0;
                    const double complex kick = cav * cexp(-I * zc / ORC_CLIGHT / rbeta * 2. * PI * resHarm * fRF);
                    for (int64_t i = 0; i < B->n; i++) {
                        if (B->lost[i] != 0) continue;
                        if ((int)floor((zmax[b] - B->z[i]) / dzBin) != k) continue;
                        B->dp[i] += creal(kick) / energy / pow(rbeta, 2);
                        B->dp[i] += creal(vb0) / 2.0 / energy / pow(rbeta, 2);
                    }
                    nin += cnt[k];
                    self += vb0 / 2.0 * (double)cnt[k];
                    cavR += creal(cav) * cnt[k]; cavI += cimag(cav) * cnt[k];
                    genR += creal(gen) * cnt[k]; genI += cimag(gen) * cnt[k];
                    indR += creal(vb + vb0) * cnt[k]; indI += cimag(vb + vb0) * cnt[k];
                }
                vb += vb0;
                double *f = info + ((size_t)b * nres + j) * 8;
                f[0] = cavR / nin; f[1] = cavI / nin; f[2] = indR / nin; f[3] = indI / nin;
                f[4] = genR / nin; f[5] = genI / nin; f[6] = creal(self / (double)nin); f[7] = cimag(self / (double)nin);
            } else {
                /* BunchMomentumUpdateDueToRFMode: bin by bin, head first */
                const double deltaL = dtBin / tF, cPsi = 2.0 * PI * resFre * dtBin;
                for (int k = 0; k < nbins; k++) {
                    const double zc = zmax[b] - k * dzBin - dzBin / 2.0;
                    const double complex genV = gen * cexp(-I * zc / ORC_CLIGHT / rbeta * 2. * PI * resHarm * fRF);
                    const double complex cav = vb + genV;
                    const double complex vb0 =
                        (-1 * 2 * PI * resFre * rs / q0 + 0.0 * I) * macro_charge[b] * (double)cnt[k] * ORC_ECHARGE;
                    for (int64_t i = 0; i < B->n; i++) {
                        if (B->lost[i] != 0) continue;
                        if ((int)floor((zmax[b] - B->z[i]) / dzBin) != k) continue;
                        B->dp[i] += creal(cav) / energy / pow(rbeta, 2);
                        B->dp[i] += creal(vb0) / 2.0 / energy / pow(rbeta, 2);
                    }
                    nin += cnt[k];
                    self += vb0 / 2.0 * (double)cnt[k];
                    cavR += creal(cav) * cnt[k]; cavI += cimag(cav) * cnt[k];
                    genR += creal(genV) * cnt[k]; genI += cimag(genV) * cnt[k];
                    indR += creal(vb + vb0) * cnt[k]; indI += cimag(vb + vb0) * cnt[k];
                    vb += vb0;
                    vb *= exp(-deltaL) * cexp(I * cPsi);
                }
                self /= (double)nin;
                double *f = info + ((size_t)b * nres + j) * 8;
                const double complex ind = (indR / nin + I * (indI / nin)) + self;
                f[0] = cavR / nin; f[1] = cavI / nin; f[2] = creal(ind); f[3] = cimag(ind);
                f[4] = genR / nin; f[5] = genI / nin; f[6] = creal(self); f[7] = cimag(self);
            }
            /* the cavity at the bunch's buckets: the sample at its own
             * bucket (k = 0) is the one the feedback reads */
            for (int k = 0; k < gap[b]; k++) {
                if (k == 0) {
                    const double dt = k * tRF - zmin[b] / ORC_CLIGHT;
                    const double complex vB = vb * exp(-dt / tF) * cexp(I * (2.0 * PI * resFre * dt));
                    const double complex vCav = vB + gen;
                    vfb[(size_t)b * nres + j] = -(vCav - req);
                }
                /* Resonator::ResonatorDynamics(tRF) */
                const double tB = tRF;
                const double sigma = 2.0 * PI * resFre / (2.0 * ql);
                const double deltaOmega = 2.0 * PI * det;
                gen *= exp(-sigma * tB) * cexp(I * deltaOmega * tB);
                const double alpha = deltaOmega * exp(-sigma * tB) * sin(deltaOmega * tB) -
                                     sigma * exp(-sigma * tB) * cos(deltaOmega * tB) + sigma;
                const double beta = sigma * exp(-sigma * tB) * sin(deltaOmega * tB) +
                                    deltaOmega * exp(-sigma * tB) * cos(deltaOmega * tB) - deltaOmega;
                const double coefB = 2 * PI * resFre / 2.0 * rs / q0 / (pow(sigma, 2) + pow(deltaOmega, 2));
                gen += coefB * (alpha * creal(ig) + beta * cimag(ig)) + I * (coefB * (-beta * creal(ig) + alpha * cimag(ig)));
            }
            /* Resonator::GetBeamInducedVol */
            const double T = exc ? tnext[b] : gap[b] * tRF;
            vb *= exp(-T / tF) * cexp(I * (2.0 * PI * resFre * T));
        }
        st[S_VB] = creal(vb); st[S_VB + 1] = cimag(vb);
        st[S_GEN] = creal(gen); st[S_GEN + 1] = cimag(gen);
    }
    /* Bunch::GetLongiKickDueToCavFB on every particle */
    for (int j = 0; j < nres; j++) {
        const double *st = state + (size_t)j * S_N;
        if ((int)st[S_MODE] != 1 || (int)st[S_DFB] != 1) continue;
        for (int b = 0; b < nb; b++) {
            const double complex c = vfb[(size_t)b * nres + j];
            const double absFB = cabs(c) * st[S_FBG];
            const double argFB = carg(c) + st[S_FBP] + st[S_FBD] * st[S_FRE] * 2 * PI;
            const double complex k = absFB * cexp(I * argFB);
            fb[(size_t)b * nres + j] = creal(k);
            orc_bunch *B = bs + b;
            for (int64_t i = 0; i < B->n; i++) B->dp[i] += creal(k) / energy / pow(rbeta, 2);
        }
    }
    free(zmin); free(zmax); free(tnext); free(vfb); free(cnt);
}
