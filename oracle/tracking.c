/*
 * tracking.c -- oracle restatement of the per-particle maps, loss marking,
 * moments and short-range wake of the reference's multi-particle turn.
 * Test infrastructure only (see cts_oracle.h).
 */
#include "cts_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.141592653589793; /* 4*atan(1.E0), include/Global.h:33 */

/* gslcblas dgemm, NoTrans/NoTrans, alpha = 1, beta = 0 (GSL 2.8
 * cblas/source_gemm_r.h): C is cleared, then for each row i and each k with
 * A[i][k] != 0 the row k of B, scaled by A[i][k], is added in order. */
void orc_dgemm(int M, int N, int K, const double *A, const double *B, double *C)
{
    for (int i = 0; i < M * N; i++) C[i] = 0.0;
    for (int i = 0; i < M; i++) {
        for (int k = 0; k < K; k++) {
            const double temp = 1.0 * A[i * K + k];
            if (temp != 0.0) {
                for (int j = 0; j < N; j++) C[i * N + j] += temp * B[k * N + j];
            }
        }
    }
}

/* src/LatticeInterActionPoint.cpp:326-431 */
void orc_section_matrices(const orc_twiss *t1, const orc_twiss *t2,
                          double b1h1[36], double h1b1[36], double h2b2[36])
{
    double H1[36], iH1[36], iH2[36], B1[36], iB1[36], iB2[36];
    memset(H1, 0, sizeof H1); memset(iH1, 0, sizeof iH1); memset(iH2, 0, sizeof iH2);
    memset(B1, 0, sizeof B1); memset(iB1, 0, sizeof iB1); memset(iB2, 0, sizeof iB2);
    for (int i = 0; i < 6; i++) { H1[i * 7] = 1.0; iH1[i * 7] = 1.0; iH2[i * 7] = 1.0; }
#define M6(m, r, c) m[(r) * 6 + (c)]
    /* H1 (386-387) */
    M6(H1, 0, 5) = -t1->eta_x;  M6(H1, 1, 5) = -t1->etap_x;
    M6(H1, 2, 5) = -t1->eta_y;  M6(H1, 3, 5) = -t1->etap_y;
    M6(H1, 4, 0) = t1->etap_x;  M6(H1, 4, 1) = -t1->eta_x;
    M6(H1, 4, 2) = t1->etap_y;  M6(H1, 4, 3) = -t1->eta_y;
    /* InvH1 (389-391) */
    M6(iH1, 0, 5) = t1->eta_x;  M6(iH1, 1, 5) = t1->etap_x;
    M6(iH1, 2, 5) = t1->eta_y;  M6(iH1, 3, 5) = t1->etap_y;
    M6(iH1, 4, 0) = -t1->etap_x; M6(iH1, 4, 1) = t1->eta_x;
    M6(iH1, 4, 2) = -t1->etap_y; M6(iH1, 4, 3) = t1->eta_y;
    /* InvH2 (393-395) */
    M6(iH2, 0, 5) = t2->eta_x;  M6(iH2, 1, 5) = t2->etap_x;
    M6(iH2, 2, 5) = t2->eta_y;  M6(iH2, 3, 5) = t2->etap_y;
    M6(iH2, 4, 0) = -t2->etap_x; M6(iH2, 4, 1) = t2->eta_x;
    M6(iH2, 4, 2) = -t2->etap_y; M6(iH2, 4, 3) = t2->eta_y;
    /* B1 (399-401) */
    M6(B1, 0, 0) = 1 / sqrt(t1->beta_x); M6(B1, 1, 0) = t1->alpha_x / sqrt(t1->beta_x);
    M6(B1, 1, 1) = sqrt(t1->beta_x);
    M6(B1, 2, 2) = 1 / sqrt(t1->beta_y); M6(B1, 3, 2) = t1->alpha_y / sqrt(t1->beta_y);
    M6(B1, 3, 3) = sqrt(t1->beta_y);
    M6(B1, 4, 4) = 1 / sqrt(t1->beta_z); M6(B1, 5, 5) = sqrt(t1->beta_z);
    /* InvB1 (404-406) */
    M6(iB1, 0, 0) = sqrt(t1->beta_x); M6(iB1, 1, 0) = -t1->alpha_x / sqrt(t1->beta_x);
    M6(iB1, 1, 1) = 1 / sqrt(t1->beta_x);
    M6(iB1, 2, 2) = sqrt(t1->beta_y); M6(iB1, 3, 2) = -t1->alpha_y / sqrt(t1->beta_y);
    M6(iB1, 3, 3) = 1 / sqrt(t1->beta_y);
    M6(iB1, 4, 4) = sqrt(t1->beta_z); M6(iB1, 5, 5) = 1 / sqrt(t1->beta_z);
    /* InvB2 (409-411) */
    M6(iB2, 0, 0) = sqrt(t2->beta_x); M6(iB2, 1, 0) = -t2->alpha_x / sqrt(t2->beta_x);
    M6(iB2, 1, 1) = 1 / sqrt(t2->beta_x);
    M6(iB2, 2, 2) = sqrt(t2->beta_y); M6(iB2, 3, 2) = -t2->alpha_y / sqrt(t2->beta_y);
    M6(iB2, 3, 3) = 1 / sqrt(t2->beta_y);
    M6(iB2, 4, 4) = sqrt(t2->beta_z); M6(iB2, 5, 5) = 1 / sqrt(t2->beta_z);
#undef M6
    /* (413-415) */
    orc_dgemm(6, 6, 6, B1, H1, b1h1);
    orc_dgemm(6, 6, 6, iH1, iB1, h1b1);
    orc_dgemm(6, 6, 6, iH2, iB2, h2b2);
}

static inline void load6(const orc_bunch *b, int64_t i, double v[6])
{
    v[0] = b->x[i]; v[1] = b->px[i]; v[2] = b->y[i];
    v[3] = b->py[i]; v[4] = b->z[i]; v[5] = b->dp[i];
}
static inline void store6(orc_bunch *b, int64_t i, const double v[6])
{
    b->x[i] = v[0]; b->px[i] = v[1]; b->y[i] = v[2];
    b->py[i] = v[3]; b->z[i] = v[4]; b->dp[i] = v[5];
}

/* src/Bunch.cpp:625-755 (the output-only accPhaseAdv bookkeeping of
 * 715-733 is not restated). */
void orc_section_map(orc_bunch *b, const orc_map *m)
{
    const double weighX = m->phase_adv_x / (2 * PI * m->nux);
    const double weighY = m->phase_adv_y / (2 * PI * m->nuy);
    double R[36], X[6], X1[6];
    for (int64_t i = 0; i < b->n; i++) {
        if (b->lost[i] != 0) continue;
        const double dp = b->dp[i];
        const double x = b->x[i] - m->eta_x * dp;
        const double px = b->px[i] - m->etap_x * dp;
        const double y = b->y[i] - m->eta_y * dp;
        const double py = b->py[i] - m->etap_y * dp;
        const double ampX = (pow(x, 2) + pow(m->alpha_x * x + m->beta_x * px, 2)) / m->beta_x;
        const double ampY = (pow(y, 2) + pow(m->alpha_y * y + m->beta_y * py, 2)) / m->beta_y;
        const double dnux = (m->chrom_x * dp + m->adtx[0] * ampX + m->adtx[1] * ampY +
                             m->adtx[2] * pow(ampX, 2) / 2 + m->adtx[3] * pow(ampY, 2) / 2 +
                             m->adtx[4] * ampX * ampY) * weighX;
        const double dnuy = (m->chrom_y * dp + m->adty[0] * ampX + m->adty[1] * ampY +
                             m->adty[2] * pow(ampX, 2) / 2 + m->adty[3] * pow(ampY, 2) / 2 +
                             m->adty[4] * ampX * ampY) * weighY;
        const double phiX = m->phase_adv_x + 2 * PI * dnux;
        const double phiY = m->phase_adv_y + 2 * PI * dnuy;
        memset(R, 0, sizeof R);
        R[0] = cos(phiX);  R[1] = sin(phiX);  R[6] = -sin(phiX);  R[7] = cos(phiX);
        R[14] = cos(phiY); R[15] = sin(phiY); R[20] = -sin(phiY); R[21] = cos(phiY);
        R[28] = 1; R[35] = 1;
        load6(b, i, X);
        orc_dgemm(6, 1, 6, m->b1h1, X, X1);
        orc_dgemm(6, 1, 6, R, X1, X);
        orc_dgemm(6, 1, 6, m->h2b2, X, X1);
        store6(b, i, X1);
    }
}

/* src/Bunch.cpp:757-795 */
void orc_skew_quad(orc_bunch *b, double k)
{
    double S[36], X[6], X1[6];
    memset(S, 0, sizeof S);
    for (int i = 0; i < 6; i++) S[i * 7] = 1.0;
    S[1 * 6 + 2] = k;
    S[3 * 6 + 0] = k;
    for (int64_t i = 0; i < b->n; i++) {
        if (b->lost[i] != 0) continue;
        load6(b, i, X);
        orc_dgemm(6, 1, 6, S, X, X1);
        store6(b, i, X1);
    }
}

/* src/Bunch.cpp:453-481 with the cavity phasor of
 * src/CavityResonator.cpp:121-132.  The phase is the imaginary part of
 * -i*z/c/beta*2*pi*h*fRF evaluated left to right; exp of a complex with a
 * zero real part is (cos, sin) of it, and the complex product's real part is
 * Vr*cos - Vi*sin. */
void orc_rf_kick(orc_bunch *b, double vr, double vi, int res_harm, double f_rf,
                 double rbeta, double energy)
{
    for (int64_t i = 0; i < b->n; i++) {
        const double th = ((((-b->z[i] / ORC_CLIGHT) / rbeta) * 2.) * PI) * (double)res_harm * f_rf;
        const double c = cos(th), s = sin(th);
        const double re = vr * c - vi * s;
        b->dp[i] += re / energy / pow(rbeta, 2);
    }
}

/* src/Bunch.cpp:494-501 */
void orc_drift(orc_bunch *b, double circ, const double a[3])
{
    for (int64_t i = 0; i < b->n; i++) {
        const double d = b->dp[i];
        b->z[i] -= circ * (a[0] * d + a[1] * pow(d, 2) + a[2] * pow(d, 3));
    }
}

/* src/Bunch.cpp:483-491 */
void orc_energy_loss(orc_bunch *b, double u0, double energy, double rbeta)
{
    for (int64_t i = 0; i < b->n; i++) b->dp[i] -= u0 / energy / pow(rbeta, 2);
}

/* src/Bunch.cpp:940-1044 */
void orc_synrad(orc_bunch *b, const double b1h1_0[36], const double h1b1_0[36],
                const double tau[3], const double emit_nat[3], const double *noise)
{
    double lambda[3], coeff[3];
    for (int i = 0; i < 3; i++) lambda[i] = exp(-1.0 / tau[i]);
    coeff[0] = sqrt(1 - pow(lambda[0], 2)) * sqrt(emit_nat[0]);
    coeff[1] = sqrt(1 - pow(lambda[1], 2)) * sqrt(emit_nat[1]);
    coeff[2] = sqrt(1 - pow(lambda[2], 4)) * sqrt(emit_nat[2]);
    double X[6], NX[6];
    const int64_t n = b->n;
    for (int64_t i = 0; i < n; i++) {
        if (b->lost[i] != 0) continue;
        load6(b, i, X);
        orc_dgemm(6, 1, 6, b1h1_0, X, NX);
        NX[0] *= lambda[0];
        NX[1] *= lambda[0];
        NX[2] *= lambda[1];
        NX[3] *= lambda[1];
        NX[5] *= lambda[2] * lambda[2];
        if (noise != NULL && n != 1) {
            NX[0] += coeff[0] * noise[0 * n + i];
            NX[1] += coeff[0] * noise[1 * n + i];
            NX[2] += coeff[1] * noise[2 * n + i];
            NX[3] += coeff[1] * noise[3 * n + i];
            NX[5] += coeff[2] * noise[4 * n + i];
        }
        orc_dgemm(6, 1, 6, h1b1_0, NX, X);
        store6(b, i, X);
    }
}

/* src/Bunch.cpp:172-200.  Codes are never reset; the count includes
 * particles lost in earlier turns and double-counts particles outside both
 * limits, as the reference does. */
double orc_mark_lost(orc_bunch *b, double t0, int harm, double ap_x, double ap_y,
                     int64_t *count_out)
{
    int64_t count = 0;
    const double zlim = t0 * ORC_CLIGHT / harm / 2;
    for (int64_t i = 0; i < b->n; i++) {
        if (fabs(b->z[i]) > zlim) { b->lost[i] = 2; count++; }
        const double t = pow(b->x[i] / ap_x, 2) + pow(b->y[i] / ap_y, 2);
        if (t > 1) { b->lost[i] = 1; count++; }
    }
    if (count_out) *count_out = count;
    return (double)(b->n - count) / (double)b->n;
}

/* src/MPBunch.cpp:1860-1885 (the start value is z[0] even if particle 0 is
 * lost). */
void orc_zminmax(const orc_bunch *b, double *zmin, double *zmax)
{
    double lo = b->z[0], hi = b->z[0];
    for (int64_t i = 0; i < b->n; i++) {
        if (b->lost[i] != 0) continue;
        if (lo > b->z[i]) lo = b->z[i];
        if (hi < b->z[i]) hi = b->z[i];
    }
    *zmin = lo - 1.e-6;
    *zmax = hi + 1.e-6;
}

/* LU with partial pivoting, unblocked right-looking form (GSL 2.8
 * linalg/lu.c LU_decomp_L2: pivot = first max |a|, column scaled by the
 * reciprocal of the pivot, rank-1 update), determinant = sign * prod(diag)
 * (gsl_linalg_LU_det). Zero matrix -> 0 (Global.cpp:184). */
double orc_det(int n, const double *a_in)
{
    double a[16];
    int allzero = 1;
    for (int i = 0; i < n * n; i++) { a[i] = a_in[i]; if (a[i] != 0.0) allzero = 0; }
    if (allzero) return 0.0;
    int signum = 1;
    for (int j = 0; j < n; j++) {
        int piv = j;
        double amax = fabs(a[j * n + j]);
        for (int i = j + 1; i < n; i++) {
            if (fabs(a[i * n + j]) > amax) { amax = fabs(a[i * n + j]); piv = i; }
        }
        if (piv != j) {
            for (int c = 0; c < n; c++) {
                double t = a[j * n + c]; a[j * n + c] = a[piv * n + c]; a[piv * n + c] = t;
            }
            signum = -signum;
        }
        const double ajj = a[j * n + j];
        if (ajj == 0.0) break; /* singular: GSL reports an error; det is 0 */
        if (j < n - 1) {
            const double inv = 1.0 / ajj;
            for (int i = j + 1; i < n; i++) a[i * n + j] *= inv;
            for (int i = j + 1; i < n; i++) {
                const double tmp = -1.0 * a[i * n + j];
                for (int c = j + 1; c < n; c++) a[i * n + c] += tmp * a[j * n + c];
            }
        }
    }
    double det = signum;
    for (int i = 0; i < n; i++) det *= a[i * n + i];
    return det;
}

/* std::accumulate(first, last, 0) with an int start value: every partial
 * sum is converted back to int (src/MPBunch.cpp:522-525). */
static double accumulate_int(const double *v, int64_t n)
{
    int acc = 0;
    for (int64_t i = 0; i < n; i++) acc = (int)(acc + v[i]);
    return (double)acc;
}

/* GetEigenEmit, src/MPBunch.cpp:500-658 */
static void eigen_emit(const orc_bunch *b, double etax, double etaxp, double etay,
                       double etayp, orc_moments *o)
{
    const int64_t n = b->n;
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n);
    double *xp = malloc(sizeof(double) * n), *yp = malloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; i++) {
        x[i] = b->x[i] - etax * b->dp[i];
        y[i] = b->y[i] - etay * b->dp[i];
        xp[i] = b->px[i] - etaxp * b->dp[i];
        yp[i] = b->py[i] - etayp * b->dp[i];
    }
    const double ax = accumulate_int(x, n), ay = accumulate_int(y, n);
    const double axp = accumulate_int(xp, n), ayp = accumulate_int(yp, n);
    double s[10] = {0};
    for (int64_t i = 0; i < n; i++) {
        x[i] -= ax; y[i] -= ay; xp[i] -= axp; yp[i] -= ayp;
        s[0] += x[i] * x[i];   s[1] += x[i] * xp[i]; s[2] += x[i] * y[i];
        s[3] += x[i] * yp[i];  s[4] += xp[i] * xp[i]; s[5] += xp[i] * y[i];
        s[6] += xp[i] * yp[i]; s[7] += y[i] * y[i];   s[8] += y[i] * yp[i];
        s[9] += yp[i] * yp[i];
    }
    free(x); free(y); free(xp); free(yp);
    for (int i = 0; i < 10; i++) s[i] /= n;
    double J[16] = {0}, S[16], CJ[16], CJ2[16];
    J[0 * 4 + 1] = 1; J[1 * 4 + 0] = -1; J[2 * 4 + 3] = 1; J[3 * 4 + 2] = -1;
    S[0] = s[0]; S[1] = s[1]; S[2] = s[2]; S[3] = s[3];
    S[4] = s[1]; S[5] = s[4]; S[6] = s[5]; S[7] = s[6];
    S[8] = s[2]; S[9] = s[5]; S[10] = s[7]; S[11] = s[8];
    S[12] = s[3]; S[13] = s[6]; S[14] = s[8]; S[15] = s[9];
    orc_dgemm(4, 4, 4, S, J, CJ);
    orc_dgemm(4, 4, 4, CJ, CJ, CJ2);
    double tr = 0;
    for (int i = 0; i < 4; i++) tr += CJ2[i * 4 + i];
    const double det = orc_det(4, S);
    o->eigen_x = sqrt(-tr + sqrt(tr * tr - 16 * det)) / 2;
    o->eigen_y = sqrt(-tr - sqrt(tr * tr - 16 * det)) / 2;
    o->emit4d = sqrt(sqrt(det));
    const double a = s[0], bb = s[2], c = s[7];
    const double lam = (a + c) / 2 + sqrt((a - c) * (a - c) / 4 + bb * bb);
    if (bb == 0) o->xy_alpha = a >= c ? 0 : PI / 2;
    else o->xy_alpha = atan2((lam - a), bb);
}

/* GetMPBunchRMS, src/MPBunch.cpp:342-498 */
void orc_rms(const orc_bunch *b, double beta_x_k, double beta_y_k,
             double eta_x0, double etap_x0, double eta_y0, double etap_y0,
             orc_moments *o)
{
    double m[6] = {0};
    int nalive = 0;
    for (int64_t i = 0; i < b->n; i++) {
        if (b->lost[i] != 0) continue;
        m[0] += b->x[i]; m[2] += b->y[i]; m[4] += b->z[i];
        m[1] += b->px[i]; m[3] += b->py[i]; m[5] += b->dp[i];
        nalive++;
    }
    for (int j = 0; j < 6; j++) m[j] /= nalive;
    double x2 = 0, y2 = 0, z2 = 0, px2 = 0, py2 = 0, pz2 = 0, xpx = 0, ypy = 0, zpz = 0;
    for (int64_t i = 0; i < b->n; i++) {
        if (b->lost[i] != 0) continue;
        x2 += pow(b->x[i], 2);   y2 += pow(b->y[i], 2);   z2 += pow(b->z[i], 2);
        px2 += pow(b->px[i], 2); py2 += pow(b->py[i], 2); pz2 += pow(b->dp[i], 2);
        xpx += b->x[i] * b->px[i]; ypy += b->y[i] * b->py[i]; zpz += b->z[i] * b->dp[i];
    }
    x2 /= nalive; y2 /= nalive; z2 /= nalive; px2 /= nalive; py2 /= nalive; pz2 /= nalive;
    xpx /= nalive; ypy /= nalive; zpz /= nalive;
    o->emit_eff[0] = sqrt(x2 * px2 - pow(xpx, 2));
    o->emit_eff[1] = sqrt(y2 * py2 - pow(ypy, 2));
    o->emit_eff[2] = sqrt(z2 * pz2 - pow(zpz, 2));
    o->rms_eff_x = sqrt(o->emit_eff[0] * beta_x_k);
    o->rms_eff_y = sqrt(o->emit_eff[1] * beta_y_k);
    x2 = y2 = z2 = px2 = py2 = pz2 = xpx = ypy = zpz = 0;
    for (int64_t i = 0; i < b->n; i++) {
        if (b->lost[i] != 0) continue;
        x2 += pow(b->x[i] - m[0], 2);  y2 += pow(b->y[i] - m[2], 2);  z2 += pow(b->z[i] - m[4], 2);
        px2 += pow(b->px[i] - m[1], 2); py2 += pow(b->py[i] - m[3], 2); pz2 += pow(b->dp[i] - m[5], 2);
        xpx += (b->x[i] - m[0]) * (b->px[i] - m[1]);
        ypy += (b->y[i] - m[2]) * (b->py[i] - m[3]);
        zpz += (b->z[i] - m[4]) * (b->dp[i] - m[5]);
    }
    x2 /= nalive; y2 /= nalive; z2 /= nalive; px2 /= nalive; py2 /= nalive; pz2 /= nalive;
    xpx /= nalive; ypy /= nalive; zpz /= nalive;
    o->c2[0] = x2; o->c2[1] = xpx; o->c2[2] = px2;
    o->c2[3] = y2; o->c2[4] = ypy; o->c2[5] = py2;
    o->c2[6] = z2; o->c2[7] = zpz; o->c2[8] = pz2;
    o->emit[0] = sqrt(x2 * px2 - pow(xpx, 2));
    o->emit[1] = sqrt(y2 * py2 - pow(ypy, 2));
    o->emit[2] = sqrt(z2 * pz2 - pow(zpz, 2));
    o->rms_x = sqrt(o->emit[0] * beta_x_k);
    o->rms_y = sqrt(o->emit[1] * beta_y_k);
    o->sigma_z = sqrt(z2);
    o->sigma_dp = sqrt(pz2);
    for (int j = 0; j < 6; j++) o->mean[j] = m[j];
    o->n_alive = nalive;
    orc_zminmax(b, &o->zmin, &o->zmax);
    eigen_emit(b, eta_x0, etap_x0, eta_y0, etap_y0, o);
}

/* GetBBRPusdoWakeFun + GetBBRWakeFun1, src/WakeFunction.cpp:558-652 */
void orc_bbr_wake(double tau, int nres, const double *lrs, const double *lq,
                  const double *lomega, const double *txrs, const double *txq,
                  const double *txomega, const double *tyrs, const double *tyq,
                  const double *tyomega, double out[3])
{
    const double sigmat = 1.e-3 / ORC_CLIGHT;
    out[0] = out[1] = out[2] = 0;
    for (int r = 0; r < nres; r++) {
        const double u = tau / sigmat;
        const double sigmaz = sigmat * ORC_CLIGHT;
        const int nz = 1000;
        double w[3] = {0, 0, 0};
        {
            const double Q = lq[r];
            const double vz = lomega[r] * sigmaz / ORC_CLIGHT;
            const double zMax = 3 * vz / (2 * Q);
            const double dz = zMax / nz;
            double sum = 0;
            for (int j = 0; j < nz; j++) {
                const double z = j * dz;
                const double term0 = -1 / 2.0 * pow(u + 2 * Q / vz * z, 2) - z;
                const double term1 = cos(z * sqrt(4 * Q * Q - 1));
                const double term2 = sin(z * sqrt(4 * Q * Q - 1)) / sqrt(4 * Q * Q - 1);
                sum += exp(term0) * (term1 - term2) * dz;
            }
            const double coeffL = sqrt(2. / PI) * lrs[r] * ORC_CLIGHT / sigmaz;
            w[2] = coeffL * sum;
        }
        if (txrs != NULL) {
            const double *Rs[2] = {txrs, tyrs}, *Qs[2] = {txq, tyq}, *Om[2] = {txomega, tyomega};
            for (int t = 0; t < 2; t++) {
                const double Q = Qs[t][r];
                const double vz = Om[t][r] * sigmaz / ORC_CLIGHT;
                const double zMax = 3 * vz / (2 * Q);
                const double dz = zMax / nz;
                double sum = 0;
                for (int j = 0; j < nz; j++) {
                    const double z = j * dz;
                    const double term0 = -1 / 2.0 * pow(u + 2 * Q / vz * z, 2) - z;
                    const double term2 = sin(z * sqrt(4 * Q * Q - 1)) / sqrt(4 * Q * Q - 1);
                    sum += 2 * Q / vz * exp(term0) * term2 * dz;
                }
                w[t] = sum * (sqrt(2. / PI) * Rs[t][r] * ORC_CLIGHT);
            }
        }
        out[0] += w[0];
        out[1] += w[1];
        out[2] += w[2];
    }
}

void orc_wake_table(int nbins, double dt, int nres, const double *lrs,
                    const double *lq, const double *lomega, double *table)
{
    for (int l = 0; l < 2 * nbins - 1; l++) {
        double w[3];
        orc_bbr_wake((l - (nbins - 1)) * dt, nres, lrs, lq, lomega, NULL, NULL, NULL,
                     NULL, NULL, NULL, w);
        table[l] = w[2];
    }
}

void orc_wake_pairs(int nbins, double dt, int nres, const double *lrs, const double *lq,
                    const double *lomega, const int32_t *counts, double *pz)
{
    for (int i = 0; i < nbins; i++) {
        for (int j = 0; j < nbins; j++) {
            double w[3];
            orc_bbr_wake((i - j) * dt, nres, lrs, lq, lomega, NULL, NULL, NULL, NULL, NULL, NULL, w);
            pz[i] += w[2] * counts[j];
        }
    }
}

/* BunchTransferDueToSRWake, src/MPBunch.cpp:1887-2026, longitudinal part
 * (the transverse potentials are output only and not restated). */
int64_t orc_sr_wake(orc_bunch *b, int nbins, int mode, double sigma_z_stale,
                    double grid_lo, double grid_hi, double macro_charge,
                    double energy, int nres, const double *lrs, const double *lq,
                    const double *lomega, double *poten, int32_t *counts,
                    double *edges, const double *wake_table)
{
    double poszMin, poszMax;
    if (mode == 0) {
        double zMin, zMax;
        orc_zminmax(b, &zMin, &zMax);
        poszMin = zMin - 2 * sigma_z_stale;
        poszMax = zMax + 2 * sigma_z_stale;
    } else {
        poszMin = grid_lo;
        poszMax = grid_hi;
    }
    const double dzBin = (poszMax - poszMin) / nbins;
    const double dtBin = dzBin / ORC_CLIGHT;
    edges[0] = poszMin;
    edges[1] = dzBin;
    int32_t *bin = malloc(sizeof(int32_t) * (b->n > 0 ? b->n : 1));
    int64_t skipped = 0;
    for (int i = 0; i < nbins; i++) counts[i] = 0;
    for (int64_t i = 0; i < b->n; i++) {
        /* int() truncates toward zero, so q in (-1, 0) lands in bin 0 as in
         * the reference; indices outside [0, nbins) are skipped. */
        const double q = (b->z[i] - poszMin) / dzBin;
        int idx = (q > -1.0 && q < (double)nbins) ? (int)q : -1;
        bin[i] = idx;
        if (idx < 0) { skipped++; continue; }
        counts[idx]++;
    }
    double *table = NULL;
    if (wake_table == NULL) {
        table = malloc(sizeof(double) * (2 * nbins - 1));
        orc_wake_table(nbins, dtBin, nres, lrs, lq, lomega, table);
        wake_table = table;
    }
    double *pz = poten + 2 * nbins;
    for (int i = 0; i < nbins; i++) {
        for (int j = 0; j < nbins; j++) {
            const int partNumInBin = counts[j];
            pz[i] += wake_table[(i - j) + (nbins - 1)] * partNumInBin;
        }
        pz[i] *= (-1) * ORC_ECHARGE * macro_charge / energy;
    }
    for (int64_t i = 0; i < b->n; i++) {
        if (bin[i] < 0 || b->lost[i] != 0) continue;
        b->dp[i] += pz[bin[i]];
    }
    free(bin);
    free(table);
    return skipped;
}
