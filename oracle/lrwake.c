/* lrwake.c -- CPU oracle (test infrastructure only) of the long-range
 * (coupled-bunch) wake: WakeFunction::InitialLRWake's table
 * (src/WakeFunction.cpp:49-205), the resistive-wall and broadband-resonator
 * wake functions it tabulates (GetRWLRWakeFun, src/WakeFunction.cpp:499-555;
 * GetBBRWakeFun, src/WakeFunction.cpp:655-720), the Yokoya factors of
 * elliptic pipe sectors (RWWakeParaReadIn, src/WakeFunction.cpp:292-350;
 * GetyokoyaFactor and its L functions, src/WakeFunction.cpp:723-880), the
 * linear interpolation GetTotLRWakeLinearFit (src/WakeFunction.cpp:204-229)
 * and the per-turn sum of MPBeam::LRWakeBeamIntaction
 * (src/MPBeam.cpp:3232-3351).
 *
 * The Yokoya series uses GSL's gamma, factorial, 2F1 and complete elliptic
 * K in the reference; here Gamma(m + 1/2)/Gamma(1/2) is the product
 * (1/2)(3/2)...(m - 1/2), 2F1(1/2, b; c; x) (0 <= x < 1) its power series
 * summed to convergence, and K(k) the arithmetic-geometric mean.  Pinned
 * by the round-pipe value (all factors 1, the reference skips the series)
 * and the flat-chamber limits 1, pi^2/24, pi^2/12, -pi^2/24, pi^2/24
 * (tests/test_lrwake.py).  Parity of the rest: closed forms of the wake
 * functions (tau^{-1/2} and tau^{-3/2} laws, resonator at tau = 0).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "cts_oracle.h"

#define PI 3.14159265358979323846

/* include/Global.h:37-39 */
static const double kZ0 = 120 * PI;
static const double kFactorGaussSI = 120 * PI * ORC_CLIGHT / (4 * PI);

/* Gamma(m + 1/2) / Gamma(1/2) for integer m >= 0 */
static double half_gamma_ratio(int m)
{
    double r = 1.0;
    for (int k = 0; k < m; k++) r *= 0.5 + k;
    return r;
}

static double factorial(int m)
{
    double r = 1.0;
    for (int k = 2; k <= m; k++) r *= k;
    return r;
}

/* 2F1(1/2, b; c; x) for 0 <= x < 1 by its power series */
static double hyp2f1_half(double b, double c, double x)
{
    double term = 1.0, sum = 1.0;
    for (int k = 0; k < 2000000; k++) {
        term *= (0.5 + k) * (b + k) / ((c + k) * (k + 1.0)) * x;
        sum += term;
        if (fabs(term) < 1e-17 * fabs(sum)) break;
    }
    return sum;
}

/* complete elliptic integral of the first kind, modulus k */
static double ellint_k(double k)
{
    double a = 1.0, g = sqrt(1.0 - k * k);
    for (int i = 0; i < 64 && fabs(a - g) > 1e-16 * a; i++) {
        const double an = 0.5 * (a + g);
        g = sqrt(a * g);
        a = an;
    }
    return PI / (2.0 * a);
}

/* Lfunction1/2 and Ldfunction1/2 (src/WakeFunction.cpp:760-880) */
static double lfun1(int r, int t, double mu)
{
    const int d = abs(r - t);
    if (r != t)
        return sqrt(2.) * PI * exp(-(2 * d + 1) * mu) * half_gamma_ratio(d) / factorial(d) *
               hyp2f1_half(d + 0.5, d + 1, exp(-4 * mu));
    return 2 * sqrt(2) * exp(-mu) * ellint_k(sqrt(exp(-4 * mu)));
}
static double lfun2(int r, int t, double mu)
{
    const int s = r + t;
    return sqrt(2.) * PI * exp(-(2 * s + 1) * mu) * half_gamma_ratio(s) / factorial(s) *
           hyp2f1_half(s + 0.5, s + 1, exp(-4 * mu));
}
static double ldfun1(int r, int t, double mu)
{
    const int d = abs(r - t);
    return sqrt(2.) * PI * exp(-(2 * d + 1) * mu) * half_gamma_ratio(d) / factorial(d) *
           hyp2f1_half(d + 0.5, d + 1, exp(-4 * mu));
}
static double ldfun2(int r, int t, double mu)
{
    const int s = r + t;
    return sqrt(2.) * PI * exp(-(2 * s + 3) * mu) * half_gamma_ratio(s + 1) / factorial(s + 1) *
           hyp2f1_half(s + 1.5, s + 2, exp(-4 * mu));
}

/* GetyokoyaFactor(a, b) for a > b (src/WakeFunction.cpp:723-757) */
static void yokoya_ab(double a, double b, double out[5])
{
    const double f = sqrt(fabs(a * a - b * b));
    const double mu = acosh(a / f);
    double coef[5];
    coef[0] = 2 * sqrt(2) / PI * b / f;
    coef[1] = sqrt(2) / PI * pow(b / f, 3);
    coef[2] = coef[1];
    coef[3] = -coef[1];
    coef[4] = coef[1];
    double t[5] = {0, 0, 0, 0, 0};
    for (int L = 0; L < 50; L++) {
        const int dl = L == 0 ? 2 : 1;
        for (int P = 0; P < 50; P++) {
            const int dp = P == 0 ? 2 : 1;
            const double l0 = lfun1(P, L, mu) + lfun2(P, L, mu);
            const double l1 = ldfun1(P, L, mu) + ldfun2(P, L, mu);
            const double l2 = ldfun1(P, L, mu) - ldfun2(P, L, mu);
            const double sg = ((P + L) % 2) ? -1.0 : 1.0;
            t[0] += l0 * sg * 1. / cosh(2 * P * mu) / cosh(2 * L * mu) / dp / dl;
            t[1] += l1 * sg * 1. / cosh((2 * P + 1) * mu) / cosh((2 * L + 1) * mu) * (2 * P + 1) * (2 * L + 1);
            t[2] += l2 * sg * 1. / sinh((2 * P + 1) * mu) / sinh((2 * L + 1) * mu) * (2 * P + 1) * (2 * L + 1);
            t[3] += l0 * sg * 1. / cosh(2 * P * mu) / cosh(2 * L * mu) * pow(2 * P, 2) / dl;
            t[4] += l0 * sg * 1. / cosh(2 * P * mu) / cosh(2 * L * mu) * pow(2 * P, 2) / dl;
        }
    }
    for (int i = 0; i < 5; i++) out[i] = coef[i] * t[i];
}

/* the sector's factors as RWWakeParaReadIn assigns them
 * (src/WakeFunction.cpp:321-345) */
void orc_lrw_yokoya(double rx, double ry, double out[5])
{
    for (int i = 0; i < 5; i++) out[i] = 1.0;
    if (rx > ry) {
        yokoya_ab(rx, ry, out);
    } else if (rx < ry) {
        yokoya_ab(ry, rx, out);
        double t = out[1]; out[1] = out[2]; out[2] = t;
        t = out[3]; out[3] = out[4]; out[4] = t;
    }
}

/* GetRWLRWakeFun (src/WakeFunction.cpp:499-555); sectors[7] per sector */
static void rw_wake(const double *sec, const double *yok, int nsec, double bx, double by, double tau, double w[3])
{
    w[0] = w[1] = w[2] = 0.0;
    if (tau == 0) return;
    for (int i = 0; i < nsec; i++) {
        const double *s = sec + 7 * i;
        const double radius = s[0] >= s[1] ? s[1] : s[0];
        const double sig = s[6] * kFactorGaussSI;
        w[0] += 1. / pow(radius, 3) * sqrt(1.E0 / sig) * s[2] * s[5] * yok[5 * i + 1] * s[3];
        w[1] += 1. / pow(radius, 3) * sqrt(1.E0 / sig) * s[2] * s[5] * yok[5 * i + 2] * s[4];
        w[2] += 1. / radius * sqrt(1.E0 / sig) * s[2] * s[5] * yok[5 * i + 0];
    }
    w[0] = w[0] * (-2.) / PI / pow(fabs(tau), 0.5) * kFactorGaussSI / bx;
    w[1] = w[1] * (-2.) / PI / pow(fabs(tau), 0.5) * kFactorGaussSI / by;
    w[2] = w[2] / (2.) / PI / ORC_CLIGHT / pow(fabs(tau), 1.5) * kFactorGaussSI;
    (void)kZ0;
}

/* GetBBRWakeFun (src/WakeFunction.cpp:655-720); res[9] per resonator,
 * frequencies in Hz */
static void bbr_wake(const double *res, int nres, double bx, double by, double tau, double w[3])
{
    w[0] = w[1] = w[2] = 0.0;
    for (int i = 0; i < nres; i++) {
        const double *r = res + 9 * i;
        const double txo = r[5] * 2 * PI, tyo = r[8] * 2 * PI;
        double alpha = txo / 2.0 / r[4];
        double omegab = sqrt(pow(txo, 2) - pow(alpha, 2));
        double coeff = ORC_CLIGHT * r[3] * txo / r[4] / omegab * exp(alpha * tau) * sin(omegab * tau);
        w[0] += coeff / bx;
        alpha = tyo / 2.0 / r[7];
        omegab = sqrt(pow(tyo, 2) - pow(alpha, 2));
        coeff = ORC_CLIGHT * r[6] * tyo / r[7] / omegab * exp(alpha * tau) * sin(omegab * tau);
        w[1] += coeff / by;
    }
    for (int i = 0; i < nres; i++) {
        const double *r = res + 9 * i;
        const double lo = r[2] * 2 * PI;
        const double alpha = lo / 2.0 / r[1];
        const double omegab = sqrt(pow(lo, 2) - pow(alpha, 2));
        const double coeff = 2 * alpha * r[0] * exp(alpha * tau);
        const double c1 = cos(omegab * tau);
        const double c2 = sin(omegab * tau) * alpha / omegab;
        if (tau == 0) w[2] += coeff / 2;
        else w[2] += coeff * (c1 + c2);
    }
}

/* InitialLRWake's table (src/WakeFunction.cpp:110-190): np points at
 * time[i] = i dt, dt = T_rf / 2; w, rw, bbr are [3][np].  Returns np, or
 * -1 when it exceeds cap. */
int orc_lrw_table(int nturns, int harmonic, double t0, int nsec, const double *sectors, int nres,
                  const double *res, double beta_x, double beta_y, int cap, double *time, double *w,
                  double *rw, double *bbr)
{
    const double trf = t0 / harmonic;
    const double dt = trf / 2;
    const int np = (int)(nturns * harmonic * trf / dt) + 2;
    if (np > cap) return -1;
    double *yok = (double *)malloc(sizeof(double) * 5 * (nsec > 0 ? nsec : 1));
    for (int i = 0; i < nsec; i++) orc_lrw_yokoya(sectors[7 * i], sectors[7 * i + 1], yok + 5 * i);
    for (int i = 0; i < np; i++) {
        const double tau = -i * dt;
        double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
        if (nsec > 0) rw_wake(sectors, yok, nsec, beta_x, beta_y, tau, a);
        if (nres > 0) bbr_wake(res, nres, beta_x, beta_y, tau, b);
        time[i] = -tau;
        for (int j = 0; j < 3; j++) {
            w[j * np + i] = b[j] + a[j];
            if (rw) rw[j * np + i] = a[j];
            if (bbr) bbr[j * np + i] = b[j];
        }
    }
    free(yok);
    return np;
}

/* One turn of MPBeam::LRWakeBeamIntaction's sum (src/MPBeam.cpp:3270-3346)
 * after the history shift: pos[(nturns - 1 - n) * nb * 3 + 3 i + {0,1,2}] =
 * bunch i's centroid x, y, z n turns ago (pos[nturns - 1] the present one);
 * harm[i] = bunchHarmNum, ne[i] = electronNumPerBunch.  kick[3 j + k] =
 * lRWakeForceAver[k] of bunch j.  Returns -1 if some tau > 0 (the
 * reference exits there), else 0. */
int orc_lrw_kick(int nturns, int nb, const int *harm, int harmonic, double t0, double rbeta, double energy,
                 const double *ne, const double *pos, const double *w, int np, double *kick)
{
    const double trf = t0 / (double)harmonic;
    const double dt = trf / 2;  /* lwaketime[1] - lwaketime[0] */
    for (int j = 0; j < nb; j++) {
        double f[3] = {0, 0, 0};
        const double zj = pos[(size_t)(nturns - 1) * nb * 3 + 3 * j + 2];
        for (int n = 0; n < nturns; n++) {
            int i0, i1;
            if (n == 0) { i0 = 0; i1 = j - 1; }
            else if (n == nturns - 1) { i0 = j; i1 = nb - 1; }
            else { i0 = 0; i1 = nb - 1; }
            const double *pn = pos + (size_t)(nturns - 1 - n) * nb * 3;
            for (int i = i0; i <= i1; i++) {
                const int ntau = harm[i] - harm[j] - n * harmonic;
                const double tau = ntau * trf + (zj - pn[3 * i + 2]) / ORC_CLIGHT / rbeta;
                if (tau > 0) return -1;
                const int index = (int)floor(-tau / dt);
                if (index + 1 >= np) return -2;
                const double v1 = -tau / dt - index, v0 = 1 - v1;
                double wk[3];
                for (int k = 0; k < 3; k++) wk[k] = w[k * np + index] * v0 + w[k * np + index + 1] * v1;
                f[0] -= wk[0] * ne[i] * pn[3 * i];
                f[1] -= wk[1] * ne[i] * pn[3 * i + 1];
                f[2] -= wk[2] * ne[i];
            }
        }
        for (int k = 0; k < 3; k++) kick[3 * j + k] = f[k] * ORC_ECHARGE / energy / pow(rbeta, 2);
    }
    return 0;
}
