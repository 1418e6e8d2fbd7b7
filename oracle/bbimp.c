/* bbimp.c -- CPU oracle (test infrastructure only) of the broadband
 * impedance in the frequency domain, longitudinal, dipolar and quadrupolar
 * planes: MPBunch::BBImpBunchInteraction (src/MPBunch.cpp:2143-2312) with
 * the grid of BoardBandImp::ReadInImp (src/BoardBandImp.cpp:71-90) and the
 * smoothing of GetSmoothedBunchProfileGassionFilter
 * (src/MPBunch.cpp:2348-2410).
 * FFTW's r2c and c2r transforms are written as direct sums (c2r of odd
 * length n: y_j = Re X_0 + 2 sum_{k=1}^{(n-1)/2} Re(X_k e^{2 pi i j k/n})),
 * and the GSL 2.8 Gaussian filter (gsl_filter_gaussian, the dependency the
 * reference links, not part of it) from its published definition: window
 * K = 51, kernel exp(-alpha^2 x^2/2) on x = (i - 25)/25 scaled to unit sum,
 * the input padded with its end values.  Pinned in tests/test_bbimp.py by a
 * constant resistive impedance (V = -R I), a linear reactive one (the
 * spectral derivative of a cosine), a constant profile (smoothing changes
 * nothing), a single spike (it returns the normalised kernel) and a
 * quadrupolar Z_Q = -i R (V_Q = +R I, and the px kick), a dipolar
 * Z_D = -i R (V_D = +R I_x from the smoothed x-weighted profile, and the
 * px kick), and a bunch at one common offset x0, whose dipolar kick equals
 * its quadrupolar kick for the same impedance table.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "cts_oracle.h"

#define PI 3.14159265358979323846
#define TAPS 51

void orc_bbi_smooth(double *p, int n)
{
    const int half = (TAPS - 1) / 2;
    const double alpha = 10.0;
    double k[TAPS], sum = 0;
    for (int i = 0; i < TAPS; i++) {
        const double xi = ((double)i - 0.5 * (TAPS - 1.0)) / (0.5 * (TAPS - 1.0));
        k[i] = exp(-0.5 * alpha * alpha * xi * xi);
        sum += k[i];
    }
    for (int i = 0; i < TAPS; i++) k[i] *= 1.0 / sum;
    double *y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int j = 0; j < TAPS; j++) {
            int q = i + j - half;
            q = q < 0 ? 0 : (q >= n ? n - 1 : q);
            s += p[q] * k[j];
        }
        y[i] = s;
    }
    double s0 = 0, s1 = 0;
    for (int i = 0; i < n; i++) {
        s0 += p[i];
        s1 += y[i];
        p[i] = y[i];
    }
    s0 = s0 / s1;
    for (int i = 0; i < n; i++) p[i] *= s0;
    free(y);
}

/* X_k = sum_j p_j e^{-2 pi i j k / n}, k < nimp (FFTW r2c) */
static void spectrum(const double *p, int nimp, double *xr, double *xi)
{
    const int n = 2 * nimp - 1;
    for (int k = 0; k < nimp; k++) {
        double sr = 0, si = 0;
        for (int j = 0; j < n; j++) {
            const double ph = -2 * PI * (double)(((int64_t)j * k) % n) / n;
            sr += p[j] * cos(ph);
            si += p[j] * sin(ph);
        }
        xr[k] = sr;
        xi[k] = si;
    }
}

/* out_j = s c2r(X Z)_j / n, with X Z times i when rot (the quadrupolar planes) */
static void product_c2r(const double *xr, const double *xi, int nimp, const double *zre, const double *zim, int rot,
                        double s, double *out)
{
    const int n = 2 * nimp - 1;
    double *yr = (double *)malloc(sizeof(double) * nimp), *yi = (double *)malloc(sizeof(double) * nimp);
    for (int k = 0; k < nimp; k++) {
        const double ar = xr[k] * zre[k] - xi[k] * zim[k], ai = xr[k] * zim[k] + xi[k] * zre[k];
        yr[k] = rot ? -ai : ar;
        yi[k] = rot ? ar : ai;
    }
    for (int j = 0; j < n; j++) {
        double t = 0;
        for (int k = 1; k < nimp; k++) {
            const double ph = 2 * PI * (double)(((int64_t)j * k) % n) / n;
            t += yr[k] * cos(ph) - yi[k] * sin(ph);
        }
        out[j] = s * (yr[0] + 2 * t) / n;
    }
    free(yr);
    free(yi);
}

/* wake[j] = -c2r(r2c(p) Z)[j] / n, n = 2 nimp - 1 */
void orc_bbi_potential(const double *p, int nimp, const double *zre, const double *zim, double *wake)
{
    double *xr = (double *)malloc(sizeof(double) * nimp), *xi = (double *)malloc(sizeof(double) * nimp);
    spectrum(p, nimp, xr, xi);
    product_c2r(xr, xi, nimp, zre, zim, 0, -1.0, wake);
    free(xr);
    free(xi);
}

/* One bunch's interaction: survivors binned on the grid
 * (int)((z + z_max + dz/2)/dz) of n = 2 nimp - 1 bins; three profiles, each
 * a running sum per particle as the reference forms them: the charge
 * profile (src/MPBunch.cpp:2166-2171) and, for the dipolar planes, the
 * x- and y-weighted ones (2256-2261, 2290-2295); each smoothed over the
 * bucket on its own.  The induced voltages, in the reference's flag order:
 * longitudinal V = -c2r(X Z)/n (planes[0]); dipolar V_Dx = c2r(X_x Z_Dx i)/n
 * (planes[1]), V_Dy likewise (planes[2]); quadrupolar V_Q = c2r(X Z_Q i)/n
 * (planes[3], planes[4]).  Kicks in the reference's order: delta +=
 * V/(E beta^2); px += V_Qx/(E beta^2) x/beta_x, py likewise; px +=
 * V_Dx/(E beta^2)/beta_x, py likewise.  zre, zim [5][nimp] and wake [5][n]
 * in flag order; prof [3][n] (charge, x-weighted, y-weighted). */
void orc_bbi_kick(orc_bunch *b, int nimp, double f_max, const int *planes, const double *zre, const double *zim,
                  double bx, double by, double macro_charge, double rbeta, double t0, int harmonic, double energy,
                  double *wake, double *prof)
{
    const int n = 2 * nimp - 1;
    const double dt = 1.0 / (2 * f_max), dz = dt * ORC_CLIGHT * rbeta;
    const double zmin = -dz * (nimp - 1);
    const double qc = macro_charge * ORC_ECHARGE / dz * ORC_CLIGHT;
    int *idx = (int *)malloc(sizeof(int) * (size_t)(b->n > 0 ? b->n : 1));
    for (int i = 0; i < 3 * n; i++) prof[i] = 0;
    for (int64_t i = 0; i < b->n; i++) {
        idx[i] = -1;
        if (b->lost[i] != 0) continue;
        const double q = (b->z[i] - zmin + dz / 2) / dz;
        if (!(q > -1.0 && q < (double)n)) continue;
        idx[i] = (int)q;
        prof[idx[i]] += qc;
        prof[n + idx[i]] += qc * b->x[i];
        prof[2 * n + idx[i]] += qc * b->y[i];
    }
    const double rf_len = t0 / harmonic * ORC_CLIGHT;
    const int is = (int)((-rf_len / 2. - zmin + dz / 2) / dz);
    const int ie = (int)((rf_len / 2. - zmin + dz / 2) / dz);
    for (int r = 0; r < 3; r++) orc_bbi_smooth(prof + (size_t)r * n + is, ie - is);
    double *xr = (double *)malloc(sizeof(double) * nimp), *xi = (double *)malloc(sizeof(double) * nimp);
    for (int q = 0; q < 5; q++)
        for (int j = 0; j < n; j++) wake[(size_t)q * n + j] = 0;
    /* profile of each plane: charge for Z and the quadrupolar planes */
    const int src[5] = {0, 1, 2, 0, 0};
    for (int r = 0; r < 3; r++) {
        int used = 0;
        for (int q = 0; q < 5; q++) used |= planes[q] && src[q] == r;
        if (!used) continue;
        spectrum(prof + (size_t)r * n, nimp, xr, xi);
        for (int q = 0; q < 5; q++)
            if (planes[q] && src[q] == r)
                product_c2r(xr, xi, nimp, zre + (size_t)q * nimp, zim + (size_t)q * nimp, q > 0, q == 0 ? -1.0 : 1.0,
                            wake + (size_t)q * n);
    }
    const double eb = energy * pow(rbeta, 2);
    for (int64_t i = 0; i < b->n; i++) {
        const int k = idx[i];
        if (k < 0) continue;
        if (planes[0]) b->dp[i] += wake[k] / eb;
        if (planes[3]) b->px[i] += wake[3 * n + k] / eb * b->x[i] / bx;
        if (planes[4]) b->py[i] += wake[4 * n + k] / eb * b->y[i] / by;
        if (planes[1]) b->px[i] += wake[n + k] / eb / bx;
        if (planes[2]) b->py[i] += wake[2 * n + k] / eb / by;
    }
    free(xr);
    free(xi);
    free(idx);
}
