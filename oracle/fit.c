/*
 * fit.c -- weighted least-squares line, as FittingGSL::FitALinear calls it
 * (src/FittingGSL.cpp:231-243 -> gsl_fit_wlinear).  Test infrastructure
 * only (see cts_oracle.h).
 *
 * gsl_fit_wlinear (GSL 2.8 fit/linear.c) forms running weighted means of
 * x and y, then running weighted means of dx^2 and dx*dy about them;
 * slope = <dx dy>/<dx^2>, intercept = <y> - <x>*slope.  Points with
 * non-positive weight are skipped.
 */
#include "cts_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>

void orc_fit_wlinear(const double *x, const double *w, const double *y, int n,
                     double *c0, double *c1)
{
    double W = 0, wm_x = 0, wm_y = 0, wm_dx2 = 0, wm_dxdy = 0;
    for (int i = 0; i < n; i++) {
        const double wi = w[i];
        if (wi > 0) {
            W += wi;
            wm_x += (x[i] - wm_x) * (wi / W);
            wm_y += (y[i] - wm_y) * (wi / W);
        }
    }
    W = 0;
    for (int i = 0; i < n; i++) {
        const double wi = w[i];
        if (wi > 0) {
            const double dx = x[i] - wm_x;
            const double dy = y[i] - wm_y;
            W += wi;
            wm_dx2 += (dx * dx - wm_dx2) * (wi / W);
            wm_dxdy += (dx * dy - wm_dxdy) * (wi / W);
        }
    }
    const double b = wm_dxdy / wm_dx2;
    *c1 = b;
    *c0 = wm_y - wm_x * b;
}

/* ------------------------------------------------------------------------
 * GetCBMGR's analysis (src/MPBeam.cpp:1496-1891) with direct DFTs.
 * mean[n][b][6] are bunch centroids at sample n; outputs [n][plane][k] as
 * in cts_cbm_analysis; any output may be NULL. */
static const double ORC_PI = 3.141592653589793;

static double complex orc_phasor(int pl, double q, double pq, const double *tunes, const double *twiss, int hb,
                                 int harmonics)
{
    const double ph = -2.0 * ORC_PI * tunes[pl] * hb / harmonics;   /* 1532-1534 */
    const double complex rot = cexp(I * ph);
    if (pl < 2) {
        const double al = twiss[2 * pl], be = twiss[2 * pl + 1];     /* 1536-1537 */
        return (q / sqrt(be) - I * (sqrt(be) * pq + al / sqrt(be) * q)) * rot;
    }
    const double bz = twiss[4];                                      /* 1538, alpha_z = 0 */
    return (q / sqrt(bz) + I * (sqrt(bz) * pq)) * rot;
}

/* forward DFT over the bunches (gsl_fft_complex_forward, 1548-1550) */
static void orc_dft(const double complex *in, double complex *out, int n, int sign)
{
    for (int k = 0; k < n; k++) {
        double complex acc = 0;
        for (int j = 0; j < n; j++) {
            const double ang = sign * 2.0 * ORC_PI * (double)((long long)j * k % n) / n;
            acc += in[j] * (cos(ang) + I * sin(ang));
        }
        out[k] = acc;
    }
}

static double orc_rate(const double *v, int i0, int dim, int interval, double t0)
{
    double *fx = malloc(sizeof(double) * dim), *fw = malloc(sizeof(double) * dim), *fy = malloc(sizeof(double) * dim);
    for (int n = 0; n < dim; n++) {                                  /* 1801-1816 */
        fw[n] = 1.0;
        fx[n] = (double)n * interval;
        fy[n] = log(v[n + i0]);
    }
    double c0, c1;
    orc_fit_wlinear(fx, fw, fy, dim, &c0, &c1);
    free(fx); free(fw); free(fy);
    return c1 / t0;
}

int orc_cbm_analysis(const double *mean, int ns, int nb, const int *harm, int harmonics, const double *tunes,
                     const double *twiss, int interval, int fit_start, int fit_end, double t0,
                     double *ideal_amp, double *ideal_arg, double *iq_amp, double *iq_arg, double *ana_re,
                     double *ana_im, double *hil_amp, double *hil_arg, double *rate_ideal, double *rate_iq,
                     double *rate_hil, double *rate_bunch)
{
    const size_t NS = (size_t)ns, NB = (size_t)nb;
#define AT(n, pl, b) ((((size_t)(n)) * 3 + (size_t)(pl)) * NB + (size_t)(b))
    double *a_id = calloc(NS * 3 * NB, sizeof(double)), *a_iq = calloc(NS * 3 * NB, sizeof(double));
    double *a_hi = calloc(NS * 3 * NB, sizeof(double));
    double complex *ana = calloc(NS * 3 * NB, sizeof(double complex));
    const size_t mx = NS > NB ? NS : NB;
    double complex *z = calloc(mx, sizeof(double complex)), *zf = calloc(mx, sizeof(double complex));
    for (size_t n = 0; n < NS; n++)
        for (int pl = 0; pl < 3; pl++) {
            for (size_t b = 0; b < NB; b++) {
                const double *q = mean + (n * NB + b) * 6;
                z[b] = orc_phasor(pl, q[2 * pl], q[2 * pl + 1], tunes, twiss, harm[b], harmonics);
            }
            orc_dft(z, zf, nb, -1);
            for (size_t k = 0; k < NB; k++) {                        /* 1561-1568 */
                const double re = creal(zf[k]), im = cimag(zf[k]);
                a_id[AT(n, pl, k)] = sqrt(pow(re, 2) + pow(im, 2));
                if (ideal_arg) ideal_arg[AT(n, pl, k)] = atan2(re, im);
            }
        }
    /* IQ decomposition (1578-1620): f_j = (j + nu) f0, t_i = h_i T0 / h */
    const double f0 = 1 / t0, trf = t0 / harmonics;
    for (size_t n = 0; n < NS; n++)
        for (int pl = 0; pl < 3; pl++)
            for (size_t j = 0; j < NB; j++) {
                const double fj = ((double)j + tunes[pl]) * f0;
                double complex acc = 0;
                for (size_t i = 0; i < NB; i++)
                    acc += mean[(n * NB + i) * 6 + 2 * pl] * cexp(-I * 2.0 * ORC_PI * fj * (harm[i] * trf));
                acc /= (double)nb;
                a_iq[AT(n, pl, j)] = cabs(acc);
                if (iq_arg) iq_arg[AT(n, pl, j)] = carg(acc);
            }
    /* analytic signals (1936-2011): one-sided spectrum, DC kept */
    for (int pl = 0; pl < 3; pl++)
        for (size_t b = 0; b < NB; b++) {
            for (size_t n = 0; n < NS; n++) z[n] = mean[(n * NB + b) * 6 + 2 * pl];
            orc_dft(z, zf, ns, -1);
            for (size_t i = 1; i <= NS / 2; i++) zf[i] *= 2.0;
            if (NS % 2 == 0) zf[NS / 2] /= 2.0;
            for (size_t i = NS / 2 + 1; i < NS; i++) zf[i] = 0;
            orc_dft(zf, z, ns, +1);
            for (size_t n = 0; n < NS; n++) ana[AT(n, pl, b)] = z[n] / (double)ns;
        }
    /* modes of (Re, -Im) of the analytic signals (1671-1731) */
    for (size_t n = 0; n < NS; n++)
        for (int pl = 0; pl < 3; pl++) {
            for (size_t b = 0; b < NB; b++) {
                const double complex a = ana[AT(n, pl, b)];
                z[b] = orc_phasor(pl, creal(a), -cimag(a), tunes, twiss, harm[b], harmonics);
            }
            orc_dft(z, zf, nb, -1);
            for (size_t k = 0; k < NB; k++) {
                const double re = creal(zf[k]), im = cimag(zf[k]);
                a_hi[AT(n, pl, k)] = sqrt(pow(re, 2) + pow(im, 2));
                if (hil_arg) hil_arg[AT(n, pl, k)] = atan2(re, im);
            }
        }
    for (size_t i = 0; i < NS * 3 * NB; i++) {
        if (ideal_amp) ideal_amp[i] = a_id[i];
        if (iq_amp) iq_amp[i] = a_iq[i];
        if (hil_amp) hil_amp[i] = a_hi[i];
        if (ana_re) ana_re[i] = creal(ana[i]);
        if (ana_im) ana_im[i] = cimag(ana[i]);
    }
    int rc = 0;
    const int i0 = fit_start / interval, dim = fit_end / interval - i0;
    if (rate_ideal || rate_iq || rate_hil || rate_bunch) {
        if (i0 < 0 || dim < 2 || fit_end / interval > ns) {
            rc = -1;
        } else {
            double *v = malloc(sizeof(double) * NS);
            for (int pl = 0; pl < 3; pl++)
                for (size_t k = 0; k < NB; k++) {
                    const size_t r = (size_t)pl * NB + k;
                    for (size_t n = 0; n < NS; n++) v[n] = a_id[AT(n, pl, k)];
                    if (rate_ideal) rate_ideal[r] = orc_rate(v, i0, dim, interval, t0);
                    for (size_t n = 0; n < NS; n++) v[n] = a_iq[AT(n, pl, k)];
                    if (rate_iq) rate_iq[r] = orc_rate(v, i0, dim, interval, t0);
                    for (size_t n = 0; n < NS; n++) v[n] = a_hi[AT(n, pl, k)];
                    if (rate_hil) rate_hil[r] = orc_rate(v, i0, dim, interval, t0);
                    for (size_t n = 0; n < NS; n++) v[n] = cabs(ana[AT(n, pl, k)]);
                    if (rate_bunch) rate_bunch[r] = orc_rate(v, i0, dim, interval, t0);
                }
            free(v);
        }
    }
#undef AT
    free(a_id); free(a_iq); free(a_hi); free(ana); free(z); free(zf);
    return rc;
}
