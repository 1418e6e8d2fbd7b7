/*
 * faddeeva.c -- oracle restatement of the Faddeeva function w(z) and of the
 * Bassetti-Erskine field used by the beam-ion kicks.  Test infrastructure
 * only (see cts_oracle.h).
 *
 * w(z) follows the region structure of the reference's vendored Faddeeva
 * package (S. G. Johnson, MIT licence; src/Faddeeva.cpp:682-974):
 *   - the real and imaginary axes are special cases (684-689);
 *   - a continued fraction (Gautschi 1970; Poppe & Wijers 1990) for large |z|
 *     (715-780);
 *   - the Zaghloul & Ali (2011, ACM TOMS Algorithm 916) sums for |x| < 10
 *     (819-925), with the table exp(-a^2 n^2) computed here;
 *   - the large-|x| sums otherwise (926-970).
 * The reference's Chebyshev fits for erfcx and Im w(x) on the axes
 * (src/Faddeeva.cpp:976-1890) are replaced by this project's own
 * evaluation: erfcx(x) = exp(x^2) erfc(x) with x^2 split exactly, a
 * continued fraction for x >= 25, Dawson's Taylor series for |Im w(x)|
 * with |x| < 0.5 and the sums above at y = 0 otherwise.  The region
 * thresholds and the continued-fraction depth are the reference's (the
 * 1e-13 known answers near x = 6 are rounding-limited and depend on them).
 */
#include "cts_oracle.h"

#include <float.h>
#include <math.h>

static const double PI = 3.141592653589793;
static const double ISPI = 0.56418958354775628694807945156; /* 1/sqrt(pi) */

/* Algorithm 916 parameters for relerr = DBL_EPSILON:
 * a = pi / sqrt(-log(eps/2)), c = 2a/pi. */
static double g_a, g_c, g_a2, g_ea2n2[64];
static int g_init = 0;

static void fad_init(void)
{
    if (g_init) return;
    g_a = PI / sqrt(-log(DBL_EPSILON * 0.5));
    g_c = (2 / PI) * g_a;
    g_a2 = g_a * g_a;
    for (int n = 1; n <= 63; n++) g_ea2n2[n - 1] = exp(-g_a2 * (double)(n * n));
    g_ea2n2[63] = 0.0;
    g_init = 1;
}

static inline double sinc_(double x, double sinx)
{
    return fabs(x) < 1e-4 ? 1 - (0.1666666666666666666667) * x * x : sinx / x;
}

static inline double sinh_taylor(double x)
{
    return x * (1 + (x * x) * (0.1666666666666666666667 + 0.00833333333333333333333 * (x * x)));
}

/* erfcx(x) = exp(x^2) erfc(x). */
double orc_erfcx(double x)
{
    if (isnan(x)) return x;
    if (x < 0) {
        if (x < -26.7) return HUGE_VAL;
        return 2 * exp(x * x) - orc_erfcx(-x);
    }
    if (x < 25.0) {
        const double h = x * x;
        const double l = fma(x, x, -h); /* x^2 = h + l exactly */
        return exp(h) * (1.0 + l) * erfc(x);
    }
    if (isinf(x)) return 0.0;
    /* Laplace continued fraction erfc(x) = exp(-x^2)/sqrt(pi) *
     * 1/(x + (1/2)/(x + 1/(x + (3/2)/(x + ...)))), evaluated bottom-up. */
    double t = x;
    for (int k = 24; k >= 1; k--) t = x + (0.5 * k) / t;
    return ISPI / t;
}

/* Number of levels of the continued fraction: the depth fit of the
 * reference's package (src/Faddeeva.cpp:758-759).  Near x = 6, |y| = 0.1
 * the result is rounding-limited and only this depth keeps the reference's
 * own 1e-13 known answer at 6.01+0.11i. */
static inline double cf_depth(double x, double ya)
{
    return floor(3.9 + 11.398 / (0.08254 * x + 0.1421 * ya + 0.2023));
}

static void w_general(double zr, double zi, double *wr_out, double *wi_out)
{
    fad_init();
    const double a = g_a, c = g_c, a2 = g_a2;
    const double relerr = DBL_EPSILON;
    const double x = fabs(zr);
    const double y = zi, ya = fabs(y);
    double retr = 0, reti = 0;
    double sum1 = 0, sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0;

    if (ya > 7 || (x > 6 && (ya > 0.1 || (x > 8 && ya > 1e-10) || x > 28))) {
        const double xs = y < 0 ? -zr : zr;
        if (x + ya > 4000) {
            if (x + ya > 1e7) {
                if (x > ya) {
                    const double yax = ya / xs;
                    const double denom = ISPI / (xs + yax * ya);
                    retr = denom * yax; reti = denom;
                } else if (isinf(ya)) {
                    if (isnan(x) || y < 0) { *wr_out = NAN; *wi_out = NAN; }
                    else { *wr_out = 0; *wi_out = 0; }
                    return;
                } else {
                    const double xya = xs / ya;
                    const double denom = ISPI / (xya * xs + ya);
                    retr = denom; reti = denom * xya;
                }
            } else {
                const double dr = xs * xs - ya * ya - 0.5, di = 2 * xs * ya;
                const double denom = ISPI / (dr * dr + di * di);
                retr = denom * (xs * di - ya * dr);
                reti = denom * (xs * dr + ya * di);
            }
        } else {
            double nu = cf_depth(x, ya);
            double wr = xs, wi = ya;
            for (nu = 0.5 * (nu - 1); nu > 0.4; nu -= 0.5) {
                const double denom = nu / (wr * wr + wi * wi);
                wr = xs - wr * denom;
                wi = ya + wi * denom;
            }
            const double denom = ISPI / (wr * wr + wi * wi);
            retr = denom * wi;
            reti = denom * wr;
        }
        if (y < 0) {
            /* w(z) = 2 exp(-z^2) - w(-z); cexp(-inf + i*anything) = 0
             * (C99 Annex G), which an overflowing exponent reaches */
            const double re = (ya - xs) * (xs + ya);
            const double ph = 2 * xs * y;
            double er_c = 0.0, er_s = 0.0;
            if (!(isinf(re) && re < 0)) {
                const double er = exp(re);
                er_c = er * cos(ph);
                er_s = er * sin(ph);
            }
            *wr_out = 2.0 * er_c - retr;
            *wi_out = 2.0 * er_s - reti;
        } else {
            *wr_out = retr; *wi_out = reti;
        }
        return;
    } else if (x < 10) {
        double prod2ax = 1, prodm2ax = 1, expx2;
        if (isnan(y)) { *wr_out = y; *wi_out = y; return; }
        if (x < 5e-4) {
            const double x2 = x * x;
            expx2 = 1 - x2 * (1 - 0.5 * x2);
            const double ax2 = (2 * a) * x;
            const double exp2ax = 1 + ax2 * (1 + ax2 * (0.5 + 0.166666666666666666667 * ax2));
            const double expm2ax = 1 - ax2 * (1 - ax2 * (0.5 - 0.166666666666666666667 * ax2));
            for (int n = 1; 1; ++n) {
                const double coef = g_ea2n2[n - 1] * expx2 / (a2 * (n * n) + y * y);
                prod2ax *= exp2ax;
                prodm2ax *= expm2ax;
                sum1 += coef;
                sum2 += coef * prodm2ax;
                sum3 += coef * prod2ax;
                sum5 += coef * (2 * a) * n * sinh_taylor((2 * a) * n * x);
                if (coef * prod2ax < relerr * sum3) break;
            }
        } else {
            expx2 = exp(-x * x);
            const double exp2ax = exp((2 * a) * x), expm2ax = 1 / exp2ax;
            for (int n = 1; 1; ++n) {
                const double coef = g_ea2n2[n - 1] * expx2 / (a2 * (n * n) + y * y);
                prod2ax *= exp2ax;
                prodm2ax *= expm2ax;
                sum1 += coef;
                sum2 += coef * prodm2ax;
                sum4 += (coef * prodm2ax) * (a * n);
                sum3 += coef * prod2ax;
                sum5 += (coef * prod2ax) * (a * n);
                if ((coef * prod2ax) * (a * n) < relerr * sum5) break;
            }
        }
        const double expx2erfcxy = y > -6 ? expx2 * orc_erfcx(y) : 2 * exp(y * y - x * x);
        if (y > 5) {
            const double sinxy = sin(x * y);
            retr = (expx2erfcxy - c * y * sum1) * cos(2 * x * y) +
                   (c * x * expx2) * sinxy * sinc_(x * y, sinxy);
            reti = 0;
        } else {
            const double xs = zr;
            const double sinxy = sin(xs * y);
            const double sin2xy = sin(2 * xs * y), cos2xy = cos(2 * xs * y);
            const double coef1 = expx2erfcxy - c * y * sum1;
            const double coef2 = c * xs * expx2;
            retr = coef1 * cos2xy + coef2 * sinxy * sinc_(xs * y, sinxy);
            reti = coef2 * sinc_(2 * xs * y, sin2xy) - coef1 * sin2xy;
        }
    } else {
        if (isnan(x)) { *wr_out = x; *wi_out = x; return; }
        if (isnan(y)) { *wr_out = y; *wi_out = y; return; }
        retr = exp(-x * x);
        reti = 0;
        const double n0 = floor(x / a + 0.5);
        const double dx = a * n0 - x;
        sum3 = exp(-dx * dx) / (a2 * (n0 * n0) + y * y);
        sum5 = a * n0 * sum3;
        const double exp1 = exp(4 * a * dx);
        double exp1dn = 1;
        int dn;
        for (dn = 1; n0 - dn > 0; ++dn) {
            const double np = n0 + dn, nm = n0 - dn;
            double tp = exp(-((a * dn + dx) * (a * dn + dx)));
            double tm = tp * (exp1dn *= exp1);
            tp /= (a2 * (np * np) + y * y);
            tm /= (a2 * (nm * nm) + y * y);
            sum3 += tp + tm;
            sum5 += a * (np * tp + nm * tm);
            if (a * (np * tp + nm * tm) < relerr * sum5) goto finish;
        }
        while (1) {
            /* the reference advances dn before evaluating the exponent here
             * (src/Faddeeva.cpp:961-963); the terms are below 1e-40 of sum5 */
            const double np = n0 + dn++;
            const double tp = exp(-((a * dn + dx) * (a * dn + dx))) / (a2 * (np * np) + y * y);
            sum3 += tp;
            sum5 += a * np * tp;
            if (a * np * tp < relerr * sum5) goto finish;
        }
    }
finish:
    *wr_out = retr + (0.5 * c) * y * (sum2 + sum3);
    *wi_out = reti + (0.5 * c) * copysign(sum5 - sum4, zr);
}

/* Im w(x) = (2/sqrt(pi)) * Dawson(x).  |x| < 0.5: Dawson's Taylor series
 * x * sum (-2x^2)^n / (2n+1)!! (the sums below cancel there); otherwise
 * the general evaluation at y = 0. */
double orc_w_im(double x)
{
    if (fabs(x) < 0.5) {
        const double m2x2 = -2.0 * x * x;
        double term = x, sum = x;
        for (int n = 1; n < 40; n++) {
            term = term * m2x2 / (2 * n + 1);
            sum += term;
            if (fabs(term) < 1e-17 * fabs(sum)) break;
        }
        return 2 * ISPI * sum;
    }
    double wr, wi;
    w_general(x, 0.0, &wr, &wi);
    return wi;
}

/* src/Faddeeva.cpp:682-689 then the general evaluation. */
void orc_faddeeva_w(double re, double im, double *wr, double *wi)
{
    if (re == 0.0) {
        *wr = orc_erfcx(im);
        *wi = re;
        return;
    }
    if (im == 0) {
        *wr = exp(-(re * re));
        *wi = orc_w_im(re);
        return;
    }
    w_general(re, im, wr, wi);
}

/* Bunch::BassettiErskine1, src/Bunch.cpp:119-170 */
void orc_bassetti_erskine(double posx, double posy, double sx, double sy,
                          double *fx, double *fy)
{
    const double ax = fabs(posx), ay = fabs(posy);
    const double sigma = sqrt(2 * pow(sx, 2) - 2 * pow(sy, 2));
    const double r = sy / sx;
    const double coeffBE = -1 * sqrt(PI) / sigma;
    double w1r, w1i, w2r, w2i;
    orc_faddeeva_w(ax / sigma, ay / sigma, &w1r, &w1i);
    orc_faddeeva_w(r * ax / sigma, ay / sigma / r, &w2r, &w2i);
    const double z3 = -pow(ax / sx, 2) / 2 - pow(ay / sy, 2) / 2;
    const double w3 = -exp(z3);
    const double w0r = coeffBE * (w1r + w3 * w2r);
    const double w0i = coeffBE * (w1i + w3 * w2i);
    double tfx = w0i, tfy = w0r;
    if (posx <= 0) tfx = -tfx;
    if (posy <= 0) tfy = -tfy;
    *fx = tfx;
    *fy = tfy;
}

/* Bunch::GaussianField, src/Bunch.cpp:202-211 */
void orc_gaussian_field(double posx, double posy, double sx, double sy,
                        double *fx, double *fy)
{
    (void)sy;
    const double r2 = pow(posx, 2) + pow(posy, 2);
    const double sigma = sx;
    const double sigma2 = sigma * sigma;
    *fx = -posx / r2 * (1 - exp(-r2 / 2 / sigma2));
    *fy = -posy / r2 * (1 - exp(-r2 / 2 / sigma2));
}

/* Field branch of MPBunch::SSIonBunchInteraction (src/MPBunch.cpp:1128-1147). */
void orc_ion_field(double x, double y, double sx, double sy, double *fx, double *fy)
{
    if (fabs(x / sx) + fabs(y / sy) < 1.0E-5) {
        *fx = 0; *fy = 0;
    } else if ((sx - sy) / sy > 1.e-4) {
        orc_bassetti_erskine(x, y, sx, sy, fx, fy);
    } else if ((sx - sy) / sy < -1.e-4) {
        orc_bassetti_erskine(y, x, sy, sx, fy, fx);
    } else {
        orc_gaussian_field(x, y, sx, sy, fx, fy);
    }
}
