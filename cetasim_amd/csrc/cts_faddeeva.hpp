// cts_faddeeva.hpp -- FP64 Faddeeva function w(z) and the Bassetti-Erskine
// field on the device.
//
// Region structure of the reference's vendored Faddeeva package (S. G.
// Johnson, MIT; src/Faddeeva.cpp:682-974): axis special cases, continued
// fraction (Gautschi 1970, Poppe & Wijers 1990) for large |z|, large-|x|
// sums near the real axis.  Where the package sums Zaghloul & Ali's
// series (ACM TOMS Algorithm 916, |x| < 10), the upper half-plane uses
// Weideman's rational approximation instead (J. A. C. Weideman, SIAM J.
// Numer. Anal. 31 (1994) 1497-1518): w(z) = 2 p(Z)/(L - iz)^2 +
// pi^{-1/2}/(L - iz), Z = (L + iz)/(L - iz), p of degree N - 1 = 39,
// L = (N/sqrt 2)^{1/2}; |w| relative error below 4e-14 there (1e-15
// typical), one division and 40 complex multiply-adds instead of a
// division-heavy series; the lower half-plane keeps the series.  erfcx and Im w(x) on the axes are this project's own
// evaluations (exp(x^2)erfc(x) with an exact x^2 split, Laplace continued
// fraction for x >= 25, Dawson's Taylor series for |x| < 0.5).  FP64
// throughout: the asymptotic and lower-half-plane branches cancel.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace cts {

// exp(-a^2 n^2), n = 1..64, a = pi/sqrt(-log(eps/2)), and {a, c = 2a/pi,
// a^2}.  The library is built without relocatable device code, so every
// translation unit that evaluates w(z) holds its own copy of the tables and
// exports an upload function that fills it; cts_create calls each of them.
static __constant__ double c_ea2n2[64];
static __constant__ double c_fad[3];
constexpr int kWeidN = 40;
static __constant__ double c_weid[kWeidN + 1];  // p's coefficients, highest degree first, then L

// Host: the tables above.
struct FadTables {
    double ea2n2[64];
    double fad[3];
    double weid[kWeidN + 1];
};
inline void build_fad_tables(FadTables &t)
{
    const double pi = 3.141592653589793;
    // Algorithm 916 constants and exp(-a^2 n^2)
    t.fad[0] = pi / std::sqrt(-std::log(2.220446049250313e-16 * 0.5));
    t.fad[1] = (2 / pi) * t.fad[0];
    t.fad[2] = t.fad[0] * t.fad[0];
    for (int n = 1; n <= 63; n++) t.ea2n2[n - 1] = std::exp(-t.fad[2] * (double)(n * n));
    t.ea2n2[63] = 0.0;
    // Weideman: f(t) = e^{-t^2}(L^2 + t^2) at t = L tan(theta/2), theta = k pi/M,
    // k = -M+1..M-1 (and 0 at k = -M), M = 2N; the coefficients are the real
    // parts of the 2M-point DFT of f in FFT order, terms 1..N
    const int N = kWeidN, M = 2 * N, M2 = 2 * M;
    const double L = std::sqrt(N / std::sqrt(2.0));
    double f[2 * 2 * kWeidN];
    f[0] = 0.0;
    for (int k = -M + 1; k <= M - 1; k++) {
        const double tt = L * std::tan(0.5 * k * pi / M);
        f[k + M] = std::exp(-tt * tt) * (L * L + tt * tt);
    }
    double g[2 * 2 * kWeidN];
    for (int j = 0; j < M2; j++) g[j] = f[(j + M) % M2];  // fftshift
    for (int m = 1; m <= N; m++) {
        double acc = 0.0;
        for (int j = 0; j < M2; j++) acc += g[j] * std::cos(2 * pi * (double)((j * m) % M2) / M2);
        t.weid[N - m] = acc / M2;
    }
    t.weid[N] = L;
}

constexpr double kIsPi = 0.56418958354775628694807945156;  // 1/sqrt(pi)

__device__ __forceinline__ double fad_sinc(double x, double sinx)
{
    return fabs(x) < 1e-4 ? 1 - (0.1666666666666666666667) * x * x : sinx / x;
}
__device__ __forceinline__ double fad_sinh_taylor(double x)
{
    return x * (1 + (x * x) * (0.1666666666666666666667 + 0.00833333333333333333333 * (x * x)));
}

__device__ inline double erfcx_pos(double x)
{
    if (x < 25.0) {
        const double h = x * x;
        const double l = fma(x, x, -h);
        return exp(h) * (1.0 + l) * erfc(x);
    }
    if (isinf(x)) return 0.0;
    double t = x;
    for (int k = 24; k >= 1; k--) t = x + (0.5 * k) / t;
    return kIsPi / t;
}

__device__ inline double cts_erfcx(double x)
{
    if (isnan(x)) return x;
    if (x < 0) {
        if (x < -26.7) return HUGE_VAL;
        return 2 * exp(x * x) - erfcx_pos(-x);
    }
    return erfcx_pos(x);
}

// Weideman's approximation for Im z >= 0 (see the header comment).
__device__ __forceinline__ void w_weideman(double zr, double zi, double &wr, double &wi)
{
#pragma clang fp contract(fast)
    const double L = c_weid[kWeidN];
    // 1/(L - iz) = u, Z = (L + iz) u
    const double dr = L + zi, di = -zr;
    const double inv = 1.0 / (dr * dr + di * di);
    const double ur = dr * inv, ui = -di * inv;
    const double nr = L - zi, ni = zr;
    const double Zr = nr * ur - ni * ui, Zi = nr * ui + ni * ur;
    double pr = c_weid[0], pi = 0.0;
#pragma unroll 8
    for (int k = 1; k < kWeidN; k++) {
        const double t = pr * Zr - pi * Zi + c_weid[k];
        pi = pr * Zi + pi * Zr;
        pr = t;
    }
    const double u2r = ur * ur - ui * ui, u2i = 2.0 * ur * ui;
    wr = 2.0 * (pr * u2r - pi * u2i) + kIsPi * ur;
    wi = 2.0 * (pr * u2i + pi * u2r) + kIsPi * ui;
}

// Continued fraction / asymptotic forms for large |z| (the package's
// first region, src/Faddeeva.cpp:740-790), for xs = Re z reflected to
// Im z >= 0 and ya = |Im z|; x = |xs|.
__device__ __forceinline__ void w_cf(double xs, double ya, double x, double &retr, double &reti, bool &inf_case)
{
    inf_case = false;
    if (x + ya > 4000) {
        if (x + ya > 1e7) {
            if (x > ya) {
                const double yax = ya / xs;
                const double denom = kIsPi / (xs + yax * ya);
                retr = denom * yax; reti = denom;
            } else if (isinf(ya)) {
                inf_case = true;
                retr = 0; reti = 0;
            } else {
                const double xya = xs / ya;
                const double denom = kIsPi / (xya * xs + ya);
                retr = denom; reti = denom * xya;
            }
        } else {
            const double dr = xs * xs - ya * ya - 0.5, di = 2 * xs * ya;
            const double denom = kIsPi / (dr * dr + di * di);
            retr = denom * (xs * di - ya * dr);
            reti = denom * (xs * dr + ya * di);
        }
    } else {
        // depth fit of the reference's package (src/Faddeeva.cpp:758-759)
        double nu = floor(3.9 + 11.398 / (0.08254 * x + 0.1421 * ya + 0.2023));
        double wr = xs, wi = ya;
        for (nu = 0.5 * (nu - 1); nu > 0.4; nu -= 0.5) {
            const double denom = nu / (wr * wr + wi * wi);
            wr = xs - wr * denom;
            wi = ya + wi * denom;
        }
        const double denom = kIsPi / (wr * wr + wi * wi);
        retr = denom * wi;
        reti = denom * wr;
    }
}

// The package's sums for |x| >= 10 next to the real axis
// (src/Faddeeva.cpp:880-945); returns w.
__device__ inline void w_large_x(double zr, double y, double x, double &wr_out, double &wi_out)
{
    const double a = c_fad[0], c = c_fad[1], a2 = c_fad[2];
    const double relerr = 2.220446049250313e-16;
    double sum3, sum5;
    const double retr = exp(-x * x);
    const double n0 = floor(x / a + 0.5);
    const double dx = a * n0 - x;
    sum3 = exp(-dx * dx) / (a2 * (n0 * n0) + y * y);
    sum5 = a * n0 * sum3;
    const double exp1 = exp(4 * a * dx);
    double exp1dn = 1;
    int dn;
    bool done = false;
    for (dn = 1; n0 - dn > 0; ++dn) {
        const double np = n0 + dn, nm = n0 - dn;
        double tp = exp(-((a * dn + dx) * (a * dn + dx)));
        double tm = tp * (exp1dn *= exp1);
        tp /= (a2 * (np * np) + y * y);
        tm /= (a2 * (nm * nm) + y * y);
        sum3 += tp + tm;
        sum5 += a * (np * tp + nm * tm);
        if (a * (np * tp + nm * tm) < relerr * sum5) { done = true; break; }
    }
    for (int guard = 0; !done && guard < 200; guard++) {
        const double np = n0 + dn++;
        const double tp = exp(-((a * dn + dx) * (a * dn + dx))) / (a2 * (np * np) + y * y);
        sum3 += tp;
        sum5 += a * np * tp;
        if (a * np * tp < relerr * sum5) done = true;
    }
    wr_out = retr + (0.5 * c) * y * sum3;
    wi_out = (0.5 * c) * copysign(sum5, zr);
}

__device__ __forceinline__ bool w_cf_region(double x, double ya)
{
    return ya > 7 || (x > 6 && (ya > 0.1 || (x > 8 && ya > 1e-10) || x > 28));
}

// w(z) for Re z >= 0, Im z >= 0, both finite and not on an axis (the
// arguments of the Bassetti-Erskine field): continued fraction far out,
// Weideman's approximation near the beam, the large-x sums near the axis.
__device__ inline void w_quadrant1(double x, double y, double &wr, double &wi)
{
    if (w_cf_region(x, y)) {
        bool inf_case;
        w_cf(x, y, x, wr, wi, inf_case);
        if (inf_case) { wr = 0; wi = 0; }
    } else if (x < 10) {
        w_weideman(x, y, wr, wi);
    } else {
        w_large_x(x, y, x, wr, wi);
    }
}

__device__ inline void w_general(double zr, double zi, double &wr_out, double &wi_out)
{
    const double a = c_fad[0], c = c_fad[1], a2 = c_fad[2];
    const double relerr = 2.220446049250313e-16;
    const double x = fabs(zr);
    const double y = zi, ya = fabs(y);
    double retr = 0, reti = 0;
    double sum1 = 0, sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0;

    if (w_cf_region(x, ya)) {
        const double xs = y < 0 ? -zr : zr;
        bool inf_case;
        w_cf(xs, ya, x, retr, reti, inf_case);
        if (inf_case) {
            if (isnan(x) || y < 0) { wr_out = NAN; wi_out = NAN; }
            else { wr_out = 0; wi_out = 0; }
            return;
        }
        if (y < 0) {
            // w(z) = 2 exp(-z^2) - w(-z); exp(-inf + i*anything) = 0
            const double re = (ya - xs) * (xs + ya);
            double er_c = 0.0, er_s = 0.0;
            if (!(isinf(re) && re < 0)) {
                const double er = exp(re);
                double s, cph;
                sincos(2 * xs * y, &s, &cph);
                er_c = er * cph;
                er_s = er * s;
            }
            wr_out = 2.0 * er_c - retr;
            wi_out = 2.0 * er_s - reti;
        } else {
            wr_out = retr; wi_out = reti;
        }
        return;
    } else if (x < 10) {
        double prod2ax = 1, prodm2ax = 1, expx2;
        if (isnan(y)) { wr_out = y; wi_out = y; return; }
        if (y >= 0) {
            w_weideman(zr, y, wr_out, wi_out);
            return;
        }
        if (x < 5e-4) {
            const double x2 = x * x;
            expx2 = 1 - x2 * (1 - 0.5 * x2);
            const double ax2 = (2 * a) * x;
            const double exp2ax = 1 + ax2 * (1 + ax2 * (0.5 + 0.166666666666666666667 * ax2));
            const double expm2ax = 1 - ax2 * (1 - ax2 * (0.5 - 0.166666666666666666667 * ax2));
            for (int n = 1; n <= 64; ++n) {
                const double coef = c_ea2n2[n - 1] * expx2 / (a2 * (n * n) + y * y);
                prod2ax *= exp2ax;
                prodm2ax *= expm2ax;
                sum1 += coef;
                sum2 += coef * prodm2ax;
                sum3 += coef * prod2ax;
                sum5 += coef * (2 * a) * n * fad_sinh_taylor((2 * a) * n * x);
                if (coef * prod2ax < relerr * sum3) break;
            }
        } else {
            expx2 = exp(-x * x);
            const double exp2ax = exp((2 * a) * x), expm2ax = 1 / exp2ax;
            for (int n = 1; n <= 64; ++n) {
                const double coef = c_ea2n2[n - 1] * expx2 / (a2 * (n * n) + y * y);
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
        const double expx2erfcxy = y > -6 ? expx2 * cts_erfcx(y) : 2 * exp(y * y - x * x);
        if (y > 5) {
            const double sinxy = sin(x * y);
            retr = (expx2erfcxy - c * y * sum1) * cos(2 * x * y) +
                   (c * x * expx2) * sinxy * fad_sinc(x * y, sinxy);
            reti = 0;
        } else {
            const double xs = zr;
            const double sinxy = sin(xs * y);
            double sin2xy, cos2xy;
            sincos(2 * xs * y, &sin2xy, &cos2xy);
            const double coef1 = expx2erfcxy - c * y * sum1;
            const double coef2 = c * xs * expx2;
            retr = coef1 * cos2xy + coef2 * sinxy * fad_sinc(xs * y, sinxy);
            reti = coef2 * fad_sinc(2 * xs * y, sin2xy) - coef1 * sin2xy;
        }
    } else {
        if (isnan(x)) { wr_out = x; wi_out = x; return; }
        if (isnan(y)) { wr_out = y; wi_out = y; return; }
        w_large_x(zr, y, x, wr_out, wi_out);
        return;
    }
    wr_out = retr + (0.5 * c) * y * (sum2 + sum3);
    wi_out = reti + (0.5 * c) * copysign(sum5 - sum4, zr);
}

__device__ inline double cts_w_im(double x)
{
    if (fabs(x) < 0.5) {
        const double m2x2 = -2.0 * x * x;
        double term = x, sum = x;
        for (int n = 1; n < 40; n++) {
            term = term * m2x2 / (2 * n + 1);
            sum += term;
            if (fabs(term) < 1e-17 * fabs(sum)) break;
        }
        return 2 * kIsPi * sum;
    }
    double wr, wi;
    w_general(x, 0.0, wr, wi);
    return wi;
}

__device__ inline void faddeeva_w(double re, double im, double &wr, double &wi)
{
    if (re == 0.0) { wr = cts_erfcx(im); wi = re; return; }
    if (im == 0) { wr = exp(-(re * re)); wi = cts_w_im(re); return; }
    w_general(re, im, wr, wi);
}

// w(z) for Re z >= 0, Im z >= 0 (any finite values, axes included).
__device__ inline void faddeeva_w_q1(double re, double im, double &wr, double &wi)
{
    if (re == 0.0) { wr = cts_erfcx(im); wi = re; return; }
    if (im == 0) { wr = exp(-(re * re)); wi = cts_w_im(re); return; }
    w_quadrant1(re, im, wr, wi);
}

// Bunch::BassettiErskine1 (src/Bunch.cpp:119-170)
__device__ inline void bassetti_erskine(double posx, double posy, double sx, double sy,
                                        double &fx, double &fy)
{
    const double ax = fabs(posx), ay = fabs(posy);
    const double sigma = sqrt(2 * (sx * sx) - 2 * (sy * sy));
    const double r = sy / sx;
    const double coeffBE = -1 * sqrt(kPi) / sigma;
    double w1r, w1i, w2r, w2i;
    faddeeva_w_q1(ax / sigma, ay / sigma, w1r, w1i);
    faddeeva_w_q1(r * ax / sigma, ay / sigma / r, w2r, w2i);
    const double tx = ax / sx, ty = ay / sy;
    const double z3 = -(tx * tx) / 2 - (ty * ty) / 2;
    const double w3 = -exp(z3);
    const double w0r = coeffBE * (w1r + w3 * w2r);
    const double w0i = coeffBE * (w1i + w3 * w2i);
    fx = posx <= 0 ? -w0i : w0i;
    fy = posy <= 0 ? -w0r : w0r;
}

// ---------------------------------------------------------------- kicks
// The field of one Gaussian beam evaluated at many points (the kicks of a
// bunch passage): the beam's constants once, then per point
// multiplications by reciprocals, and both w(z) of the Bassetti-Erskine
// form by Weideman's approximation in one interleaved evaluation.  The
// kicks take it on all of Re z < 100, 0 <= Im z < 1000, where its |w|
// relative error stays below 3e-14 (checked against the reference's
// package on the quadrant's regions: largest next to the real axis for
// 6 < Re z < 10), so that the lanes of a wave do not split between the
// package's regions; points beyond take faddeeva_w_q1.  Same formulas as
// ion_field otherwise, rounding within a few ulps.
__device__ inline void gaussian_field(double posx, double posy, double sx, double &fx, double &fy);

struct BeField {
    int32_t mode;          // 0: Bassetti-Erskine, 1: the same with x and y exchanged, 2: round beam
    double isx, isy;       // 1/sx, 1/sy of the beam as given (zero test)
    double isig, r, ir;    // for the ordered beam SX > SY: 1/sqrt(2 (SX^2 - SY^2)), SY/SX, SX/SY
    double coeff;          // -sqrt(pi) * isig
    double hx, hy;         // 1/(2 SX^2), 1/(2 SY^2)
    double sx;             // round beam: sigma_x
};

__device__ __host__ inline BeField be_plan(double sx, double sy)
{
    BeField F;
    F.isx = 1.0 / sx;
    F.isy = 1.0 / sy;
    F.mode = ((sx - sy) / sy > 1.e-4) ? 0 : ((sx - sy) / sy < -1.e-4) ? 1 : 2;
    const double SX = F.mode == 1 ? sy : sx, SY = F.mode == 1 ? sx : sy;
    F.isig = 1.0 / sqrt(2 * (SX * SX) - 2 * (SY * SY));
    F.r = SY / SX;
    F.ir = SX / SY;
    F.coeff = -1 * sqrt(kPi) * F.isig;
    F.hx = 0.5 / (SX * SX);
    F.hy = 0.5 / (SY * SY);
    F.sx = sx;
    return F;
}

__device__ __forceinline__ bool w_kick_region(double x, double y)
{
    return x < 100.0 && y < 1000.0;
}

// w_weideman at two points, the two Horner recurrences interleaved.
__device__ __forceinline__ void w_weideman2(const double zr[2], const double zi[2], double wr[2], double wi[2])
{
#pragma clang fp contract(fast)
    const double L = c_weid[kWeidN];
    double ur[2], ui[2], Zr[2], Zi[2], pr[2], pi[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const double dr = L + zi[q], di = -zr[q];
        const double inv = 1.0 / (dr * dr + di * di);
        ur[q] = dr * inv;
        ui[q] = -di * inv;
        const double nr = L - zi[q], ni = zr[q];
        Zr[q] = nr * ur[q] - ni * ui[q];
        Zi[q] = nr * ui[q] + ni * ur[q];
        pr[q] = c_weid[0];
        pi[q] = 0.0;
    }
#pragma unroll 8
    for (int k = 1; k < kWeidN; k++) {
        const double ck = c_weid[k];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double t = pr[q] * Zr[q] - pi[q] * Zi[q] + ck;
            pi[q] = pr[q] * Zi[q] + pi[q] * Zr[q];
            pr[q] = t;
        }
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const double u2r = ur[q] * ur[q] - ui[q] * ui[q], u2i = 2.0 * ur[q] * ui[q];
        wr[q] = 2.0 * (pr[q] * u2r - pi[q] * u2i) + kIsPi * ur[q];
        wi[q] = 2.0 * (pr[q] * u2i + pi[q] * u2r) + kIsPi * ui[q];
    }
}

// ion_field with the beam's constants from be_plan.
__device__ __forceinline__ void be_field(double x, double y, const BeField &F, double &fx, double &fy)
{
    if (fabs(x) * F.isx + fabs(y) * F.isy < 1.0E-5) {
        fx = 0;
        fy = 0;
        return;
    }
    if (F.mode == 2) {
        gaussian_field(x, y, F.sx, fx, fy);
        return;
    }
    const bool sw = F.mode == 1;
    const double X = sw ? y : x, Y = sw ? x : y;
    const double ax = fabs(X), ay = fabs(Y);
    double zr[2], zi[2], wr[2], wi[2];
    zr[0] = ax * F.isig;
    zi[0] = ay * F.isig;
    zr[1] = F.r * zr[0];
    zi[1] = zi[0] * F.ir;
    if (w_kick_region(zr[0], zi[0]) && w_kick_region(zr[1], zi[1])) {
        w_weideman2(zr, zi, wr, wi);
    } else {
        faddeeva_w_q1(zr[0], zi[0], wr[0], wi[0]);
        faddeeva_w_q1(zr[1], zi[1], wr[1], wi[1]);
    }
    const double w3 = -exp(-(ax * ax) * F.hx - (ay * ay) * F.hy);
    const double w0r = F.coeff * (wr[0] + w3 * wr[1]);
    const double w0i = F.coeff * (wi[0] + w3 * wi[1]);
    const double fX = X <= 0 ? -w0i : w0i, fY = Y <= 0 ? -w0r : w0r;
    fx = sw ? fY : fX;
    fy = sw ? fX : fY;
}

// Bunch::GaussianField (src/Bunch.cpp:202-211)
__device__ inline void gaussian_field(double posx, double posy, double sx, double &fx, double &fy)
{
    const double r2 = posx * posx + posy * posy;
    const double sigma2 = sx * sx;
    const double e = 1 - exp(-r2 / 2 / sigma2);
    fx = -posx / r2 * e;
    fy = -posy / r2 * e;
}

// Branches of MPBunch::SSIonBunchInteraction (src/MPBunch.cpp:1128-1147)
__device__ inline void ion_field(double x, double y, double sx, double sy, double &fx, double &fy)
{
    if (fabs(x / sx) + fabs(y / sy) < 1.0E-5) {
        fx = 0; fy = 0;
    } else if ((sx - sy) / sy > 1.e-4) {
        bassetti_erskine(x, y, sx, sy, fx, fy);
    } else if ((sx - sy) / sy < -1.e-4) {
        bassetti_erskine(y, x, sy, sx, fy, fx);
    } else {
        gaussian_field(x, y, sx, fx, fy);
    }
}

}  // namespace cts
