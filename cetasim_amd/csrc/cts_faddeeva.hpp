// cts_faddeeva.hpp -- FP64 Faddeeva function w(z) = exp(-z^2) erfc(-iz) and
// the Bassetti-Erskine field on the device.
//
// w(z) everywhere (faddeeva_w; host and device).  The first quadrant
// x = Re z >= 0, y = Im z >= 0 is evaluated directly; the rest follows from
// w(-x + iy) = conj w(x + iy) and, below the real axis,
// w(z) = 2 exp(-z^2) - w(-z).  By region of the first quadrant:
//  * |z| < 1: the Maclaurin series w(z) = sum_n (iz)^n / Gamma(n/2 + 1).
//  * y >= 8 or x >= 15: the Laplace continued fraction
//    w(z) = (i/sqrt(pi)) / (z - (1/2) / (z - 1 / (z - (3/2) / (z - ...)))),
//    its depth set by |z| (its truncation error behaves as the asymptotic
//    series' next term, (2n+1)!!/(2|z|^2)^(n+1)); next to the real axis it is
//    completed by exp(-z^2), which the fraction does not carry; beyond
//    |z| = 1e7 the two leading terms of the asymptotic series.
//  * elsewhere: the trapezoidal rule for w(z) = (i/pi) int exp(-t^2)/(z - t) dt
//    on nodes t = c + n h, h = 0.4, |t| <= 7, with the pole correction
//    2 exp(-z^2) / (1 - exp(-2 pi i (z - c) / h)) (Chiarella & Reichel,
//    Math. Comp. 22 (1968) 137; Matta & Reichel, Math. Comp. 25 (1971) 339);
//    what remains is the aliasing error exp(-pi^2/h^2) ~ 2e-27.  For y >= 0.5
//    the nodes lie symmetric about 0 (c = h/2) and are summed in pairs, which
//    leaves Im w proportional to x; below, they are centred on x (x - c is an
//    odd multiple of h/2), so no node comes near the pole and the correction
//    reduces to exp(-z^2) (1 - tanh(pi y / h)); Re w is then a sum of positive
//    terms proportional to y plus that correction, exact to the last digits
//    however small it is.
// On the axes: w(iy) = erfcx(y), and w(x) = exp(-x^2) + i Im w(x) with Im w(x)
// from Dawson's Maclaurin series for |x| < 0.5, the centred trapezoidal rule
// for |x| < 15 and the continued fraction beyond.
//
// Near the beam the kicks take Weideman's rational approximation instead
// (J. A. C. Weideman, SIAM J. Numer. Anal. 31 (1994) 1497-1518; be_field
// below), one division and 40 complex multiply-adds per point.  FP64
// throughout: the far-field and lower-half-plane forms cancel.
#pragma once

#include <hip/hip_runtime.h>

#include "cts_device.hpp"

#include <cmath>

namespace cts {

// Weideman's coefficients.  The library is built without relocatable device
// code, so every translation unit that evaluates the kick field holds its
// own copy of the table and exports an upload function that fills it;
// cts_create calls each of them.
constexpr int kWeidN = 40;
static __constant__ double c_weid[kWeidN + 1];  // p's coefficients, highest degree first, then L

// Host: the table above.
struct FadTables {
    double weid[kWeidN + 1];
};
inline void build_fad_tables(FadTables &t)
{
    const double pi = 3.141592653589793;
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
constexpr double kFadH = 0.4;                              // trapezoid step
constexpr double kFadTMax = 7.0;                           // nodes |t| <= 7
constexpr double kFadPiOverH = 7.8539816339744830962;      // pi / h

// exp(-t^2) with t^2 split exactly (t^2 = hi + lo, fma), so the large
// arguments of the node weights keep full relative accuracy
__host__ __device__ inline double fad_gauss(double t)
{
    const double hi = t * t;
    const double lo = fma(t, t, -hi);
    return exp(-hi) * (1.0 - lo);
}

__host__ __device__ inline double erfcx_pos(double x)
{
    if (x < 25.0) {
        const double h = x * x;
        const double l = fma(x, x, -h);
        return exp(h) * (1.0 + l) * erfc(x);
    }
    if (std::isinf(x)) return 0.0;
    double t = x;
    for (int k = 24; k >= 1; k--) t = x + (0.5 * k) / t;
    return kIsPi / t;
}

__host__ __device__ inline double cts_erfcx(double x)
{
    if (std::isnan(x)) return x;
    if (x < 0) {
        if (x < -26.7) return HUGE_VAL;
        return 2 * exp(x * x) - erfcx_pos(-x);
    }
    return erfcx_pos(x);
}

// Maclaurin series of w for |z| < 1 (48 terms: the last is below 1e-17 |w|).
// (iz)^n by complex products keeps Im w proportional to x next to the
// imaginary axis; 1/Gamma(n/2 + 1) by c_{n+2} = c_n / (n/2 + 1).
__host__ __device__ inline void fad_maclaurin(double x, double y, double &wr, double &wi)
{
    const double ur = -y, ui = x;  // iz
    double pr = 1.0, pi = 0.0;      // (iz)^n
    double ce = 1.0, co = 2.0 * kIsPi;  // 1/Gamma(n/2+1), n even / odd
    double sr = 1.0, si = 0.0;
    for (int n = 1; n < 48; n++) {
        const double t = pr * ur - pi * ui;
        pi = pr * ui + pi * ur;
        pr = t;
        double c;
        if (n & 1) {
            c = co;
            co = co / (0.5 * n + 1.0);
        } else {
            ce = ce / (0.5 * n);
            c = ce;
        }
        sr += c * pr;
        si += c * pi;
    }
    wr = sr;
    wi = si;
}

// i / (sqrt(pi) z) (1 + 1/(2 z^2)) for |z| >= 1e7, with Smith's division
// (no overflow for |z| up to DBL_MAX).
__host__ __device__ inline void fad_asymptotic(double x, double y, double &wr, double &wi)
{
    double ir, ii;  // 1/z
    if (fabs(x) >= fabs(y)) {
        const double r = y / x, d = x + y * r;
        ir = 1.0 / d;
        ii = -r / d;
    } else {
        const double r = x / y, d = x * r + y;
        ir = r / d;
        ii = -1.0 / d;
    }
    // 1 + (1/z)^2 / 2
    const double fr = 1.0 + 0.5 * (ir * ir - ii * ii), fi = ir * ii;
    // i (1/z) f / sqrt(pi)
    const double gr = ir * fr - ii * fi, gi = ir * fi + ii * fr;
    wr = -kIsPi * gi;
    wi = kIsPi * gr;
}

// Laplace continued fraction with n levels, bottom up: r = z, then
// r = z - (k/2)/r for k = n..1; w = i / (sqrt(pi) r).
__host__ __device__ inline void fad_cf(double x, double y, int n, double &wr, double &wi)
{
    double rr = x, ri = y;
    for (int k = n; k >= 1; k--) {
        const double d = 0.5 * k / (rr * rr + ri * ri);
        rr = x - rr * d;
        ri = y + ri * d;
    }
    const double d = kIsPi / (rr * rr + ri * ri);
    wr = ri * d;
    wi = rr * d;
}

__host__ __device__ inline int fad_cf_depth(double r2)
{
    return r2 < 100.0 ? 18 : r2 < 400.0 ? 11 : r2 < 2500.0 ? 7 : r2 < 1e5 ? 4 : 2;
}

// Trapezoidal rule, nodes +-(n + 1/2) h summed in pairs, with the pole
// correction; y >= 0.5.  With t^2 the node, A = x^2 - y^2 - t^2, B = 2xy:
// i (1/(z - t) + 1/(z + t)) = 2 (y (|z|^2 + t^2) + i x (|z|^2 - t^2)) / (A^2 + B^2).
__host__ __device__ inline void fad_trap_sym(double x, double y, double &wr, double &wi)
{
    const double z2 = x * x + y * y, d2 = x * x - y * y, b2 = 4.0 * x * x * y * y;
    double sr = 0.0, si = 0.0;
    for (int n = 0; n < 18; n++) {
        const double t = (n + 0.5) * kFadH, t2 = t * t;
        const double a = d2 - t2;
        const double g = fad_gauss(t) / (a * a + b2);
        sr += g * (z2 + t2);
        si += g * (z2 - t2);
    }
    const double k = 2.0 * kFadH / 3.14159265358979323846;
    wr = k * y * sr;
    wi = k * x * si;
    // 2 exp(-z^2) / (1 + exp(2 pi y/h) exp(-i theta)), theta = 2 pi x/h:
    // with q = exp(-2 pi y/h),
    // = 2 q exp(y^2 - x^2) (exp(i(theta - 2xy)) + q exp(-2ixy)) / (1 + 2 q cos theta + q^2)
    const double ex = (y - x) * (y + x) - 2.0 * kFadPiOverH * y;
    if (ex > -745.0) {
        const double q = exp(-2.0 * kFadPiOverH * y);
        const double th = 2.0 * kFadPiOverH * x, p = 2.0 * x * y;
        const double s = 2.0 * exp(ex) / (1.0 + 2.0 * q * cos(th) + q * q);
        wr += s * (cos(th - p) + q * cos(p));
        wi += s * (sin(th - p) - q * sin(p));
    }
}

// Trapezoidal rule on nodes x + (m + 1/2) h, |t| <= 7, with the pole
// correction exp(-z^2) (1 - tanh(pi y/h)); y < 0.5.  Each node adds
// (h/pi) exp(-t^2) (y - i s) / (s^2 + y^2), s = t - x.
__host__ __device__ inline void fad_trap_centred(double x, double y, double &wr, double &wi)
{
    const int m0 = (int)ceil((-kFadTMax - x) / kFadH - 0.5);
    const int m1 = (int)floor((kFadTMax - x) / kFadH - 0.5);
    const double y2 = y * y;
    double sr = 0.0, si = 0.0;
    for (int m = m0; m <= m1; m++) {
        const double s = (m + 0.5) * kFadH;
        const double g = fad_gauss(x + s) / (s * s + y2);
        sr += g;
        si += g * s;
    }
    const double k = kFadH / 3.14159265358979323846;
    wr = k * y * sr;
    wi = -k * si;
    const double ex = (y - x) * (y + x);
    if (ex > -745.0) {
        const double e = exp(ex) * (2.0 / (1.0 + exp(2.0 * kFadPiOverH * y)));
        const double p = 2.0 * x * y;
        wr += e * cos(p);
        wi -= e * sin(p);
    }
}

// w(z) for x >= 0, y >= 0, both finite or infinite (not NaN).
__host__ __device__ inline void fad_q1(double x, double y, double &wr, double &wi)
{
    if (std::isinf(x) || std::isinf(y)) { wr = 0.0; wi = 0.0; return; }
    if (x >= 1e7 || y >= 1e7) { fad_asymptotic(x, y, wr, wi); return; }
    const double r2 = x * x + y * y;
    if (y >= 8.0 || x >= 15.0) {
        fad_cf(x, y, fad_cf_depth(r2), wr, wi);
        if (x >= 15.0 && y < 1.0) {
            // the exponentially small exp(-z^2) next to the real axis
            const double ex = (y - x) * (y + x);
            if (ex > -745.0) {
                const double e = exp(ex), p = 2.0 * x * y;
                wr += e * cos(p);
                wi -= e * sin(p);
            }
        }
        return;
    }
    if (r2 < 1.0) { fad_maclaurin(x, y, wr, wi); return; }
    if (y >= 0.5) fad_trap_sym(x, y, wr, wi);
    else fad_trap_centred(x, y, wr, wi);
}

// Im w(x) = 2 Dawson(x) / sqrt(pi) on the real axis.
__host__ __device__ inline double cts_w_im(double x)
{
    if (std::isnan(x)) return x;
    const double ax = fabs(x);
    double v;
    if (ax < 0.5) {
        const double m2x2 = -2.0 * x * x;
        double term = x, sum = x;
        for (int n = 1; n < 40; n++) {
            term = term * m2x2 / (2 * n + 1);
            sum += term;
            if (fabs(term) < 1e-17 * fabs(sum)) break;
        }
        return 2 * kIsPi * sum;
    }
    double wr;
    if (std::isinf(ax)) v = 0.0;
    else if (ax < 15.0) fad_trap_centred(ax, 0.0, wr, v);
    else fad_cf(ax, 0.0, fad_cf_depth(ax * ax), wr, v);
    return x < 0 ? -v : v;
}

__host__ __device__ inline void faddeeva_w(double re, double im, double &wr, double &wi)
{
    if (std::isnan(re) || std::isnan(im)) {
        if (re == 0.0) { wr = im; wi = 0.0; return; }  // w(i NaN)
        wr = NAN; wi = NAN; return;
    }
    if (re == 0.0) { wr = cts_erfcx(im); wi = re; return; }
    if (im == 0.0) { wr = exp(-(re * re)); wi = cts_w_im(re); return; }
    const double ax = fabs(re);
    if (im > 0.0) {
        fad_q1(ax, im, wr, wi);
        if (re < 0.0) wi = -wi;
        return;
    }
    // below the real axis: w(z) = 2 exp(-z^2) - w(-z), -z = -re + i |im|
    double vr, vi;
    fad_q1(ax, -im, vr, vi);
    if (re > 0.0) vi = -vi;  // w(-re + i|im|) = conj w(re + i|im|) for re > 0
    const double ex = (-im - ax) * (ax - im);  // im^2 - re^2 without overflow
    double er = 0.0, ei = 0.0;
    if (std::isnan(ex)) { wr = NAN; wi = NAN; return; }
    if (ex > -745.0) {
        const double e = 2.0 * exp(ex), p = -2.0 * re * im;
        er = e * cos(p);
        ei = e * sin(p);
    }
    wr = er - vr;
    wi = ei - vi;
}

// w(z) for Re z >= 0, Im z >= 0 (any finite values, axes included).
__host__ __device__ inline void faddeeva_w_q1(double re, double im, double &wr, double &wi)
{
    if (re == 0.0) { wr = cts_erfcx(im); wi = re; return; }
    if (im == 0.0) { wr = exp(-(re * re)); wi = cts_w_im(re); return; }
    fad_q1(re, im, wr, wi);
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
// form by Weideman's approximation in one interleaved evaluation:
// w(z) = 2 p(Z)/(L - iz)^2 + pi^{-1/2}/(L - iz), Z = (L + iz)/(L - iz), p of
// degree N - 1 = 39, L = (N/sqrt 2)^{1/2}.  The kicks take it on all of
// Re z < 100, 0 <= Im z < 1000, where its |w| relative error stays below
// 3e-14 (largest next to the real axis for 6 < Re z < 10), so that the lanes
// of a wave do not split between methods; points beyond take
// faddeeva_w_q1.  Same formulas as ion_field otherwise, rounding within a
// few ulps.
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

// Weideman's approximation at NP points, their Horner recurrences
// interleaved.
template <int NP>
__device__ __forceinline__ void w_weideman(const double *zr, const double *zi, double *wr, double *wi)
{
#pragma clang fp contract(fast)
    // the coefficients through an opaque pointer: their scalar loads stay in
    // the step loop instead of being hoisted out of the caller's loops and
    // kept (and spilled) in scalar registers
    const cdptr cw = reload((uint64_t)(uintptr_t)c_weid);
    const double L = cw[kWeidN];
    double ur[NP], ui[NP], Zr[NP], Zi[NP], r[NP], s[NP], b1[NP], b2[NP];
#pragma unroll
    for (int q = 0; q < NP; q++) {
        const double dr = L + zi[q], di = -zr[q];
        // 1/|L - i z|^2 from the hardware reciprocal and two Newton steps:
        // the denominator is at least L^2 (Im z >= 0) and finite in the
        // kick region, so the IEEE division's scaling and fix-up steps are
        // not needed (within an ulp of the quotient)
        const double den = dr * dr + di * di;
        double inv = __builtin_amdgcn_rcp(den);
        inv = fma(inv, fma(-den, inv, 1.0), inv);
        inv = fma(inv, fma(-den, inv, 1.0), inv);
        ur[q] = dr * inv;
        ui[q] = -di * inv;
        const double nr = L - zi[q], ni = zr[q];
        Zr[q] = nr * ur[q] - ni * ui[q];
        Zi[q] = nr * ui[q] + ni * ur[q];
        r[q] = 2.0 * Zr[q];
        s[q] = Zr[q] * Zr[q] + Zi[q] * Zi[q];
        b2[q] = cw[0];
        b1[q] = fma_vvs(r[q], b2[q], cw[1]);
    }
    // p(Z) for p's real coefficients through the remainder of p modulo the
    // real quadratic (x - Z)(x - conj Z) = x^2 - r x + s, r = 2 Re Z,
    // s = |Z|^2 (D. E. Knuth, The Art of Computer Programming vol. 2,
    // 4.6.4): b_k = c_k + r b_{k+1} - s b_{k+2} from the highest degree down,
    // then p(Z) = Z b_1 + (c_0 - s b_2).  Two fma per coefficient (the
    // coefficient as a scalar operand) instead of a complex multiply-add;
    // Z never reaches the unit circle (|Z| < 1 for Im z >= 0), where the
    // recurrence's double root at a real Z would amplify rounding, and over
    // the kick region the result stays within 1.5e-15 of |w| of the complex
    // Horner scheme's
#pragma unroll 8
    for (int k = 2; k < kWeidN - 1; k++) {
        const double ck = cw[k];
#pragma unroll
        for (int q = 0; q < NP; q++) {
            const double t = fma(r[q], b1[q], fnma_vvs(s[q], b2[q], ck));
            b2[q] = b1[q];
            b1[q] = t;
        }
    }
    const double c0 = cw[kWeidN - 1];
#pragma unroll
    for (int q = 0; q < NP; q++) {
        const double pr = Zr[q] * b1[q] + fnma_vvs(s[q], b2[q], c0), pi = Zi[q] * b1[q];
        const double u2r = ur[q] * ur[q] - ui[q] * ui[q], u2i = 2.0 * ur[q] * ui[q];
        wr[q] = 2.0 * (pr * u2r - pi * u2i) + kIsPi * ur[q];
        wi[q] = 2.0 * (pr * u2i + pi * u2r) + kIsPi * ui[q];
    }
}

// w(z) outside the kick region (x >= 100 or y >= 1000, both >= 0), as
// faddeeva_w_q1 evaluates it there: erfcx on the imaginary axis, otherwise
// the continued fraction (at most 4 levels for |z|^2 >= 1e4) or the
// asymptotic form; exp(-z^2) has underflowed.  Few registers, so the kick
// loops keep their occupancy.
__device__ __forceinline__ void w_far(double x, double y, double &wr, double &wi)
{
    if (x == 0.0) { wr = cts_erfcx(y); wi = 0.0; return; }
    if (std::isinf(x) || std::isinf(y)) { wr = 0.0; wi = 0.0; return; }
    if (x >= 1e7 || y >= 1e7) { fad_asymptotic(x, y, wr, wi); return; }
    fad_cf(x, y, fad_cf_depth(x * x + y * y), wr, wi);
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
    // Weideman's approximation for both points; a point outside the kick
    // region (rare: far from the beam) takes the far-field form instead
    // one point after the other: four waves per SIMD interleave enough
    // work, and the single recurrence needs fewer registers (fewer spills
    // in the ion turn's passage loop) than the two side by side
#pragma nounroll
    for (int q = 0; q < 2; q++) w_weideman<1>(zr + q, zi + q, wr + q, wi + q);
#pragma unroll
    for (int q = 0; q < 2; q++)
        if (!w_kick_region(zr[q], zi[q])) w_far(zr[q], zi[q], wr[q], wi[q]);
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
