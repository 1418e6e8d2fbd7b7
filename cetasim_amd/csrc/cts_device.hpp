// cts_device.hpp -- device helpers shared by the tracking kernels.
//
// All HIP sources of this library are compiled with -ffp-contract=off:
// multiply and add stay separate operations in the order written, so the
// FP64 arithmetic follows the reference's operation order
// (src/Bunch.cpp:625-755, 940-1044) and the float Box-Muller transform
// reproduces the host replica bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cts {

constexpr double kPi = 3.141592653589793;  // 4*atan(1), include/Global.h:33
constexpr double kClight = 299792458.0;
constexpr double kEcharge = 1.602017733e-19;
constexpr double kRe = 2.81794092E-15;
constexpr double kElectronMassEV = 0.51099906e6;  // include/Global.h:30
constexpr double kIonMassEV = 931.49432e+6;       // atomic mass unit, include/Global.h:32
constexpr double kBoltzmann = 1.3806505E-23;      // include/Global.h:36

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
struct u32x4 { uint32_t v[4]; };

__device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                               uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        // one 32x32 -> 64-bit product gives both halves
        const uint64_t m0 = (uint64_t)0xD2511F53u * c0, m1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(m0 >> 32), lo0 = (uint32_t)m0;
        const uint32_t hi1 = (uint32_t)(m1 >> 32), lo1 = (uint32_t)m1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    u32x4 o;
    o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

// ln(u), u in (0, 1]: u = m*2^e, m in [sqrt(1/2), sqrt(2)), 2*atanh series.
__device__ __forceinline__ float cts_logf(float u)
{
    const uint32_t b = __float_as_uint(u);
    int e = (int)((b >> 23) & 0xffu) - 127;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    const float p = 1.0f + s2 * (0.333333333f + s2 * (0.2f + s2 * (0.142857143f + s2 * 0.111111111f)));
    return (float)e * 0.693147181f + 2.0f * s * p;
}

// sin and cos of 2*pi*u2, u2 = k*2^-24 in [0, 1).
__device__ __forceinline__ void cts_sincos2pif(float u2, float &s_out, float &c_out)
{
    const float t = u2 * 4.0f;
    const int q = (int)t;
    float f = t - (float)q;
    const bool swap = f > 0.5f;
    if (swap) f = 1.0f - f;
    const float th = f * 1.57079633f;
    const float t2 = th * th;
    const float s = th * (1.0f - t2 * (0.166666667f - t2 * (0.00833333333f - t2 * 0.000198412698f)));
    const float c = 1.0f - t2 * (0.5f - t2 * (0.0416666667f - t2 * (0.00138888889f - t2 * 2.48015873e-5f)));
    const float sn = swap ? c : s, cs = swap ? s : c;
    switch (q & 3) {
    case 0: s_out = sn; c_out = cs; break;
    case 1: s_out = cs; c_out = -sn; break;
    case 2: s_out = -sn; c_out = -cs; break;
    default: s_out = -cs; c_out = sn; break;
    }
}

// Box-Muller pair from two 24-bit integers: u1 = (a+1)*2^-24 in (0, 1],
// u2 = b*2^-24 in [0, 1).
__device__ __forceinline__ void box_muller24(uint32_t a, uint32_t b, float &n0, float &n1)
{
    const float u1 = (float)(a + 1u) * 0x1p-24f;
    const float u2 = (float)b * 0x1p-24f;
    const float r = __builtin_sqrtf(-2.0f * cts_logf(u1));
    float s, c;
    cts_sincos2pif(u2, s, c);
    n0 = r * c;
    n1 = r * s;
}

__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float &n0, float &n1)
{
    box_muller24(a >> 8, b >> 8, n0, n1);
}

// --------------------------------------------- table-driven normal pairs
// The quantum-excitation normals use a Box-Muller transform whose ln and
// sincos come from two 256-entry tables (4 KB, staged in LDS by the
// kernels that draw them) and short polynomials written with explicit
// fmaf, so the oracle's C replica produces the same bits:
//   sc[j] = {sin(2 pi j/256), cos(2 pi j/256)},
//   lg[j] = {ln(1 + j/256), 256/(256 + j)}   (float roundings of the
//           double values; the host builds them with build_normal_tables).
constexpr int kNormTab = 256;
struct NormalTab {
    float2 sc[kNormTab];
    float2 lg[kNormTab];
};

// -2 ln(u1) for u1 = (a + 1) 2^-21, a < 2^21: k = a + 1 = 2^e m with m in
// [1, 2); ln m = lg[j].x + ln(1 + t), t = (m - m_j) / m_j, |t| < 2^-8, to
// third order.  Next to u1 = 1 (k > 2^21 - 2^13), where that sum would
// cancel, -ln(1 - y) with y = 1 - u1 exact, to fourth order.
__device__ __forceinline__ float neg2ln_u21(const float2 *lg, uint32_t a)
{
    const uint32_t k = a + 1u;
    const uint32_t bits = __float_as_uint((float)k);  // exact: k <= 2^21
    const int e = (int)(bits >> 23) - 127;
    const float2 L = lg[(bits >> 15) & 0xffu];
    const float t = (float)(bits & 0x7fffu) * 0x1p-23f * L.y;
    const float q = __builtin_fmaf(__builtin_fmaf(t, 0x1.555556p-2f, -0.5f), t, 1.0f);
    float lnu = __builtin_fmaf((float)(e - 21), 0x1.62e43p-1f, L.x + t * q);
    if (k > 0x1fe000u) {
        const float y = (float)(0x200000u - k) * 0x1p-21f;
        const float r = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(y, 0.25f, 0x1.555556p-2f), y, 0.5f), y, 1.0f);
        lnu = -(y * r);
    }
    return -2.0f * lnu;
}

// sin and cos of 2 pi b 2^-21, b < 2^21: table angle 2 pi (b >> 13)/256
// plus delta < 2 pi/256 (sin to delta^3, cos to delta^4).
__device__ __forceinline__ void sincos2pi_u21(const float2 *sc, uint32_t b, float &s, float &c)
{
    const float d = (float)(b & 0x1fffu) * 0x1.921fb6p-19f;  // 2 pi 2^-21
    const float d2 = d * d;
    const float sd = d * __builtin_fmaf(d2, -0x1.555556p-3f, 1.0f);
    const float cd = __builtin_fmaf(d2, __builtin_fmaf(d2, 0x1.555556p-5f, -0.5f), 1.0f);
    const float2 T = sc[b >> 13];
    s = __builtin_fmaf(T.x, cd, T.y * sd);
    c = __builtin_fmaf(T.y, cd, -(T.x * sd));
}

__device__ __forceinline__ void box_muller_u21(const NormalTab &T, uint32_t a, uint32_t b, float &n0, float &n1)
{
    const float r = __builtin_sqrtf(neg2ln_u21(T.lg, a));
    float s, c;
    sincos2pi_u21(T.sc, b, s, c);
    n0 = r * c;
    n1 = r * s;
}

// The five quantum-excitation normals (r0 r1 r2 r3 r5) of one particle:
// one Philox4x32-10 block {index, turn, 0} cut into six 21-bit uniforms
// (bits 0-20, 21-41, 42-62 of the two 64-bit halves), three Box-Muller
// pairs, the sixth normal unused.  21-bit radii reach 5.4 sigma.
__device__ __forceinline__ void qe_normals(const NormalTab &T, uint64_t seed, uint32_t turn, uint64_t index,
                                           double out[5])
{
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t i0 = (uint32_t)index, i1 = (uint32_t)(index >> 32);
    const u32x4 a = philox4x32_10(i0, i1, turn, 0u, k0, k1);
    const uint64_t lo = ((uint64_t)a.v[1] << 32) | a.v[0], hi = ((uint64_t)a.v[3] << 32) | a.v[2];
    const uint32_t m21 = 0x1fffffu;
    float n0, n1, n2, n3, n4, n5;
    box_muller_u21(T, (uint32_t)lo & m21, (uint32_t)(lo >> 21) & m21, n0, n1);
    box_muller_u21(T, (uint32_t)(lo >> 42) & m21, (uint32_t)hi & m21, n2, n3);
    box_muller_u21(T, (uint32_t)(hi >> 21) & m21, (uint32_t)(hi >> 42) & m21, n4, n5);
    out[0] = n0; out[1] = n1; out[2] = n2; out[3] = n3; out[4] = n4;
}

// Stage the normal tables from global memory into LDS (all threads of the
// block take part; the caller synchronizes).
__device__ __forceinline__ void load_normal_tab(NormalTab &T, const float2 *g, int tid, int nthreads)
{
    for (int j = tid; j < 2 * kNormTab; j += nthreads) {
        if (j < kNormTab) T.sc[j] = g[j];
        else T.lg[j - kNormTab] = g[j];
    }
}

// ------------------------------------------------------- section matrices
// B1H1, InvH1InvB1 and InvH2InvB2 (src/LatticeInterActionPoint.cpp:326-431)
// share one structural pattern: row 0 {0,5}, row 1 {0,1,5}, row 2 {2,5},
// row 3 {2,3,5}, row 4 {0,1,2,3,4}, row 5 {5}.  The 16 entries are stored in
// that order; the product adds them in ascending column order as gslcblas
// dgemm does (entries that are exactly zero only change the sign of a zero).
struct SparseMat { double m[16]; };

__device__ __forceinline__ void spmv(const SparseMat &A, const double x[6], double y[6])
{
    y[0] = A.m[0] * x[0] + A.m[1] * x[5];
    y[1] = A.m[2] * x[0] + A.m[3] * x[1] + A.m[4] * x[5];
    y[2] = A.m[5] * x[2] + A.m[6] * x[5];
    y[3] = A.m[7] * x[2] + A.m[8] * x[3] + A.m[9] * x[5];
    y[4] = A.m[10] * x[0] + A.m[11] * x[1] + A.m[12] * x[2] + A.m[13] * x[3] + A.m[14] * x[4];
    y[5] = A.m[15] * x[5];
}

// Constant-address-space view of a device buffer.  reload() hides the
// address from the optimizer, so the loads it feeds stay inside the loop as
// scalar loads instead of being hoisted and then spilled.
typedef const __attribute__((address_space(4))) double *cdptr;

__device__ __forceinline__ cdptr reload(uint64_t a)
{
    asm volatile("" : "+s"(a));
    return (cdptr)a;
}

// a * b + c with c a scalar-register operand, as one v_fma_f64 (the same
// fused rounding as fma()).  Left to itself the compiler forms a Horner
// step with a constant addend as two moves of c into the accumulator and a
// v_fmac_f64: three vector instructions for one.
__device__ __forceinline__ double fma_vvs(double a, double b, double c)
{
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

// -(a * b) + c, likewise (the negation as the instruction's modifier)
__device__ __forceinline__ double fnma_vvs(double a, double b, double c)
{
    double r;
    asm("v_fma_f64 %0, -%1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

__device__ __forceinline__ void spmv_k(cdptr A, const double x[6], double y[6])
{
#pragma clang fp contract(fast)
    y[0] = A[0] * x[0] + A[1] * x[5];
    y[1] = A[2] * x[0] + A[3] * x[1] + A[4] * x[5];
    y[2] = A[5] * x[2] + A[6] * x[5];
    y[3] = A[7] * x[2] + A[8] * x[3] + A[9] * x[5];
    y[4] = A[10] * x[0] + A[11] * x[1] + A[12] * x[2] + A[13] * x[3] + A[14] * x[4];
    y[5] = A[15] * x[5];
}


// ---------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Sum over the 64 lanes by DPP moves (no LDS): pairs within quads, quads
// within rows (half mirror, mirror), then rows through the row_bcast15 and
// row_bcast31 broadcasts, which leave the total in lane 63; returned to
// every lane.  The association order differs from wave_sum's.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_dpp(double v)
{
    v += dpp_f64<0xb1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4e, 0xf>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141, 0xf>(v);  // row_half_mirror
    v += dpp_f64<0x140, 0xf>(v);  // row_mirror
    v += dpp_f64<0x142, 0xa>(v);  // row_bcast15 into rows 1, 3
    v += dpp_f64<0x143, 0xc>(v);  // row_bcast31 into rows 2, 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}
// Four wave sums at once, each step applied to the four values before the
// next, so the DPP moves of one value fill the wait states of another; the
// sums are exact in lane 63 only (rows the broadcast steps skip keep
// unspecified values: the moves carry no old value).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64_any(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ void dpp_step4(double v[4])
{
    double t[4];
#pragma unroll
    for (int j = 0; j < 4; j++) t[j] = dpp_f64_any<CTRL, ROWS>(v[j]);
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] += t[j];
}
__device__ __forceinline__ void wave_sum4_dpp(double v[4])
{
    dpp_step4<0xb1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
    dpp_step4<0x4e, 0xf>(v);   // quad_perm [2, 3, 0, 1]
    dpp_step4<0x141, 0xf>(v);  // row_half_mirror
    dpp_step4<0x140, 0xf>(v);  // row_mirror
    dpp_step4<0x142, 0xa>(v);  // row_bcast15 into rows 1, 3
    dpp_step4<0x143, 0xc>(v);  // row_bcast31 into rows 2, 3
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int lo = __builtin_amdgcn_readlane(__double2loint(v[j]), 63);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(v[j]), 63);
        v[j] = __hiloint2double(hi, lo);
    }
}
__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace cts
