// cts_mlp.hpp -- the learned field surrogate: a 3 -> 64 -> 64 -> 2 MLP with
// tanh-form GELU activations for the normalized field of a Gaussian beam.
//
// Inputs (features) of one field point, after ordering the beam so that
// sigma_x >= sigma_y:  u = x / sigma_x, v = y / sigma_y, r = sigma_y / sigma_x,
//   a0 = |u| / 3 - 1,  a1 = |v| / 3 - 1,  a2 = (ln r - ln(0.02) / 2) / (-ln(0.02) / 2),
// each in [-1, 1] on the trained domain |u|, |v| <= 6, r in [0.02, 1].
// Outputs o0, o1 encode the field G = F * sigma_x of a unit line charge
// (F as the Bassetti-Erskine kick of src/Bunch.cpp:119-211):
//   G_x = u * b_x * exp(mu_x + s_x * o0),   G_y = -v * exp(mu_y + s_y * o1),
//   b_x = -(1 - e^{-rho^2 / (1 + r)}) / rho^2,   rho^2 = u^2 + (v r)^2,
// b_x is exact for the horizontal field in the near field (-1/(1+r), the
// linear field of the beam) and far out (the line charge, 1/rho); the
// vertical field of a flat beam falls off with u on the scale of sigma_x,
// not of rho, so it is learned without an envelope.  The form is odd in u
// and v as the field is, and the network's absolute error is the field's
// relative error.
//
// Parameter vector (FP32, 4546 values): W1[64][3], b1[64], W2[64][64]
// (row = output), b2[64], W3[2][64], b3[2].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cts {

constexpr int kMlpH = 64;
constexpr int kMlpParams = 4546;
constexpr int MP_W1 = 0, MP_B1 = 192, MP_W2 = 256, MP_B2 = 4352, MP_W3 = 4416, MP_B3 = 4544;

// domain of the surrogate (outside: exact Bassetti-Erskine)
constexpr double kSurUMax = 6.0;
constexpr double kSurRMin = 0.02;
constexpr double kSurZero = 1.0e-3;  // |u| + |v| below this: exact path
constexpr float kSurLnRMin = -3.91202300542814606f;  // ln(0.02)

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// The network's arithmetic is contracted to FMAs (the rest of the library
// keeps multiply and add separate).
//
// GELU, tanh form: 0.5 z (1 + tanh(a)), a = sqrt(2/pi) (z + 0.044715 z^3),
// evaluated as z * sigma(2a) = z / (1 + 2^(-2a log2 e)) with the hardware
// exp2 and reciprocal; dg is its derivative, r + 2 z r (1 - r) a'(z).
constexpr float kGeluK0 = 0.7978845608028654f, kGeluK1 = 0.044715f;
constexpr float kGeluC = -2.0f * 0.7978845608028654f * 1.4426950408889634f;  // -2 k0 log2(e)
__device__ __forceinline__ float gelu_sig(float z)
{
#pragma clang fp contract(fast)
    const float z2 = z * z;
    const float q = kGeluC + (kGeluC * kGeluK1) * z2;
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(q * z));
}
__device__ __forceinline__ float gelu_f(float z, float &dg)
{
#pragma clang fp contract(fast)
    const float r = gelu_sig(z);
    dg = r + 2.0f * kGeluK0 * z * r * (1.0f - r) * (1.0f + 3.0f * kGeluK1 * z * z);
    return z * r;
}
__device__ __forceinline__ float gelu_only(float z) { return z * gelu_sig(z); }

// x = hi + lo with hi, lo bfloat16 (round to nearest even): the product of
// two such pairs by three bf16 MFMAs (hi*hi + hi*lo + lo*hi) carries ~16
// significant bits.
__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo)
{
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// Feature encoding of one point (u, v already taken as |u|, |v|).
__device__ __forceinline__ void sur_features(double au, double av, double r, float a[3])
{
    a[0] = (float)(au * (1.0 / 3.0) - 1.0);
    a[1] = (float)(av * (1.0 / 3.0) - 1.0);
    a[2] = ((float)log(r) - 0.5f * kSurLnRMin) / (-0.5f * kSurLnRMin);
}

// Envelope factors b_x, b_y (see the header comment); FP32.
__device__ __forceinline__ void sur_envelope(float au, float av, float r, float &bx, float &by)
{
    const float vr = av * r;
    const float rho2 = au * au + vr * vr;
    const float qx = rho2 / (1.0f + r);
    // (1 - e^{-q}) / q
    const float fx = qx < 1e-4f ? 1.0f - 0.5f * qx : -expm1f(-qx) / qx;
    bx = -fx / (1.0f + r);
    by = -1.0f;
}

constexpr int kPb = 72;  // LDS row pitch of bf16 tiles (144 B): fragment reads spread over banks
constexpr int kPf = 65;  // LDS row pitch of FP32 rows

__device__ __forceinline__ v4f mfma3(v8bf ah, v8bf al, v8bf bh, v8bf bl, v4f c)
{
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
    return c;
}

__device__ __forceinline__ v8bf frag(const __bf16 *base)
{
    return *reinterpret_cast<const v8bf *>(base);
}

// ------------------------------------------------------------- inference
// Network outputs for the 64 points of a wave (lane = point), computed as
// Z2^T = W2 H1^T so that the first layer is produced directly in the B
// operand layout and the second layer's result lands where the last layer
// needs it; no LDS besides the weights.  wl: LDS weights, see
// load_mlp_lds.  Returns o0, o1 of the lane's own point.
struct MlpLds {
    __bf16 w2h[64 * kPb], w2l[64 * kPb];  // W2[n][k]
    float w1[64 * 4];                     // W1[k][0..2], b1[k]
    float w3[64 * 4];                     // b2[n], W3[0][n], W3[1][n], -
    float b3[2];
};

__device__ __forceinline__ void load_mlp_lds(MlpLds &L, const float *W, int tid, int nthreads)
{
    for (int i = tid; i < 4096; i += nthreads) {
        const int n = i >> 6, k = i & 63;
        __bf16 hi, lo;
        split_bf16(W[MP_W2 + i], hi, lo);
        L.w2h[n * kPb + k] = hi;
        L.w2l[n * kPb + k] = lo;
    }
    for (int k = tid; k < 64; k += nthreads) {
        L.w1[4 * k] = W[MP_W1 + 3 * k];
        L.w1[4 * k + 1] = W[MP_W1 + 3 * k + 1];
        L.w1[4 * k + 2] = W[MP_W1 + 3 * k + 2];
        L.w1[4 * k + 3] = W[MP_B1 + k];
        L.w3[4 * k] = W[MP_B2 + k];
        L.w3[4 * k + 1] = W[MP_W3 + k];
        L.w3[4 * k + 2] = W[MP_W3 + 64 + k];
        L.w3[4 * k + 3] = 0.0f;
    }
    if (tid == 0) { L.b3[0] = W[MP_B3]; L.b3[1] = W[MP_B3 + 1]; }
}

__device__ __forceinline__ void mlp_forward_wave(const MlpLds &L, float a0, float a1, float a2, float &out0,
                                                 float &out1)
{
#pragma clang fp contract(fast)
    const int l = threadIdx.x & 63;
    const int lr = l & 15, lg = (l >> 4) & 3;
    float q0 = 0.0f, q1 = 0.0f;
#pragma unroll 1
    for (int nb = 0; nb < 4; nb++) {
        // first layer for samples s = 16 nb + lr in the B layout: features
        // k = 32 kb + 8 lg + i
        const float x0 = __shfl(a0, 16 * nb + lr, 64);
        const float x1 = __shfl(a1, 16 * nb + lr, 64);
        const float x2 = __shfl(a2, 16 * nb + lr, 64);
        v8bf bh[2], bl[2];
        int opw = 0;
        asm volatile("" : "+v"(opw));
#pragma unroll
        for (int kb = 0; kb < 2; kb++)
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int k = 32 * kb + 8 * lg + i;
                const float4 w = *reinterpret_cast<const float4 *>(&L.w1[4 * k + opw]);
                const float z = w.w + w.x * x0 + w.y * x1 + w.z * x2;
                __bf16 hi, lo;
                split_bf16(gelu_only(z), hi, lo);
                bh[kb][i] = hi;
                bl[kb][i] = lo;
            }
        float s0 = 0.0f, s1 = 0.0f;
        // re-read the weight fragments from LDS in every block instead of
        // keeping them in registers across the loop (occupancy)
        int opq = 0;
        asm volatile("" : "+v"(opq));
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            v4f acc = {0, 0, 0, 0};
#pragma unroll
            for (int kb = 0; kb < 2; kb++) {
                const int ao = (16 * mb + lr) * kPb + 32 * kb + 8 * lg + opq;
                acc = mfma3(frag(L.w2h + ao), frag(L.w2l + ao), bh[kb], bl[kb], acc);
            }
            // rows n = 16 mb + 4 lg + t, column s = 16 nb + lr
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const float4 e = *reinterpret_cast<const float4 *>(&L.w3[4 * (16 * mb + 4 * lg + t) + opq]);
                const float h = gelu_only(acc[t] + e.x);
                s0 += e.y * h;
                s1 += e.z * h;
            }
        }
        // sum over the four lane groups that share the column
        s0 += __shfl_xor(s0, 16, 64);
        s0 += __shfl_xor(s0, 32, 64);
        s1 += __shfl_xor(s1, 16, 64);
        s1 += __shfl_xor(s1, 32, 64);
        // lane l = 16 lg + lr owns sample s = l, i.e. the block nb = lg
        if (nb == lg) {
            q0 = s0;
            q1 = s1;
        }
    }
    out0 = q0 + L.b3[0];
    out1 = q1 + L.b3[1];
}

// Surrogate output normalisation (target mean and scale of o0, o1).
struct SurNorm { float mx, sx, my, sy; };

// Field (F_x, F_y) of a Gaussian beam of rms sizes (sx, sy) at (x, y) per
// unit line charge, from the surrogate, for the lanes whose point lies in
// its domain; returns false for the other lanes (and for inactive lanes),
// which take the exact path.  Every lane of the wave must call it: the
// network runs on the whole wave's points at once.
__device__ __forceinline__ bool sur_field(const MlpLds &L, const SurNorm &nm, bool active, double x, double y,
                                          double sx, double sy, double &fx, double &fy)
{
    const bool swap = sx < sy;
    const double X = swap ? y : x, Y = swap ? x : y, SX = swap ? sy : sx, SY = swap ? sx : sy;
    const double u = X / SX, v = Y / SY, r = SY / SX;
    const double au = fabs(u), av = fabs(v);
    const bool dom = active && au <= kSurUMax && av <= kSurUMax && r >= kSurRMin && r <= 1.0 &&
                     au + av >= kSurZero;
    // a wave with no lane inside the trained domain skips the network (the
    // callers run this in wave-uniform control flow, as the MFMA needs)
    if (!__any(dom)) return false;
    float a[3];
    sur_features(dom ? au : 1.0, dom ? av : 1.0, dom ? r : 0.5, a);
    float o0, o1;
    mlp_forward_wave(L, a[0], a[1], a[2], o0, o1);
    if (!dom) return false;
    float bx, by;
    sur_envelope((float)au, (float)av, (float)r, bx, by);
    const double gx = (double)((float)au * bx * __expf(nm.mx + nm.sx * o0));
    const double gy = (double)((float)av * by * __expf(nm.my + nm.sy * o1));
    const double FX = (u < 0 ? -gx : gx) / SX, FY = (v < 0 ? -gy : gy) / SX;
    fx = swap ? FY : FX;
    fy = swap ? FX : FY;
    return true;
}

}  // namespace cts
