// cts_mlp.hpp -- the learned field surrogate: a 3 -> 64 -> 64 -> 2 MLP with
// tanh-form GELU activations for the normalized field of a Gaussian beam.
//
// Inputs (features) of one field point, after ordering the beam so that
// sigma_x >= sigma_y:  u = x / sigma_x, v = y / sigma_y, r = sigma_y / sigma_x,
//   a0 = |u| / 3 - 1,  a1 = |v| / 3 - 1,  a2 = (ln r - ln(0.02) / 2) / (-ln(0.02) / 2),
// each in [-1, 1] on the trained domain |u|, |v| <= 6, r in [0.02, 1].
// Outputs o0, o1 encode the field G = F * sigma_x of a unit line charge
// (F as the Bassetti-Erskine kick of src/Bunch.cpp:119-211) through
//   G_x = -u * exp(mu_x + s_x * o0),   G_y = -v * exp(mu_y + s_y * o1),
// which is odd in u and v as the field is, and makes the network's
// absolute error the field's relative error.
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

// GELU, tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))), with
// tanh(a) = 1 - 2 / (1 + e^{2a}); dg is its derivative.
__device__ __forceinline__ float gelu_f(float z, float &dg)
{
    const float k0 = 0.7978845608028654f, k1 = 0.044715f;
    const float z2 = z * z;
    const float a = k0 * z * (1.0f + k1 * z2);
    const float e = __expf(2.0f * a);
    const float t = 1.0f - 2.0f / (1.0f + e);
    dg = 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * k0 * (1.0f + 3.0f * k1 * z2);
    return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_only(float z)
{
    const float k0 = 0.7978845608028654f, k1 = 0.044715f;
    const float a = k0 * z * (1.0f + k1 * z * z);
    const float t = 1.0f - 2.0f / (1.0f + __expf(2.0f * a));
    return 0.5f * z * (1.0f + t);
}

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

}  // namespace cts
