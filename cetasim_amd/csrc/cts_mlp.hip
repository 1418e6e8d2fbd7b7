// cts_mlp.hip -- training and inference of the field surrogate (gfx950).
//
// The two 64x64 products of every forward and backward pass run on the
// bf16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulate) with each
// operand split into a bf16 high and low part (three products per tile), so
// the products carry ~16 significant bits; the 3-input first layer, the
// 2-output last layer and the activations are FP32 vector code.
//
// MFMA 16x16x32 operand layout (one wave, lane l):
//   A (16 x 32, row-major):  row l % 16, columns 8 (l / 16) .. +7
//   B (32 x 16):             column l % 16, rows 8 (l / 16) .. +7
//   D (16 x 16):             rows 4 (l / 16) + t, column l % 16, t = 0..3
#include "cts_internal.hpp"
#include "cts_mlp.hpp"

namespace cts {

// ---------------------------------------------------------------- samples
// Training / validation samples: |u|, |v| uniform in (0, 6], r log-uniform
// in [0.02, 1] from Philox4x32-10 (key = seed, counter = {i, 0, 0, tag});
// the field is the FP64 Bassetti-Erskine path (sigma_x = 1, sigma_y = r).
// Each record: a0 a1 a2 t0 t1 with t = log(G / (u b)) (see cts_mlp.hpp).
__global__ void __launch_bounds__(256)
k_mlp_samples(int64_t n, uint64_t seed, uint32_t tag, float *out, double *exact)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u32x4 o = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 0u, tag, (uint32_t)seed,
                                  (uint32_t)(seed >> 32));
    // |u|, |v| >= kSurZero / 2 keeps every sample off the zero-field branch
    const double lo = 0.5 * kSurZero;
    const double au = lo + (kSurUMax - lo) * (((double)(o.v[0] >> 8) + 0.5) * 0x1p-24);
    const double av = lo + (kSurUMax - lo) * (((double)(o.v[1] >> 8) + 0.5) * 0x1p-24);
    const double r = exp(log(kSurRMin) * (1.0 - ((double)(o.v[2] >> 8) + 0.5) * 0x1p-24));
    double gx, gy;
    ion_field(au, av * r, 1.0, r, gx, gy);
    float a[3];
    sur_features(au, av, r, a);
    float bx, by;
    sur_envelope((float)au, (float)av, (float)r, bx, by);
    float *rec = out + i * 5;
    rec[0] = a[0];
    rec[1] = a[1];
    rec[2] = a[2];
    rec[3] = (float)log(gx / (au * (double)bx));
    rec[4] = (float)log(gy / (av * (double)by));
    if (exact) {
        exact[i * 5 + 0] = au;
        exact[i * 5 + 1] = av;
        exact[i * 5 + 2] = r;
        exact[i * 5 + 3] = gx;
        exact[i * 5 + 4] = gy;
    }
}

// ------------------------------------------------------------- training
// One workgroup = one wave; it runs `chunks` batches of 64 samples and
// accumulates the gradient of the mean squared error of the normalized
// targets, then writes its partial gradient (+ loss) to part[wg].
struct MlpTrainParams {
    const float *params;    // kMlpParams
    const float *samples;   // [n][5]
    int64_t n;
    float *part;            // [grid][kMlpParams + 1]
    float norm[4];          // mu_x, s_x, mu_y, s_y of the targets
    float inv_batch;
    uint64_t seed;
    uint32_t step;
    int32_t chunks;
    int32_t random;         // 0: samples taken in order (gradient checks)
};

__global__ void __launch_bounds__(64)
k_mlp_train(MlpTrainParams P)
{
#pragma clang fp contract(fast)
    __shared__ __attribute__((aligned(16))) __bf16 w2h[64 * kPb], w2l[64 * kPb];    // W2[n][k]
    __shared__ __attribute__((aligned(16))) __bf16 w2th[64 * kPb], w2tl[64 * kPb];  // W2^T[k][n]
    __shared__ __attribute__((aligned(16))) __bf16 h1h[64 * kPb], h1l[64 * kPb];    // H1[s][k]
    __shared__ __attribute__((aligned(16))) __bf16 h1th[64 * kPb], h1tl[64 * kPb];  // H1^T[k][s]
    __shared__ __attribute__((aligned(16))) __bf16 d2h[64 * kPb], d2l[64 * kPb];    // dZ2[s][n]
    __shared__ __attribute__((aligned(16))) __bf16 d2th[64 * kPb], d2tl[64 * kPb];  // dZ2^T[n][s]
    __shared__ float z1s[64 * kPf];  // z1, then dz1   [s][k]
    __shared__ float z2s[64 * kPf];  // z2, then h2    [s][n]
    __shared__ float xs[64 * 4];
    __shared__ float dos[64 * 2];
    const int l = threadIdx.x;
    const int lr = l & 15, lg = l >> 4;
    const float *W = P.params;

    for (int i = l; i < 4096; i += 64) {
        const int n = i >> 6, k = i & 63;
        __bf16 hi, lo;
        split_bf16(W[MP_W2 + i], hi, lo);
        w2h[n * kPb + k] = hi; w2l[n * kPb + k] = lo;
        w2th[k * kPb + n] = hi; w2tl[k * kPb + n] = lo;
    }
    // per-feature accumulators (lane = feature j) and the dW2 tiles
    float gw1[3] = {0, 0, 0}, gb1 = 0, gb2 = 0, gw3[2] = {0, 0}, gb3[2] = {0, 0}, loss = 0;
    v4f gw2[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) gw2[a][b] = (v4f){0, 0, 0, 0};
    const float w3_0l = W[MP_W3 + l], w3_1l = W[MP_W3 + 64 + l];
    (void)w3_0l; (void)w3_1l;
    __syncthreads();

    for (int c = 0; c < P.chunks; c++) {
        const int64_t gid = ((int64_t)blockIdx.x * P.chunks + c) * 64 + l;
        int64_t idx;
        if (P.random) {
            const u32x4 o = philox4x32_10((uint32_t)gid, (uint32_t)(gid >> 32), P.step, 0x4d4c5054u,
                                          (uint32_t)P.seed, (uint32_t)(P.seed >> 32));
            idx = (int64_t)(((uint64_t)o.v[0] * (uint64_t)P.n) >> 32);
        } else {
            idx = gid < P.n ? gid : P.n - 1;
        }
        const float *rec = P.samples + idx * 5;
        const float a0 = rec[0], a1 = rec[1], a2 = rec[2];
        const float y0 = (rec[3] - P.norm[0]) / P.norm[1], y1 = (rec[4] - P.norm[2]) / P.norm[3];
        const bool valid = P.random || gid < P.n;
        xs[l * 4 + 0] = a0; xs[l * 4 + 1] = a1; xs[l * 4 + 2] = a2;
        // layer 1 (lane = sample)
        for (int k = 0; k < 64; k++) {
            const float z = W[MP_B1 + k] + W[MP_W1 + 3 * k] * a0 + W[MP_W1 + 3 * k + 1] * a1 + W[MP_W1 + 3 * k + 2] * a2;
            z1s[l * kPf + k] = z;
            __bf16 hi, lo;
            split_bf16(gelu_only(z), hi, lo);
            h1h[l * kPb + k] = hi; h1l[l * kPb + k] = lo;
            h1th[k * kPb + l] = hi; h1tl[k * kPb + l] = lo;
        }
        __syncthreads();
        // layer 2: Z2[s][n] = sum_k H1[s][k] W2[n][k]
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int nb = 0; nb < 4; nb++) {
                v4f acc = {0, 0, 0, 0};
#pragma unroll
                for (int kb = 0; kb < 2; kb++) {
                    const int ao = (16 * mb + lr) * kPb + 32 * kb + 8 * lg;
                    const int bo = (16 * nb + lr) * kPb + 32 * kb + 8 * lg;
                    acc = mfma3(frag(h1h + ao), frag(h1l + ao), frag(w2h + bo), frag(w2l + bo), acc);
                }
                const int n = 16 * nb + lr;
#pragma unroll
                for (int t = 0; t < 4; t++) z2s[(16 * mb + 4 * lg + t) * kPf + n] = acc[t] + W[MP_B2 + n];
            }
        __syncthreads();
        // output layer, loss and dZ2 (lane = sample)
        float o0 = W[MP_B3], o1 = W[MP_B3 + 1];
        for (int n = 0; n < 64; n++) {
            const float h = gelu_only(z2s[l * kPf + n]);
            o0 += W[MP_W3 + n] * h;
            o1 += W[MP_W3 + 64 + n] * h;
        }
        const float e0 = valid ? o0 - y0 : 0.0f, e1 = valid ? o1 - y1 : 0.0f;
        loss += e0 * e0 + e1 * e1;
        const float d0 = 2.0f * e0 * P.inv_batch, d1 = 2.0f * e1 * P.inv_batch;
        dos[l * 2] = d0; dos[l * 2 + 1] = d1;
        for (int n = 0; n < 64; n++) {
            float dg;
            const float h = gelu_f(z2s[l * kPf + n], dg);
            const float dz = (d0 * W[MP_W3 + n] + d1 * W[MP_W3 + 64 + n]) * dg;
            __bf16 hi, lo;
            split_bf16(dz, hi, lo);
            d2h[l * kPb + n] = hi; d2l[l * kPb + n] = lo;
            d2th[n * kPb + l] = hi; d2tl[n * kPb + l] = lo;
            z2s[l * kPf + n] = h;
        }
        __syncthreads();
        // per-feature sums over the samples (lane = feature j)
        for (int s = 0; s < 64; s++) {
            const float h = z2s[s * kPf + l];
            gw3[0] += dos[2 * s] * h;
            gw3[1] += dos[2 * s + 1] * h;
            gb2 += (float)d2h[s * kPb + l] + (float)d2l[s * kPb + l];
        }
        gb3[0] += d0;
        gb3[1] += d1;
        // dH1 = dZ2 W2 -> dZ1 = dH1 * gelu'(z1), in place of z1
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int nb = 0; nb < 4; nb++) {
                v4f acc = {0, 0, 0, 0};
#pragma unroll
                for (int kb = 0; kb < 2; kb++) {
                    const int ao = (16 * mb + lr) * kPb + 32 * kb + 8 * lg;
                    const int bo = (16 * nb + lr) * kPb + 32 * kb + 8 * lg;
                    acc = mfma3(frag(d2h + ao), frag(d2l + ao), frag(w2th + bo), frag(w2tl + bo), acc);
                }
                const int k = 16 * nb + lr;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int s = 16 * mb + 4 * lg + t;
                    float dg;
                    (void)gelu_f(z1s[s * kPf + k], dg);
                    z1s[s * kPf + k] = acc[t] * dg;
                }
            }
        // dW2 += dZ2^T H1: rows n, columns k, sum over the samples
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int nb = 0; nb < 4; nb++)
#pragma unroll
                for (int kb = 0; kb < 2; kb++) {
                    const int ao = (16 * mb + lr) * kPb + 32 * kb + 8 * lg;
                    const int bo = (16 * nb + lr) * kPb + 32 * kb + 8 * lg;
                    gw2[mb][nb] = mfma3(frag(d2th + ao), frag(d2tl + ao), frag(h1th + bo), frag(h1tl + bo), gw2[mb][nb]);
                }
        __syncthreads();
        for (int s = 0; s < 64; s++) {
            const float dz = z1s[s * kPf + l];
            gw1[0] += dz * xs[s * 4 + 0];
            gw1[1] += dz * xs[s * 4 + 1];
            gw1[2] += dz * xs[s * 4 + 2];
            gb1 += dz;
        }
        __syncthreads();
    }
    float *out = P.part + (int64_t)blockIdx.x * (kMlpParams + 1);
    out[MP_W1 + 3 * l] = gw1[0];
    out[MP_W1 + 3 * l + 1] = gw1[1];
    out[MP_W1 + 3 * l + 2] = gw1[2];
    out[MP_B1 + l] = gb1;
    out[MP_B2 + l] = gb2;
    out[MP_W3 + l] = gw3[0];
    out[MP_W3 + 64 + l] = gw3[1];
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int nb = 0; nb < 4; nb++)
#pragma unroll
            for (int t = 0; t < 4; t++) out[MP_W2 + (16 * mb + 4 * lg + t) * 64 + 16 * nb + lr] = gw2[mb][nb][t];
    // db3 and the loss: sums over the lanes (samples)
    float s0 = gb3[0], s1 = gb3[1], sl = loss;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, 64);
        s1 += __shfl_xor(s1, o, 64);
        sl += __shfl_xor(sl, o, 64);
    }
    if (l == 0) {
        out[MP_B3] = s0;
        out[MP_B3 + 1] = s1;
        out[kMlpParams] = sl;
    }
}

// Sum of the workgroup partials (fixed order) and one Adam step.
__global__ void __launch_bounds__(256)
k_mlp_adam(float *params, float *m, float *v, const float *part, int nwg, float lr, float b1, float b2,
           float eps, float bc1, float bc2, float *grad_out, float *loss_out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    float g = 0.0f;
    for (int w = 0; w < nwg; w++) g += part[(int64_t)w * (kMlpParams + 1) + p];
    if (p == kMlpParams) {
        if (loss_out) *loss_out = g;
        return;
    }
    if (grad_out) grad_out[p] = g;
    if (lr == 0.0f) return;
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

// Raw network outputs for n feature vectors (tests and validation).
__global__ void __launch_bounds__(256)
k_mlp_eval(const float *W, const float *feat, int64_t n, int32_t stride, float *out)
{
    __shared__ __attribute__((aligned(16))) MlpLds L;
    load_mlp_lds(L, W, threadIdx.x, blockDim.x);
    __syncthreads();
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        float a0 = 0, a1 = 0, a2 = 0;
        if (i < n) {
            a0 = feat[i * stride];
            a1 = feat[i * stride + 1];
            a2 = feat[i * stride + 2];
        }
        float o0, o1;
        mlp_forward_wave(L, a0, a1, a2, o0, o1);  // the whole wave takes part
        if (i < n) {
            out[2 * i] = o0;
            out[2 * i + 1] = o1;
        }
    }
}

// Field error of the surrogate against the exact FP64 values in
// exact[n][5] = |u| |v| r Gx Gy: per point |G_sur - G| / |G| (vector norm),
// written to err[n].
__global__ void __launch_bounds__(256)
k_mlp_field_error(const float *W, const double *exact, int64_t n, float4 norm, double *err)
{
    __shared__ __attribute__((aligned(16))) MlpLds L;
    load_mlp_lds(L, W, threadIdx.x, blockDim.x);
    __syncthreads();
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        double au = 1, av = 1, r = 1;
        if (i < n) { au = exact[i * 5]; av = exact[i * 5 + 1]; r = exact[i * 5 + 2]; }
        float a[3];
        sur_features(au, av, r, a);
        float o0, o1;
        mlp_forward_wave(L, a[0], a[1], a[2], o0, o1);
        if (i < n) {
            float bx, by;
            sur_envelope((float)au, (float)av, (float)r, bx, by);
            const double gx = au * (double)bx * exp((double)(norm.x + norm.y * o0));
            const double gy = av * (double)by * exp((double)(norm.z + norm.w * o1));
            const double ex = exact[i * 5 + 3], ey = exact[i * 5 + 4];
            err[i] = sqrt(((gx - ex) * (gx - ex) + (gy - ey) * (gy - ey)) / (ex * ex + ey * ey));
        }
    }
}

hipError_t launch_mlp_samples(int64_t n, uint64_t seed, uint32_t tag, float *out, double *exact, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_samples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, seed, tag, out, exact);
    return hipGetLastError();
}

hipError_t launch_mlp_train(const float *params, const float *samples, int64_t n, float *part, const float norm[4],
                            float inv_batch, uint64_t seed, uint32_t step, int chunks, int random, int nwg,
                            hipStream_t st)
{
    MlpTrainParams P;
    P.params = params; P.samples = samples; P.n = n; P.part = part;
    for (int i = 0; i < 4; i++) P.norm[i] = norm[i];
    P.inv_batch = inv_batch; P.seed = seed; P.step = step; P.chunks = chunks; P.random = random;
    hipLaunchKernelGGL(k_mlp_train, dim3(nwg), dim3(64), 0, st, P);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *params, float *m, float *v, const float *part, int nwg, float lr, float b1,
                           float b2, float eps, float bc1, float bc2, float *grad_out, float *loss_out,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, st, params, m, v, part, nwg,
                       lr, b1, b2, eps, bc1, bc2, grad_out, loss_out);
    return hipGetLastError();
}

hipError_t launch_mlp_eval(const float *W, const float *feat, int64_t n, int stride, float *out, hipStream_t st)
{
    int64_t nwg = (n + 255) / 256;
    if (nwg > 2048) nwg = 2048;
    if (nwg < 1) nwg = 1;
    hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)nwg), dim3(256), 0, st, W, feat, n, stride, out);
    return hipGetLastError();
}

hipError_t launch_mlp_field_error(const float *W, const double *exact, int64_t n, const float norm[4], double *err,
                                  hipStream_t st)
{
    int64_t nwg = (n + 255) / 256;
    if (nwg > 2048) nwg = 2048;
    if (nwg < 1) nwg = 1;
    const float4 nm = make_float4(norm[0], norm[1], norm[2], norm[3]);
    hipLaunchKernelGGL(k_mlp_field_error, dim3((unsigned)nwg), dim3(256), 0, st, W, exact, n, nm, err);
    return hipGetLastError();
}

hipError_t upload_faddeeva_tables_mlp(const FadTables &t)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(c_weid), t.weid, (kWeidN + 1) * sizeof(double));
}

}  // namespace cts
