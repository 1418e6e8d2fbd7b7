// cts_bbi.hip -- broadband impedance in the frequency domain, longitudinal,
// dipolar and quadrupolar planes (MPBunch::BBImpBunchInteraction,
// src/MPBunch.cpp:2143-2312).
//
// The reference bins a bunch on the centred grid of BoardBandImp::ReadInImp
// (src/BoardBandImp.cpp:71-90: N = 2 nimp - 1 bins of dz = c beta/(2 f_max)
// from -z_max), smooths the profile over one RF bucket, takes its real FFT,
// multiplies by Z_k, and returns V = -c2r(...)/N.  With Z extended as a
// Hermitian sequence (the c2r transform ignores Im Z_0) that is the circular
// convolution V_j = -sum_i p_i g_{(j - i) mod N} with the real kernel
//   g_m = (1/N) [Re Z_0 + 2 sum_{k=1}^{nimp-1} Re(Z_k e^{2 pi i k m / N})],
// computed once on the host.  The transverse planes multiply the spectrum by
// i Z and keep the + sign; the quadrupolar ones use the charge profile, the
// dipolar ones the x- and y-weighted profiles (sum of charge times x per
// bin).  Per turn: an LDS histogram per workgroup segment (with the
// dipolar planes, the per-bin sums of x and y as two-limb fixed-point
// integers, so they do not depend on the order of the atomics), one block
// per bunch for the smoothing and the convolutions, and a kick sweep.  The
// smoothing is GSL's gsl_filter_gaussian (GSL 2.8, the version the
// reference builds against) with K = 51, alpha = 10 and
// GSL_FILTER_END_PADVALUE: y_i = sum_j w_j x_{clamp(i + j - 25)},
// w_j = exp(-alpha^2 ((j - 25)/25)^2 / 2) normalised to sum 1, followed by
// the reference's rescaling to the input's sum (src/MPBunch.cpp:2373-2410).
#include <cmath>
#include <string>
#include <vector>

#include "cts_internal.hpp"

namespace cts {

int bbimp_table(const cts_bbimp_config &c, const cts_ring &ring, BbiTable &t, std::string &err)
{
    // planes in the reference's flag order: Z, Dx, Dy, Qx, Qy
    const double *zr[5] = {c.zz_re, c.zdx_re, c.zdy_re, c.zqx_re, c.zqy_re};
    const double *zi[5] = {c.zz_im, c.zdx_im, c.zdy_im, c.zqx_im, c.zqy_im};
    bool given = true;
    for (int q = 0; q < 5; q++) {
        t.planes[q] = c.flags[q] != 0;
        given = given && (!t.planes[q] || (zr[q] && zi[q]));
    }
    const int n = 2 * c.nimp - 1;
    const bool dip = t.planes[1] || t.planes[2];
    const int max_bins = dip ? kBbiDipMaxBins : kBbiMaxBins;
    if (c.nimp < 2 || n > max_bins || !(c.f_max > 0) || !given) {
        err = "cts_set_bbimp: nimp in [2, " + std::to_string((max_bins + 1) / 2) +
              "] (the bound with the dipolar planes on), f_max > 0, the impedance of every plane switched on";
        return CTS_ERR_ARG;
    }
    const bool tx = t.planes[1] || t.planes[3], ty = t.planes[2] || t.planes[4];
    if ((tx && !(c.beta_x > 0)) || (ty && !(c.beta_y > 0))) {
        err = "cts_set_bbimp: beta_x, beta_y > 0 for the transverse planes";
        return CTS_ERR_ARG;
    }
    t.bx = tx ? c.beta_x : 1.0;
    t.by = ty ? c.beta_y : 1.0;
    const int nimp = c.nimp;
    const double dt = 1.0 / (2 * c.f_max);
    t.n = n;
    t.dz = dt * kClight * ring.rbeta;
    const double zmax = t.dz * (nimp - 1);
    t.zmin = -zmax;
    // one RF bucket, indexed as GetSmoothedBunchProfileGassionFilter does
    const double rf_len = ring.t0 / ring.harmonic * kClight;
    t.is = (int)((-rf_len / 2. - t.zmin + t.dz / 2) / t.dz);
    t.ie = (int)((rf_len / 2. - t.zmin + t.dz / 2) / t.dz);
    if (-rf_len / 2. - t.zmin + t.dz / 2 < 0 || t.ie > n || t.ie <= t.is) {
        err = "cts_set_bbimp: the impedance grid (z_max " + std::to_string(zmax) +
              " m) must cover one RF bucket (" + std::to_string(rf_len) + " m)";
        return CTS_ERR_ARG;
    }
    t.g.assign(5 * (size_t)n, 0.0);
    for (int q = 0; q < 5; q++) {
        if (!t.planes[q]) continue;
        // the transverse planes multiply the spectrum by i Z: (re, im) -> (-im, re)
        auto re = [&](int k) { return q == 0 ? zr[q][k] : -zi[q][k]; };
        auto im = [&](int k) { return q == 0 ? zi[q][k] : zr[q][k]; };
        for (int m = 0; m < n; m++) {
            double s = 0;
            for (int k = 1; k < nimp; k++) {
                // exact reduction of the phase k m / n
                const double ph = 2 * M_PI * (double)(((int64_t)k * m) % n) / n;
                s += re(k) * std::cos(ph) - im(k) * std::sin(ph);
            }
            t.g[(size_t)q * n + m] = (re(0) + 2 * s) / n;
        }
    }
    const int half = (kBbiTaps - 1) / 2;
    const double alpha = 10.0;
    double sum = 0;
    for (int j = 0; j < kBbiTaps; j++) {
        const double xj = (j - half) / (double)half;
        t.taps[j] = std::exp(-0.5 * alpha * alpha * xj * xj);
        sum += t.taps[j];
    }
    for (int j = 0; j < kBbiTaps; j++) t.taps[j] /= sum;
    return CTS_OK;
}

namespace {

// bin of z as the reference indexes it, (int)((z - zmin + dz/2)/dz), or -1
// off the grid (the reference does not test the range)
__device__ __forceinline__ int bbi_bin(double z, const BbiParams &P)
{
    const double q = (z - P.zmin + P.dz / 2) / P.dz;
    return (q > -1.0 && q < (double)P.n) ? (int)q : -1;
}

// survivors before this turn's loss marking
__device__ __forceinline__ bool bbi_alive(uint8_t f) { return (f & 3u) == 0u || (f & kNewLost); }

constexpr int kUnroll = 4;  // particles per thread per step of the sweeps

// histogram copies in LDS: one per wave when they fit in 32 KB (a bunch
// fills few bins, so the waves' atomics would collide on one copy); one
// with the dipolar sums beside it
__host__ __device__ constexpr int bbi_copies(int n, bool dip) { return (!dip && n <= 2048) ? kBlock / 64 : 1; }

// a coordinate as two integer limbs, v ~ hi 2^-30 + lo 2^-62, hi =
// floor(v 2^30), 0 <= lo <= 2^32: the limb sums of up to 2^31 values with
// |v| < 1 stay inside 64 bits, and their total does not depend on the order
// of the additions (the conversion is exact up to 2^-85 for |v| < 2^-30
// below zero, a fixed rounding of the value itself)
__device__ __forceinline__ void fx2_split(double v, unsigned long long &hi, unsigned long long &lo)
{
    const double h = floor(ldexp(v, 30));
    hi = (unsigned long long)(long long)h;
    lo = (unsigned long long)(long long)rint(ldexp(v - ldexp(h, -30), 62));
}
__device__ __forceinline__ double fx2_value(unsigned long long hi, unsigned long long lo)
{
    return ldexp((double)(long long)hi, -30) + ldexp((double)(long long)lo, -62);
}

}  // namespace

__global__ void __launch_bounds__(kBlock) k_bbi_bin(BbiParams P)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t hl[];
    const int tid = threadIdx.x;
    const bool dipx = P.planes[1] != 0, dipy = P.planes[2] != 0, dip = dipx || dipy;
    const int nc = bbi_copies(P.n, dip);
    uint32_t *mine = hl + (tid / 64 % nc) * P.n;
    // dipolar sums [n][4]: x hi, x lo, y hi, y lo (8-byte aligned after the counts)
    unsigned long long *ds = reinterpret_cast<unsigned long long *>(hl + ((nc * P.n + 1) & ~1));
    for (int s = P.wg_seg[blockIdx.x]; s < P.wg_seg[blockIdx.x + 1]; s++) {
        const int b = P.seg_bunch[s];
        if (P.scale[b] == 0.0) continue;  // uniform: the whole block skips
        for (int j = tid; j < nc * P.n; j += kBlock) hl[j] = 0u;
        if (dip)
            for (int j = tid; j < 4 * P.n; j += kBlock) ds[j] = 0ull;
        __syncthreads();
        // four particles per thread in flight: all loads before any use
        const int64_t e = P.seg_end[s];
        for (int64_t i0 = P.seg_begin[s] + tid; i0 < e; i0 += kUnroll * kBlock) {
            double zz[kUnroll], xx[kUnroll], yy[kUnroll];
            uint8_t ff[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int64_t i = i0 + (int64_t)u * kBlock;
                ff[u] = i < e ? P.flag[i] : (uint8_t)3u;
                zz[u] = i < e ? P.z[i] : 0.0;
                xx[u] = (i < e && dipx) ? P.x[i] : 0.0;
                yy[u] = (i < e && dipy) ? P.y[i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                if (!bbi_alive(ff[u])) continue;
                const int k = bbi_bin(zz[u], P);
                if (k < 0) continue;
                atomicAdd(&mine[k], 1u);
                if (dipx) {
                    unsigned long long hi, lo;
                    fx2_split(xx[u], hi, lo);
                    atomicAdd(&ds[4 * k], hi);
                    atomicAdd(&ds[4 * k + 1], lo);
                }
                if (dipy) {
                    unsigned long long hi, lo;
                    fx2_split(yy[u], hi, lo);
                    atomicAdd(&ds[4 * k + 2], hi);
                    atomicAdd(&ds[4 * k + 3], lo);
                }
            }
        }
        __syncthreads();
        for (int j = tid; j < P.n; j += kBlock) {
            uint32_t c = 0;
            for (int q = 0; q < nc; q++) c += hl[q * P.n + j];
            if (c) atomicAdd(&P.hist[(int64_t)b * P.n + j], c);
        }
        if (dip)
            for (int j = tid; j < 4 * P.n; j += kBlock) {
                const unsigned long long v = ds[j];
                if (v) atomicAdd(&P.dsum[(int64_t)b * 4 * P.n + j], v);
            }
        __syncthreads();
    }
}

// The smoothed profile over [is, ie) of one bunch from its raw profile
// raw(i) (the reference's GetSmoothedBunchProfileGassionFilter), raw values
// elsewhere; written to prof[0, n).
template <typename Raw>
__device__ void bbi_smooth(const BbiParams &P, Raw raw, double *prof, double (*red)[kBlock])
{
    const int tid = threadIdx.x;
    const int nr = P.ie - P.is, half = (kBbiTaps - 1) / 2;
    double s0 = 0, s1 = 0;
    for (int i = tid; i < nr; i += kBlock) {
        double y = 0;
        for (int j = 0; j < kBbiTaps; j++) {
            int k = i + j - half;
            k = k < 0 ? 0 : (k >= nr ? nr - 1 : k);
            y += P.taps[j] * raw(P.is + k);
        }
        s0 += raw(P.is + i);
        s1 += y;
        prof[P.is + i] = y;
    }
    red[0][tid] = s0;
    red[1][tid] = s1;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) { red[0][tid] += red[0][tid + w]; red[1][tid] += red[1][tid + w]; }
        __syncthreads();
    }
    const double r = red[1][0] != 0.0 ? red[0][0] / red[1][0] : 0.0;
    __syncthreads();  // red is reused by the next profile
    for (int i = tid; i < P.n; i += kBlock) {
        if (i >= P.is && i < P.ie) prof[i] *= r;
        else prof[i] = raw(i);
    }
    __syncthreads();
}

// one block per bunch: the profiles, their smoothing over [is, ie), the
// convolutions; the histogram and dipolar sums of the bunch are cleared for
// the next turn (fills, no memset node)
__global__ void __launch_bounds__(kBlock) k_bbi_wake(BbiParams P)
{
    __shared__ double red[2][kBlock];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int64_t nbn = (int64_t)P.nb * P.n;
    uint32_t *cnt = P.hist + (int64_t)b * P.n;
    unsigned long long *ds = P.dsum ? P.dsum + (int64_t)b * 4 * P.n : nullptr;
    double *prof[3] = {P.prof + (int64_t)b * P.n, P.prof + nbn + (int64_t)b * P.n,
                       P.prof + 2 * nbn + (int64_t)b * P.n};
    const double sc = P.scale[b];
    if (sc == 0.0) {
        for (int j = tid; j < P.n; j += kBlock) {
            for (int q = 0; q < 5; q++) P.wake[((int64_t)q * P.nb + b) * P.n + j] = 0.0;
            for (int r = 0; r < 3; r++) prof[r][j] = 0.0;
            cnt[j] = 0u;
            if (ds)
                for (int c = 0; c < 4; c++) ds[4 * j + c] = 0ull;
        }
        return;
    }
    bbi_smooth(P, [&](int i) { return (double)cnt[i] * sc; }, prof[0], red);
    if (P.planes[1]) bbi_smooth(P, [&](int i) { return sc * fx2_value(ds[4 * i], ds[4 * i + 1]); }, prof[1], red);
    if (P.planes[2]) bbi_smooth(P, [&](int i) { return sc * fx2_value(ds[4 * i + 2], ds[4 * i + 3]); }, prof[2], red);
    // longitudinal V = -IDFT(I Z); transverse V = +IDFT(I i Z), the dipolar
    // planes from the weighted profiles
    const int src[5] = {0, 1, 2, 0, 0};
    for (int q = 0; q < 5; q++) {
        if (!P.planes[q]) continue;
        const double *g = P.g + (int64_t)q * P.n;
        const double *pr = prof[src[q]];
        double *wake = P.wake + ((int64_t)q * P.nb + b) * P.n;
        for (int j = tid; j < P.n; j += kBlock) {
            double v = 0;
            int m = j;  // (j - i) mod n for i = 0, 1, ...
            for (int i = 0; i < P.n; i++) {
                v += pr[i] * g[m];
                m = m == 0 ? P.n - 1 : m - 1;
            }
            wake[j] = q == 0 ? -v : v;
        }
    }
    __syncthreads();
    for (int j = tid; j < P.n; j += kBlock) {
        cnt[j] = 0u;
        if (ds)
            for (int c = 0; c < 4; c++) ds[4 * j + c] = 0ull;
    }
}

__global__ void __launch_bounds__(kBlock) k_bbi_kick(BbiParams P)
{
    const int tid = threadIdx.x;
    const int64_t nbn = (int64_t)P.nb * P.n;
    const bool tx = P.planes[1] || P.planes[3], ty = P.planes[2] || P.planes[4];
    for (int s = P.wg_seg[blockIdx.x]; s < P.wg_seg[blockIdx.x + 1]; s++) {
        const int b = P.seg_bunch[s];
        if (P.scale[b] == 0.0) continue;
        const double *wz = P.wake + (int64_t)b * P.n;
        const double *wdx = P.wake + nbn + (int64_t)b * P.n, *wdy = P.wake + 2 * nbn + (int64_t)b * P.n;
        const double *wqx = P.wake + 3 * nbn + (int64_t)b * P.n, *wqy = P.wake + 4 * nbn + (int64_t)b * P.n;
        const int64_t e = P.seg_end[s];
        for (int64_t i0 = P.seg_begin[s] + tid; i0 < e; i0 += kUnroll * kBlock) {
            double zz[kUnroll], d[kUnroll];
            uint8_t ff[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int64_t i = i0 + (int64_t)u * kBlock;
                ff[u] = i < e ? P.flag[i] : (uint8_t)3u;
                zz[u] = i < e ? P.z[i] : 0.0;
                d[u] = (i < e && P.planes[0]) ? P.dp[i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int64_t i = i0 + (int64_t)u * kBlock;
                if (!bbi_alive(ff[u])) continue;
                const int k = bbi_bin(zz[u], P);
                if (k < 0) continue;
                // the reference's order: Z, Qx, Qy, then Dx, Dy
                if (P.planes[0]) P.dp[i] = d[u] + wz[k] / P.energy / P.beta2;
                if (tx) {
                    double v = P.px[i];
                    if (P.planes[3]) v += wqx[k] / P.energy / P.beta2 * P.x[i] / P.bx;
                    if (P.planes[1]) v += wdx[k] / P.energy / P.beta2 / P.bx;
                    P.px[i] = v;
                }
                if (ty) {
                    double v = P.py[i];
                    if (P.planes[4]) v += wqy[k] / P.energy / P.beta2 * P.y[i] / P.by;
                    if (P.planes[2]) v += wdy[k] / P.energy / P.beta2 / P.by;
                    P.py[i] = v;
                }
            }
        }
    }
}

size_t bbi_bin_lds(int n, bool dip)
{
    const size_t counts = (size_t)bbi_copies(n, dip) * n;
    return dip ? ((counts + 1) & ~(size_t)1) * sizeof(uint32_t) + (size_t)4 * n * sizeof(unsigned long long)
               : counts * sizeof(uint32_t);
}

hipError_t launch_bbi_bin(const BbiParams &p, int nwg, hipStream_t st)
{
    const bool dip = p.planes[1] || p.planes[2];
    hipLaunchKernelGGL(k_bbi_bin, dim3(nwg), dim3(kBlock), bbi_bin_lds(p.n, dip), st, p);
    return hipGetLastError();
}
hipError_t launch_bbi_wake(const BbiParams &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_bbi_wake, dim3(p.nb), dim3(kBlock), 0, st, p);
    return hipGetLastError();
}
hipError_t launch_bbi_kick(const BbiParams &p, int nwg, hipStream_t st)
{
    hipLaunchKernelGGL(k_bbi_kick, dim3(nwg), dim3(kBlock), 0, st, p);
    return hipGetLastError();
}

}  // namespace cts
