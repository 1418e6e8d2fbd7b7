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
//
// The time-domain method (MPBunch::BBImpBunchInteractionTD,
// src/MPBunch.cpp:2029-2142, the reference's default) bins a bunch over one
// RF bucket, (int)((z + L/2)/dz) on (int)(L/dz) bins with dz the wake
// table's step, smooths the charge profile rho (charge per bin) the same
// way, takes the mean x and y of each bin, and sums the wake table over the
// bins behind each one: V_i = -sum_{j >= i} w[j - i] rho_j (the dipolar
// planes rho_j <x>_j).  No transform: the sums run over the bins where rho
// is nonzero (zeros add nothing), one block per bunch.
//
// A bunch fills a small part of the grid, so the binning histograms each
// workgroup segment over the range of bins its particles reach, in LDS
// windows of up to kBbiWin bins: one sweep over the segment finds the
// range, one sweep per window bins it.
#include <climits>
#include <cmath>
#include <string>
#include <vector>

#include "cts_internal.hpp"

namespace cts {

namespace {

// time domain: the wake table stands in for the kernels
int bbimp_table_td(const cts_bbimp_config &c, const cts_ring &ring, BbiTable &t, std::string &err)
{
    const double *w[5] = {c.wake_z, c.wake_dx, c.wake_dy, c.wake_qx, c.wake_qy};
    bool given = c.wake_pos != nullptr;
    for (int q = 0; q < 5; q++) given = given && (!t.planes[q] || w[q]);
    if (c.nwake < 2 || !given) {
        err = "cts_set_bbimp: time domain: nwake >= 2, wake_pos and the wake of every plane switched on";
        return CTS_ERR_ARG;
    }
    // dzBin and nBinBunchDen of src/MPBunch.cpp:2036-2040
    const double dz = std::fabs(c.wake_pos[1] - c.wake_pos[0]);
    const double rf_len = ring.t0 / ring.harmonic * kClight;
    const double nbins = rf_len / dz;
    if (!(dz > 0) || !(nbins >= 1.0) || !(nbins < kBbiMaxGridTD + 1.0)) {
        err = "cts_set_bbimp: time domain: the wake's step must cut the RF bucket (" + std::to_string(rf_len) +
              " m) into 1 to " + std::to_string(kBbiMaxGridTD) + " bins";
        return CTS_ERR_ARG;
    }
    t.td = 1;
    t.nw = c.nwake;
    t.n = (int)(1 * rf_len / dz);
    t.dz = dz;
    t.zmin = -1 * rf_len / 2;
    t.half = 0.0;
    t.is = 0;
    t.ie = t.n;
    t.g.assign(5 * (size_t)t.nw, 0.0);
    for (int q = 0; q < 5; q++)
        if (t.planes[q]) std::copy(w[q], w[q] + t.nw, t.g.begin() + (size_t)q * t.nw);
    return CTS_OK;
}

}  // namespace

int bbimp_table(const cts_bbimp_config &c, const cts_ring &ring, BbiTable &t, std::string &err)
{
    // planes in the reference's flag order: Z, Dx, Dy, Qx, Qy
    for (int q = 0; q < 5; q++) t.planes[q] = c.flags[q] != 0;
    const bool tx = t.planes[1] || t.planes[3], ty = t.planes[2] || t.planes[4];
    if ((tx && !(c.beta_x > 0)) || (ty && !(c.beta_y > 0))) {
        err = "cts_set_bbimp: beta_x, beta_y > 0 for the transverse planes";
        return CTS_ERR_ARG;
    }
    t.bx = tx ? c.beta_x : 1.0;
    t.by = ty ? c.beta_y : 1.0;
    int rc;
    if (c.time_domain == 1) {
        if ((rc = bbimp_table_td(c, ring, t, err))) return rc;
    } else if (c.time_domain == 0) {
        const double *zr[5] = {c.zz_re, c.zdx_re, c.zdy_re, c.zqx_re, c.zqy_re};
        const double *zi[5] = {c.zz_im, c.zdx_im, c.zdy_im, c.zqx_im, c.zqy_im};
        bool given = true;
        for (int q = 0; q < 5; q++) given = given && (!t.planes[q] || (zr[q] && zi[q]));
        const int n = 2 * c.nimp - 1;
        if (c.nimp < 2 || n > kBbiMaxGridFD || !(c.f_max > 0) || !given) {
            err = "cts_set_bbimp: nimp in [2, " + std::to_string((kBbiMaxGridFD + 1) / 2) +
                  "], f_max > 0, the impedance of every plane switched on";
            return CTS_ERR_ARG;
        }
        const int nimp = c.nimp;
        const double dt = 1.0 / (2 * c.f_max);
        t.n = n;
        t.dz = dt * kClight * ring.rbeta;
        t.half = t.dz / 2;
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
    } else {
        // the reference prints "wrong setting in BoardBandImpedance namelist"
        // every turn and applies nothing (src/MPBeam.cpp:1345)
        err = "cts_set_bbimp: time_domain 0 (frequency domain) or 1 (time domain)";
        return CTS_ERR_ARG;
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

// bin of z as the reference indexes it, (int)((z - zmin + dz/2)/dz) in the
// frequency domain, (int)((z - zmin)/dz) in the time domain, or -1 off the
// grid (the reference does not test the range)
__device__ __forceinline__ int bbi_bin(double z, const BbiParams &P)
{
    const double q = (z - P.zmin + P.half) / P.dz;
    return (q > -1.0 && q < (double)P.n) ? (int)q : -1;
}

// survivors before this turn's loss marking
__device__ __forceinline__ bool bbi_alive(uint8_t f) { return (f & 3u) == 0u || (f & kNewLost); }

constexpr int kUnroll = 4;  // particles per thread per step of the sweeps

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

// LDS of the binning: counts, then (dipolar planes) four limb sums per bin
__host__ __device__ constexpr int bbi_count_words(bool dip) { return dip ? ((kBbiDipWin + 1) & ~1) : kBbiWin; }

__global__ void __launch_bounds__(kBlock) k_bbi_bin(BbiParams P)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t hl[];
    __shared__ int krange[2];
    const int tid = threadIdx.x;
    const bool dipx = P.planes[1] != 0, dipy = P.planes[2] != 0, dip = dipx || dipy;
    // dipolar sums [window][4]: x hi, x lo, y hi, y lo (8-byte aligned after the counts)
    unsigned long long *ds = reinterpret_cast<unsigned long long *>(hl + bbi_count_words(dip));
    for (int s = P.wg_seg[blockIdx.x]; s < P.wg_seg[blockIdx.x + 1]; s++) {
        const int b = P.seg_bunch[s];
        if (P.scale[b] == 0.0) continue;  // uniform: the whole block skips
        const int64_t e = P.seg_end[s];
        // the range of bins the segment's counted particles reach; a grid
        // that fits one window is taken whole (no sweep for the range)
        const bool whole = P.n <= (dip ? kBbiDipWin : kBbiWinCopies);
        if (tid == 0) { krange[0] = whole ? 0 : INT_MAX; krange[1] = whole ? P.n - 1 : -1; }
        __syncthreads();
        int rlo = INT_MAX, rhi = -1;
        for (int64_t i0 = P.seg_begin[s] + tid; !whole && i0 < e; i0 += kUnroll * kBlock) {
            double zz[kUnroll];
            uint8_t ff[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int64_t i = i0 + (int64_t)u * kBlock;
                ff[u] = i < e ? P.flag[i] : (uint8_t)3u;
                zz[u] = i < e ? P.z[i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int k = bbi_alive(ff[u]) ? bbi_bin(zz[u], P) : -1;
                if (k >= 0) { rlo = min(rlo, k); rhi = max(rhi, k); }
            }
        }
        if (rhi >= 0) { atomicMin(&krange[0], rlo); atomicMax(&krange[1], rhi); }
        __syncthreads();
        const int klo = krange[0], khi = krange[1];
        __syncthreads();  // krange is reset by the next segment
        if (khi < klo) continue;
        // narrow ranges: one copy of the counts per wave (a bunch fills few
        // bins, so the waves' atomics would collide on one copy)
        const int nc = (!dip && khi - klo < kBbiWinCopies) ? kBlock / 64 : 1;
        const int W = dip ? kBbiDipWin : (nc > 1 ? kBbiWinCopies : kBbiWin);
        uint32_t *mine = hl + (tid / 64 % nc) * W;
        for (int w0 = klo; w0 <= khi; w0 += W) {
            const int wn = min(W, khi - w0 + 1);
            for (int j = tid; j < nc * W; j += kBlock) hl[j] = 0u;
            if (dip)
                for (int j = tid; j < 4 * wn; j += kBlock) ds[j] = 0ull;
            __syncthreads();
            // four particles per thread in flight: all loads before any use
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
                    const int k = bbi_bin(zz[u], P) - w0;
                    if (k < 0 || k >= wn) continue;
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
            for (int j = tid; j < wn; j += kBlock) {
                uint32_t c = 0;
                for (int q = 0; q < nc; q++) c += hl[q * W + j];
                if (c) atomicAdd(&P.hist[(int64_t)b * P.n + w0 + j], c);
            }
            if (dip)
                for (int j = tid; j < 4 * wn; j += kBlock) {
                    const unsigned long long v = ds[j];
                    if (v) atomicAdd(&P.dsum[((int64_t)b * P.n + w0) * 4 + j], v);
                }
            __syncthreads();
        }
    }
}

// The smoothed profile over [is, ie) of one bunch from its raw profile
// raw(i) (the reference's GetSmoothedBunchProfileGassionFilter), raw values
// elsewhere; written to prof[0, n).  raw vanishes outside [lo, hi], so the
// smoothed values vanish outside [lo - 25, hi + 25]: only those are formed.
template <typename Raw>
__device__ void bbi_smooth(const BbiParams &P, Raw raw, double *prof, double (*red)[kBlock], int lo, int hi)
{
    const int tid = threadIdx.x;
    const int nr = P.ie - P.is, half = (kBbiTaps - 1) / 2;
    const int a = max(P.is, lo - half), e = min(P.ie, hi + half + 1);
    double s0 = 0, s1 = 0;
    for (int i = a + tid; i < e; i += kBlock) {
        double y = 0;
        for (int j = 0; j < kBbiTaps; j++) {
            int k = i - P.is + j - half;
            k = k < 0 ? 0 : (k >= nr ? nr - 1 : k);
            y += P.taps[j] * raw(P.is + k);
        }
        s0 += raw(i);
        s1 += y;
        prof[i] = y;
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
        if (i >= a && i < e) prof[i] *= r;
        else if (i >= P.is && i < P.ie) prof[i] = 0.0;
        else prof[i] = raw(i);
    }
    __syncthreads();
}

// one block per bunch: the profiles, their smoothing, the induced voltages;
// the histogram and dipolar sums of the bunch are cleared for the next turn
// (fills, no memset node)
__global__ void __launch_bounds__(kBlock) k_bbi_wake(BbiParams P)
{
    __shared__ double red[2][kBlock];
    __shared__ int krange[2];
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
    // the bins holding particles
    if (tid == 0) { krange[0] = INT_MAX; krange[1] = -1; }
    __syncthreads();
    {
        int lo = INT_MAX, hi = -1;
        for (int j = tid; j < P.n; j += kBlock)
            if (cnt[j]) { lo = min(lo, j); hi = max(hi, j); }
        if (hi >= 0) { atomicMin(&krange[0], lo); atomicMax(&krange[1], hi); }
    }
    __syncthreads();
    const bool any = krange[1] >= krange[0];
    const int lo = any ? krange[0] : 0, hi = any ? krange[1] : -1;
    // every nonzero value of a profile lies in [plo, phi]
    const int plo = max(0, lo - (kBbiTaps - 1) / 2), phi = min(P.n - 1, hi + (kBbiTaps - 1) / 2);
    const int src[5] = {0, 1, 2, 0, 0};
    if (!P.td) {
        bbi_smooth(P, [&](int i) { return (double)cnt[i] * sc; }, prof[0], red, lo, hi);
        if (P.planes[1])
            bbi_smooth(P, [&](int i) { return sc * fx2_value(ds[4 * i], ds[4 * i + 1]); }, prof[1], red, lo, hi);
        if (P.planes[2])
            bbi_smooth(P, [&](int i) { return sc * fx2_value(ds[4 * i + 2], ds[4 * i + 3]); }, prof[2], red, lo, hi);
        // longitudinal V = -IDFT(I Z); transverse V = +IDFT(I i Z), the
        // dipolar planes from the weighted profiles: circular convolutions
        // over the nonzero part of the profile, in the order of i
        for (int q = 0; q < 5; q++) {
            if (!P.planes[q]) continue;
            const double *g = P.g + (int64_t)q * P.n;
            const double *pr = prof[src[q]];
            double *wake = P.wake + ((int64_t)q * P.nb + b) * P.n;
            for (int j = tid; j < P.n; j += kBlock) {
                double v = 0;
                int m = j - plo;  // (j - i) mod n for i = plo, plo + 1, ...
                if (m < 0) m += P.n;
                for (int i = plo; i <= phi; i++) {
                    v += pr[i] * g[m];
                    m = m == 0 ? P.n - 1 : m - 1;
                }
                wake[j] = q == 0 ? -v : v;
            }
        }
    } else {
        // charge per bin, smoothed over the whole grid (src/MPBunch.cpp:
        // 2048-2056); the mean x and y of each bin (2058-2082)
        bbi_smooth(P, [&](int i) { return (double)cnt[i] * sc; }, prof[0], red, lo, hi);
        for (int r = 1; r < 3; r++) {
            if (!P.planes[r]) continue;
            for (int j = tid; j < P.n; j += kBlock)
                prof[r][j] = cnt[j] ? fx2_value(ds[4 * j + 2 * (r - 1)], ds[4 * j + 2 * (r - 1) + 1]) / cnt[j] : 0.0;
        }
        __syncthreads();
        // V_i = -sum_{j >= i} w[j - i] rho_j (the dipolar planes rho_j <x>_j),
        // j over the rows of the wake and the nonzero profile (2100-2111)
        for (int q = 0; q < 5; q++) {
            if (!P.planes[q]) continue;
            const double *w = P.g + (int64_t)q * P.nw;
            const double *rho = prof[0];
            const double *avg = (q == 1 || q == 2) ? prof[q] : nullptr;
            double *wake = P.wake + ((int64_t)q * P.nb + b) * P.n;
            for (int i = tid; i < P.n; i += kBlock) {
                double v = 0;
                const int j1 = (int)min((int64_t)phi, (int64_t)i + P.nw - 1);
                if (avg)
                    for (int j = max(i, plo); j <= j1; j++) v -= w[j - i] * rho[j] * avg[j];
                else
                    for (int j = max(i, plo); j <= j1; j++) v -= w[j - i] * rho[j];
                wake[i] = v;
            }
        }
    }
    __syncthreads();
    for (int j = lo + tid; j <= hi; j += kBlock) {
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
                if (P.planes[0]) P.dp[i] = d[u] + wz[k] / P.energy / P.beta2;
                if (P.td) {
                    // the time domain's order: Z, Dx, Dy, Qx, Qy
                    // (src/MPBunch.cpp:2113-2121)
                    if (tx) {
                        double v = P.px[i];
                        if (P.planes[1]) v += wdx[k] / P.energy / P.beta2 / P.bx;
                        if (P.planes[3]) v += wqx[k] / P.energy / P.beta2 / P.bx * P.x[i];
                        P.px[i] = v;
                    }
                    if (ty) {
                        double v = P.py[i];
                        if (P.planes[2]) v += wdy[k] / P.energy / P.beta2 / P.by;
                        if (P.planes[4]) v += wqy[k] / P.energy / P.beta2 / P.by * P.y[i];
                        P.py[i] = v;
                    }
                    continue;
                }
                // the frequency domain's order: Z, Qx, Qy, then Dx, Dy
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

size_t bbi_bin_lds(bool dip)
{
    return (size_t)bbi_count_words(dip) * sizeof(uint32_t) +
           (dip ? (size_t)4 * kBbiDipWin * sizeof(unsigned long long) : 0);
}

hipError_t launch_bbi_bin(const BbiParams &p, int nwg, hipStream_t st)
{
    const bool dip = p.planes[1] || p.planes[2];
    hipLaunchKernelGGL(k_bbi_bin, dim3(nwg), dim3(kBlock), bbi_bin_lds(dip), st, p);
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
