// cts_ions.hip -- beam-ion interaction at one interaction point (gfx950).
//
// The ion cloud of each gas species lives in HBM as FP64 arrays x, y, vx,
// vy plus an alive byte, in two buffers (the turn-end compaction writes the
// other one).  Bunch passages are sequential (bunch j+1 sees the cloud
// bunch j left, src/MPBeam.cpp:2590-2609); each passage is one launch over
// the cloud that, per ion,
//   - generates the passage's new ions from the bunch's Gaussian
//     (IonGenerator, src/LatticeInterActionPoint.cpp:985-1019),
//   - accumulates the cloud moments of the state before the kick
//     (IonRMSCal, src/LatticeInterActionPoint.cpp:1113-1155),
//   - applies the Bassetti-Erskine kick of the bunch, the drift to the next
//     bunch and the loss test (src/MPBunch.cpp:1113-1160,
//     src/LatticeInterActionPoint.cpp:1158-1210).
// The last workgroup to finish reduces the moments (fixed order) into the
// electron-side kick parameters of that bunch (src/MPBunch.cpp:1165-1206).
// Lost ions stay in their slots, flagged, until the compaction at the end
// of the turn, which keeps the survivors in order.
#include "cts_internal.hpp"

namespace cts {

constexpr int kIonPart = 6;  // n, sum dx, sum dy, sum dx^2, sum dy^2, n alive after the loss test

// One new ion of species p for passage `passage`: Box-Muller pairs of the
// Philox blocks {t, passage | p << 24, turn, 2^31 | attempt/2}, rejected
// outside the 2-sigma ellipse, at most 64 attempts.
__device__ __forceinline__ void ion_generate(uint64_t seed, uint32_t turn, uint32_t passage, int p,
                                             uint32_t t, double xm, double ym, double sx, double sy,
                                             double &gx, double &gy)
{
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    double tx = xm, ty = ym;
    for (uint32_t a = 0; a < 64u; a++) {
        const u32x4 o = philox4x32_10(t, passage | ((uint32_t)p << 24), turn, 0x80000000u | (a >> 1), k0, k1);
        float n0, n1;
        if ((a & 1u) == 0) box_muller(o.v[0], o.v[1], n0, n1);
        else box_muller(o.v[2], o.v[3], n0, n1);
        tx = (double)n0 * sx + xm;
        ty = (double)n1 * sy + ym;
        const double u = (tx - xm) / sx, v = (ty - ym) / sy;
        if (u * u + v * v > 4.E0) continue;
        break;
    }
    gx = tx;
    gy = ty;
}

// SUR: the field of the bunch at each ion from the surrogate inside its
// domain (exact field outside it).
template <bool SUR>
__global__ void __launch_bounds__(kIonBlock)
k_ion_passage(IonPassParams P)
{
    constexpr int kNV = CTS_MAX_SPECIES * kIonPart;
    __shared__ double red[kIonBlock / 64][kNV];
    __shared__ uint32_t last;
    __shared__ __attribute__((aligned(16))) MlpLds L;
    if (SUR) {
        load_mlp_lds(L, P.sur_params, threadIdx.x, kIonBlock);
        __syncthreads();
    }
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = P.bunch;
    const double xm = P.bview[4 * b + 0], ym = P.bview[4 * b + 1];
    const double sx = P.bview[4 * b + 2], sy = P.bview[4 * b + 3];
    // loss boundary: min(ionLossBoundary * largest bunch size, aperture)
    const double lbx = P.loss_boundary * P.size_max[0], lby = P.loss_boundary * P.size_max[1];
    const double bx = P.ap_x > lbx ? lbx : P.ap_x;
    const double by = P.ap_y > lby ? lby : P.ap_y;

    for (int p = 0; p < P.nspec; p++) {
        const int64_t len = P.len[p];
        const bool overflow = P.n_alive[p] + P.n_new > P.max_ions;
        const int64_t end = overflow ? len : len + P.n_new;
        const double cI = P.coeffI[p];
        double acc[kIonPart] = {0, 0, 0, 0, 0, 0};
        double *X = P.x[p], *Y = P.y[p], *VX = P.vx[p], *VY = P.vy[p];
        uint8_t *A = P.alive[p];
        // whole blocks per iteration, so that every lane of a wave reaches
        // the surrogate together
        for (int64_t base = (int64_t)blockIdx.x * kIonBlock; base < end; base += (int64_t)gridDim.x * kIonBlock) {
            const int64_t i = base + tid;
            bool act = i < end;
            double x = 0.0, y = 0.0, vx = 0.0, vy = 0.0;
            if (act) {
                if (i >= len) {
                    ion_generate(P.seed, P.turn, P.passage, p, (uint32_t)(i - len), xm, ym, sx, sy, x, y);
                } else if (A[i]) {
                    x = X[i]; y = Y[i]; vx = VX[i]; vy = VY[i];
                } else {
                    act = false;
                }
            }
            const double dx = x - xm, dy = y - ym;
            double fx = 0.0, fy = 0.0;
            if (SUR) {
                const bool ok = sur_field(L, P.sur_norm, act, dx, dy, sx, sy, fx, fy);
                if (act && !ok) ion_field(dx, dy, sx, sy, fx, fy);
            } else if (act) {
                ion_field(dx, dy, sx, sy, fx, fy);
            }
            if (!act) continue;
            acc[0] += 1.0;
            acc[1] += dx;
            acc[2] += dy;
            acc[3] += dx * dx;
            acc[4] += dy * dy;
            vx += cI * fx;
            vy += cI * fy;
            x += vx * P.dt;
            y += vy * P.dt;
            const bool keep = !(fabs(x) > bx || fabs(y) > by);
            acc[5] += keep ? 1.0 : 0.0;
            X[i] = x; Y[i] = y; VX[i] = vx; VY[i] = vy;
            A[i] = keep ? 1 : 0;
        }
#pragma unroll
        for (int j = 0; j < kIonPart; j++) {
            const double r = wave_sum(acc[j]);
            if (lane == 0) red[wave][p * kIonPart + j] = r;
        }
    }
    __syncthreads();
    if (tid < kNV) {
        double r = 0.0;
        if (tid < P.nspec * kIonPart)
            for (int w = 0; w < kIonBlock / 64; w++) r += red[w][tid];
        P.part[(int64_t)blockIdx.x * kNV + tid] = r;
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) last = atomicAdd(P.cnt, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    // the last workgroup folds the partials: thread t takes workgroups
    // t, t + 256, ... in order, then a fixed wave/LDS tree (deterministic)
    {
        double r[kNV];
#pragma unroll
        for (int j = 0; j < kNV; j++) r[j] = 0.0;
        for (uint32_t w = tid; w < gridDim.x; w += kIonBlock) {
            const double *src = P.part + (int64_t)w * kNV;
#pragma unroll
            for (int j = 0; j < kNV; j++) r[j] += src[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kNV; j++) {
            const double s = wave_sum(r[j]);
            if (lane == 0) red[wave][j] = s;
        }
        __syncthreads();
        double s = 0.0;
        if (tid < kNV)
            for (int w = 0; w < kIonBlock / 64; w++) s += red[w][tid];
        __syncthreads();
        if (tid < kNV) red[0][tid] = s;
        __syncthreads();
    }
    if (tid < P.nspec) {
        const int p = tid;
        const double n = red[0][p * kIonPart + 0];
        double mx = P.mean_state[2 * p], my = P.mean_state[2 * p + 1], rx = 0, ry = 0;
        if (n >= 2) {
            const double ax = red[0][p * kIonPart + 1] / n, ay = red[0][p * kIonPart + 2] / n;
            mx = xm + ax;
            my = ym + ay;
            rx = sqrt(fmax(red[0][p * kIonPart + 3] / n - ax * ax, 0.0));
            ry = sqrt(fmax(red[0][p * kIonPart + 4] / n - ay * ay, 0.0));
        } else if (n == 1) {
            // one ion: uncentred rms, the mean keeps its last value
            // (src/LatticeInterActionPoint.cpp:1125-1129)
            rx = fabs(xm + red[0][p * kIonPart + 1]);
            ry = fabs(ym + red[0][p * kIonPart + 2]);
        }
        P.mean_state[2 * p] = mx;
        P.mean_state[2 * p + 1] = my;
        // ion number of the cloud: this bunch's macro-ion charge times the
        // macro-ion count (src/MPBunch.cpp:1172)
        const double nI0 = P.ion_q[b * P.nspec + p] * n;
        double *e = P.ion_e + ((int64_t)b * CTS_MAX_SPECIES + p) * 5;
        e[0] = 2.0 * nI0 * kRe / P.gamma;
        e[1] = mx;
        e[2] = my;
        e[3] = rx;
        e[4] = ry;
        double *c = P.cloud + ((int64_t)b * CTS_MAX_SPECIES + p) * 6;
        c[0] = n; c[1] = mx; c[2] = my; c[3] = rx; c[4] = ry; c[5] = nI0;
        const bool overflow = P.n_alive[p] + P.n_new > P.max_ions;
        if (overflow) atomicOr(P.status, 2u);
        P.n_alive[p] = (int64_t)red[0][p * kIonPart + 5];
        if (!overflow) P.len[p] += P.n_new;
    }
    if (tid == 0) *P.cnt = 0u;
}

// Bunch views for the ion phase: centroid and rms sizes of every bunch
// from the last moment record (rmsRx = sqrt(emitX * betaX[k]),
// src/MPBunch.cpp:481-482), and their maxima (MPGetBeamInfo,
// src/MPBeam.cpp:2203-2258).
__global__ void __launch_bounds__(256)
k_ion_prep(const cts_moments *cur, int nb, double beta_x, double beta_y, double *bview, double *size_max,
           uint32_t *fb_count)
{
    if (threadIdx.x == 0 && fb_count) *fb_count = 0u;
    __shared__ double mx[256], my[256];
    double ax = -INFINITY, ay = -INFINITY;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const double *c = cur[b].c2;
        const double rx = sqrt(sqrt(c[0] * c[2] - c[1] * c[1]) * beta_x);
        const double ry = sqrt(sqrt(c[3] * c[5] - c[4] * c[4]) * beta_y);
        bview[4 * b + 0] = cur[b].mean[0];
        bview[4 * b + 1] = cur[b].mean[2];
        bview[4 * b + 2] = rx;
        bview[4 * b + 3] = ry;
        ax = fmax(ax, rx);
        ay = fmax(ay, ry);
    }
    mx[threadIdx.x] = ax;
    my[threadIdx.x] = ay;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = mx[0], c = my[0];
        for (int t = 1; t < (int)blockDim.x; t++) { a = fmax(a, mx[t]); c = fmax(c, my[t]); }
        size_max[0] = a;
        size_max[1] = c;
    }
}

// Stable compaction of one species' cloud into the other buffer: chunks of
// kIonChunk slots, 16 per thread; pass 1 counts the survivors of each chunk,
// pass 2 places them after the survivors of all earlier chunks.
__global__ void __launch_bounds__(256)
k_compact_count(const uint8_t *alive, const int64_t *len_p, uint32_t *chunk_cnt)
{
    __shared__ uint32_t s[256];
    const int64_t len = *len_p;
    const int64_t base = (int64_t)blockIdx.x * kIonChunk + threadIdx.x * 16;
    uint32_t c = 0;
    for (int k = 0; k < 16; k++)
        if (base + k < len) c += alive[base + k];
    s[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) chunk_cnt[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256)
k_compact_scatter(IonCompactParams P)
{
    __shared__ uint32_t s[256];
    __shared__ uint64_t part[256];
    const int64_t len = *P.len;
    // survivors of the chunks before this one
    uint64_t a = 0;
    for (uint32_t w = threadIdx.x; w < blockIdx.x; w += 256) a += P.chunk_cnt[w];
    part[threadIdx.x] = a;
    const int64_t base = (int64_t)blockIdx.x * kIonChunk + threadIdx.x * 16;
    uint32_t c = 0;
    for (int k = 0; k < 16; k++)
        if (base + k < len) c += P.alive[base + k];
    s[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    const uint64_t off = part[0];
    // inclusive scan of the per-thread counts (Hillis-Steele)
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t v = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t dst = off + s[threadIdx.x] - c;
    for (int k = 0; k < 16; k++) {
        const int64_t i = base + k;
        if (i < len && P.alive[i]) {
            P.ox[dst] = P.x[i]; P.oy[dst] = P.y[i]; P.ovx[dst] = P.vx[i]; P.ovy[dst] = P.vy[i];
            P.oalive[dst] = 1;
            dst++;
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) *P.new_len = (int64_t)(off + s[255]);
}

// this translation unit's copy of the Faddeeva tables
hipError_t upload_faddeeva_tables_ions(const double ea2n2[64], const double fad[3])
{
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_ea2n2), ea2n2, 64 * sizeof(double));
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_fad), fad, 3 * sizeof(double));
}

hipError_t launch_ion_passage(const IonPassParams &p, int nwg, hipStream_t st)
{
    if (p.sur_params)
        hipLaunchKernelGGL(k_ion_passage<true>, dim3(nwg), dim3(kIonBlock), 0, st, p);
    else
        hipLaunchKernelGGL(k_ion_passage<false>, dim3(nwg), dim3(kIonBlock), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_ion_prep(const cts_moments *cur, int nb, double bx, double by, double *bview, double *size_max,
                           uint32_t *fb_count, hipStream_t st)
{
    hipLaunchKernelGGL(k_ion_prep, dim3(1), dim3(256), 0, st, cur, nb, bx, by, bview, size_max, fb_count);
    return hipGetLastError();
}

// len_bound: an upper bound of *p.len known on the host; the grid covers it.
hipError_t launch_ion_compact(const IonCompactParams &p, int64_t len_bound, hipStream_t st)
{
    int64_t nch = (len_bound + kIonChunk - 1) / kIonChunk;
    if (nch < 1) nch = 1;
    hipLaunchKernelGGL(k_compact_count, dim3((unsigned)nch), dim3(256), 0, st, p.alive, p.len, p.chunk_cnt);
    hipLaunchKernelGGL(k_compact_scatter, dim3((unsigned)nch), dim3(256), 0, st, p);
    return hipGetLastError();
}

}  // namespace cts
