// cts_ions.hip -- beam-ion interaction at one interaction point (gfx950).
//
// The ion cloud of each gas species lives in HBM as FP64 arrays x, y, vx,
// vy plus an alive byte, in two buffers (the turn-end compaction writes the
// other one).  In the reference the bunches pass the cloud one after the
// other (src/MPBeam.cpp:2590-2609): per passage j it generates the bunch's
// new ions (IonGenerator, src/LatticeInterActionPoint.cpp:985-1019), takes
// the cloud moments (IonRMSCal, 1113-1155), kicks the ions with the bunch's
// field and the electrons with the cloud's (src/MPBunch.cpp:1094-1210), and
// drifts the ions to the next bunch and removes the lost ones
// (src/LatticeInterActionPoint.cpp:1158-1210).
//
// An ion's motion depends only on the bunches' centroids and sizes, which
// are fixed for the whole phase (taken before it, src/MPBeam.cpp:1061-1063),
// not on the other ions.  So one launch moves every ion through all the
// passages of the turn, one ion per lane (ions born at passage j start
// there), with the same operations in the same order as passage by
// passage; the cloud moments of each passage are summed on the way.  The
// sums use 128-bit fixed point (2^-100 units): integer addition is
// associative, so the result does not depend on the order in which waves
// finish.  A second small kernel turns them into each passage's moments and
// electron-side kick parameters, in passage order.
#include "cts_internal.hpp"

namespace cts {

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

// v * 2^100 as three signed limbs of < 2^42 each,
// v = (l2 2^84 + l1 2^42 + l0) 2^-100, for |v| < 2^26.  Limb sums of up to
// 2^21 terms stay inside 64 bits, so the limbs are added with plain
// (non-returning) integer atomics and no carries.
__device__ __forceinline__ long long fx_limb(double v, int k)
{
    const double a = fabs(v);
    const double t2 = floor(ldexp(a, 16));
    const double r2 = a - ldexp(t2, -16);  // exact (Sterbenz)
    const double t1 = floor(ldexp(r2, 58));
    const double r1 = r2 - ldexp(t1, -58);  // exact
    const double t = k == 2 ? t2 : (k == 1 ? t1 : rint(ldexp(r1, 100)));
    const long long m = (long long)t;
    return v < 0 ? -m : m;
}

__device__ __forceinline__ double fx_value(const unsigned long long *acc)
{
    return ldexp((double)(long long)acc[2], -16) + ldexp((double)(long long)acc[1], -58) +
           ldexp((double)(long long)acc[0], -100);
}

// grid = (waves / 4, nspec), four independent waves per block.  Wave w <
// w_old holds the ions of slots [64 w, 64 w + 64) from passage 0; the other
// waves hold new ions of passage j0 (slot len + j0 n_new + t).  The lanes
// of a wave visit the passages in step; a lane drops out when its ion is
// lost.  SUR: the bunch's field at the ion from the surrogate inside its
// domain, the exact field elsewhere.
template <bool SUR>
__global__ void __launch_bounds__(256)
k_ion_turn(IonTurnParams P)
{
    __shared__ __attribute__((aligned(16))) MlpLds L;
    const int p = blockIdx.y;
    const int64_t len = P.len[p];
    {
        // blocks whose four waves all have nothing to do leave before the
        // (block-wide) weight load: the grid covers a bound of the old ions
        const int wb = blockIdx.x * 4;
        if (wb >= P.w_old + P.nb * P.w_new) return;
        if (wb + 3 < P.w_old && (int64_t)wb * 64 >= len) return;
    }
    if (SUR) {
        load_mlp_lds(L, P.sur_params, threadIdx.x, 256);
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const uint32_t turn = P.turn_dev ? *P.turn_dev + P.turn : P.turn;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= P.w_old + P.nb * P.w_new) return;
    if (w < P.w_old && (int64_t)w * 64 >= len) return;  // whole wave past the old ions
    int j0;
    int64_t slot;
    bool born = false;
    if (w < P.w_old) {
        j0 = 0;
        slot = (int64_t)w * 64 + lane;
    } else {
        const int k = w - P.w_old;
        j0 = k / P.w_new;
        const int t = (k % P.w_new) * 64 + lane;
        slot = len + (int64_t)j0 * P.n_new + t;
        born = t < P.n_new;
        if (slot >= P.cap) born = false;  // past the capacity: only after an overflow
    }
    double *X = P.x[p], *Y = P.y[p], *VX = P.vx[p], *VY = P.vy[p];
    uint8_t *A = P.alive[p];
    bool alive;
    double x = 0.0, y = 0.0, vx = 0.0, vy = 0.0;
    if (w < P.w_old) {
        alive = slot < len && A[slot];
        if (alive) { x = X[slot]; y = Y[slot]; vx = VX[slot]; vy = VY[slot]; }
    } else {
        alive = born;
        if (alive) {
            const double *bv = P.bview + 4 * j0;
            ion_generate(P.seed, turn, (uint32_t)j0, p, (uint32_t)(slot - len - (int64_t)j0 * P.n_new), bv[0],
                         bv[1], bv[2], bv[3], x, y);
        }
    }
    // loss boundary: min(ionLossBoundary * largest bunch size, aperture)
    const double lbx = P.loss_boundary * P.size_max[0], lby = P.loss_boundary * P.size_max[1];
    const double bx = P.ap_x > lbx ? lbx : P.ap_x;
    const double by = P.ap_y > lby ? lby : P.ap_y;
    unsigned long long *acc = P.acc + (int64_t)p * P.nb * kIonAcc * 3;
    for (int j = j0; j < P.nb; j++) {
        if (!__any(alive)) break;
        const double *bv = P.bview + 4 * j;
        const double xm = bv[0], ym = bv[1], sx = bv[2], sy = bv[3];
        const double dx = x - xm, dy = y - ym;
        double fx = 0.0, fy = 0.0;
        if (SUR) {
            const bool ok = sur_field(L, P.sur_norm, alive, dx, dy, sx, sy, fx, fy);
            if (alive && !ok) ion_field(dx, dy, sx, sy, fx, fy);
        } else if (alive) {
            ion_field(dx, dy, sx, sy, fx, fy);
        }
        bool keep = false;
        if (alive) {
            const double cI = P.coeffI[j * CTS_MAX_SPECIES + p];
            vx += cI * fx;
            vy += cI * fy;
            x += vx * P.dt[j];
            y += vy * P.dt[j];
            keep = !(fabs(x) > bx || fabs(y) > by);
        }
        // the cloud as the bunch finds it (before the kick), and the ions kept
        double c[kIonAcc];
        c[0] = alive ? 1.0 : 0.0;
        c[1] = alive ? dx : 0.0;
        c[2] = alive ? dy : 0.0;
        c[3] = alive ? dx * dx : 0.0;
        c[4] = alive ? dy * dy : 0.0;
        c[5] = keep ? 1.0 : 0.0;
#pragma unroll
        for (int v = 0; v < kIonAcc; v++) c[v] = wave_sum(c[v]);
        // lane 3 v + k adds limb k of value v
        if (lane < 3 * kIonAcc) {
            const int v = lane / 3, k = lane - 3 * v;
            double cv = c[0];
#pragma unroll
            for (int u = 1; u < kIonAcc; u++) cv = v == u ? c[u] : cv;
            const long long l = fx_limb(cv, k);
            if (l != 0) atomicAdd(acc + ((int64_t)j * kIonAcc + v) * 3 + k, (unsigned long long)l);
        }
        alive = keep;
    }
    const bool mine = (w < P.w_old) ? slot < len : born;
    if (mine) {
        X[slot] = x; Y[slot] = y; VX[slot] = vx; VY[slot] = vy;
        A[slot] = alive ? 1 : 0;
    }
}

// The passage sums converted in parallel, then passage by passage in order
// (one thread per species): cloud moments with
// the reference's one-ion and empty-cloud cases
// (src/LatticeInterActionPoint.cpp:1113-1155), the ion number
// (src/MPBunch.cpp:1172), the electron-side kick parameters, and the
// capacity test of IonsUpdate (src/LatticeInterActionPoint.cpp:1041-1049).
__global__ void __launch_bounds__(256)
k_ion_finish(IonFinishParams P)
{
    extern __shared__ double sums[];  // [nb][kIonAcc]
    for (int p = 0; p < P.nspec; p++) {
        const unsigned long long *acc = P.acc + (int64_t)p * P.nb * kIonAcc * 3;
        for (int i = threadIdx.x; i < P.nb * kIonAcc; i += blockDim.x) sums[i] = fx_value(acc + (int64_t)i * 3);
        __syncthreads();
        if (threadIdx.x == 0) {
            double mx = P.mean_state[2 * p], my = P.mean_state[2 * p + 1];
            double before = (double)P.len[p];
            bool overflow = false;
            for (int j = 0; j < P.nb; j++) {
                const double *a = sums + j * kIonAcc;
                const double n = a[0], s1 = a[1], s2 = a[2], q1 = a[3], q2 = a[4];
                const double xm = P.bview[4 * j], ym = P.bview[4 * j + 1];
                double rx = 0.0, ry = 0.0;
                if (n >= 2) {
                    const double ax = s1 / n, ay = s2 / n;
                    mx = xm + ax;
                    my = ym + ay;
                    rx = sqrt(fmax(q1 / n - ax * ax, 0.0));
                    ry = sqrt(fmax(q2 / n - ay * ay, 0.0));
                } else if (n == 1) {
                    // one ion: uncentred rms, the mean keeps its last value
                    rx = fabs(xm + s1);
                    ry = fabs(ym + s2);
                }
                const double nI0 = P.ion_q[j * P.nspec + p] * n;
                double *e = P.ion_e + ((int64_t)j * CTS_MAX_SPECIES + p) * 5;
                e[0] = 2.0 * nI0 * kRe / P.gamma;
                e[1] = mx;
                e[2] = my;
                e[3] = rx;
                e[4] = ry;
                double *c = P.cloud + ((int64_t)j * CTS_MAX_SPECIES + p) * 6;
                c[0] = n; c[1] = mx; c[2] = my; c[3] = rx; c[4] = ry; c[5] = nI0;
                if (before + P.n_new > (double)P.max_ions) overflow = true;
                before = a[5];
            }
            P.mean_state[2 * p] = mx;
            P.mean_state[2 * p + 1] = my;
            P.n_alive[p] = (int64_t)before;
            // slots written by the turn (old ions, then every passage's new
            // ones), the range the compaction reads
            const int64_t used = P.len[p] + (int64_t)P.nb * P.n_new;
            P.len[p] = used < P.cap ? used : P.cap;
            if (overflow) atomicOr(P.status, 2u);
        }
        __syncthreads();
    }
}

// Bunch views for the ion phase: centroid and rms sizes of every bunch
// from the last moment record (rmsRx = sqrt(emitX * betaX[k]),
// src/MPBunch.cpp:481-482), and their maxima (MPGetBeamInfo,
// src/MPBeam.cpp:2203-2258).
__global__ void __launch_bounds__(256)
k_ion_prep(const cts_moments *cur, int nb, double beta_x, double beta_y, double *bview, double *size_max,
           uint32_t *fb_count, unsigned long long *acc, int64_t nacc)
{
    if (threadIdx.x == 0 && fb_count) *fb_count = 0u;
    for (int64_t i = threadIdx.x; i < nacc; i += blockDim.x) acc[i] = 0ull;
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
hipError_t upload_faddeeva_tables_ions(const FadTables &t)
{
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_ea2n2), t.ea2n2, 64 * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_fad), t.fad, 3 * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_weid), t.weid, (kWeidN + 1) * sizeof(double));
    return e;
}

hipError_t launch_ion_turn(const IonTurnParams &p, int nspec, hipStream_t st)
{
    const int64_t waves = (int64_t)p.w_old + (int64_t)p.nb * p.w_new;
    const dim3 grid((unsigned)((waves + 3) / 4), (unsigned)nspec);
    if (p.sur_params)
        hipLaunchKernelGGL(k_ion_turn<true>, grid, dim3(256), 0, st, p);
    else
        hipLaunchKernelGGL(k_ion_turn<false>, grid, dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_ion_finish(const IonFinishParams &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_ion_finish, dim3(1), dim3(256), (size_t)p.nb * kIonAcc * sizeof(double), st, p);
    return hipGetLastError();
}

hipError_t launch_ion_prep(const cts_moments *cur, int nb, double bx, double by, double *bview, double *size_max,
                           uint32_t *fb_count, unsigned long long *acc, int64_t nacc, hipStream_t st)
{
    hipLaunchKernelGGL(k_ion_prep, dim3(1), dim3(256), 0, st, cur, nb, bx, by, bview, size_max, fb_count, acc,
                       nacc);
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
