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

// The slot of lane `lane` of wave w of a segment: waves below w_old hold
// the compacted ions [64 w, 64 w + 64) (below len); the others the new ions
// t of passage j_begin + q0, at slot len + q0 n_new + t (t below n_new;
// past the capacity only after an overflow).
__device__ __forceinline__ bool ion_wave_slot(int w, int w_old, int64_t len, int n_new, int w_new, int64_t cap,
                                              int lane, int64_t &slot, int &q0, int &t)
{
    if (w < w_old) {
        q0 = 0;
        t = 0;
        slot = (int64_t)w * 64 + lane;
        return slot < len;
    }
    const int k = w - w_old;
    q0 = k / w_new;
    t = (k - q0 * w_new) * 64 + lane;
    slot = len + (int64_t)q0 * n_new + t;
    return t < n_new && slot < cap;
}

// grid = (blocks, nspec); a block takes a cohort of kIonCohort consecutive
// waves of the segment at a time (striding over the grid).  Wave w <
// w_old holds the compacted ions of slots [64 w, 64 w + 64), from the
// segment's first passage; the other waves hold the new ions of passage j0
// (slot len + (j0 - j_begin) n_new + t), which join at their passage.  The
// cohort walks the passages in step; after each passage, when the ions
// still alive in the waves already walking fit into fewer waves, they are
// packed (in wave and lane order) into the first of them through LDS, so
// the lanes that evaluate the field stay mostly occupied as ions are lost.
// Every ion sees the same operations in the same order as passage by
// passage.  The cloud moments of each passage are summed per wave with DPP
// moves and ballots into LDS, added over the cohort's waves in wave order
// and turned into fixed-point limbs once per cohort (integer atomics, so
// the sums do not depend on the order in which blocks finish); each block
// adds its sums to one copy of the global sums at the end.  SUR: the
// bunch's field at the ion from the surrogate inside its domain, the exact
// field elsewhere.
template <bool SUR>
__global__ void __launch_bounds__(kIonBlock, kIonWavesPerEU)
k_ion_turn(IonTurnParams P)
{
    __shared__ __attribute__((aligned(16))) MlpLds L;
    __shared__ double sbv[kIonSegment][4], scI[kIonSegment], sdt[kIonSegment];
    __shared__ BeField sF[kIonSegment];
    __shared__ unsigned long long sacc[kIonSegment * kIonAcc * 3];
    __shared__ double wpart[kIonCohort][kIonSegment][kIonAcc];  // per wave, per passage
    __shared__ double rp[4][kIonBlock];                          // packing: x, y, vx, vy
    __shared__ int32_t rs[kIonBlock], ro[kIonBlock];             // packing: slot, first wave
    __shared__ int32_t wn[kIonCohort];                           // alive ions of each wave (-1: not walking)
    __shared__ uint32_t wkept[kIonCohort];                       // survivors of each first wave
    __shared__ int32_t s_qmin;
    const int p = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t len = P.len[p];
    // waves: the ions alive at the start of the segment (compacted), then
    // the new ions of every passage of the segment
    const int w_old = (int)((len + 63) / 64);
    const int nseg = P.j_end - P.j_begin;
    const int total = w_old + nseg * P.w_new;
    const int ncoh = (total + kIonCohort - 1) / kIonCohort;
    if ((int)blockIdx.x >= ncoh) return;  // whole block idle (uniform)
    if (SUR) load_mlp_lds(L, P.sur_params, tid, kIonBlock);
    for (int i = tid; i < nseg * 4; i += kIonBlock) sbv[i >> 2][i & 3] = P.bview[4 * P.j_begin + i];
    for (int i = tid; i < nseg; i += kIonBlock) {
        scI[i] = P.coeffI[(P.j_begin + i) * CTS_MAX_SPECIES + p];
        sdt[i] = P.dt[P.j_begin + i];
        sF[i] = be_plan(P.bview[4 * (P.j_begin + i) + 2], P.bview[4 * (P.j_begin + i) + 3]);
    }
    for (int i = tid; i < nseg * kIonAcc * 3; i += kIonBlock) sacc[i] = 0ull;
    const uint32_t turn = P.turn_dev ? *P.turn_dev + P.turn : P.turn;
    double *X = P.x[p], *Y = P.y[p], *VX = P.vx[p], *VY = P.vy[p];
    uint8_t *A = P.alive[p];
    // loss boundary: min(ionLossBoundary * largest bunch size, aperture)
    const double lbx = P.loss_boundary * P.size_max[0], lby = P.loss_boundary * P.size_max[1];
    const double bx = P.ap_x > lbx ? lbx : P.ap_x;
    const double by = P.ap_y > lby ? lby : P.ap_y;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int c = blockIdx.x; c < ncoh; c += gridDim.x) {
        const int w = c * kIonCohort + wv;
        // slot: this lane's ion in the buffer (-1: none); q0: the wave's
        // first passage, counted from j_begin (nseg: no ions)
        int q0 = nseg, t = 0;
        int64_t slot = -1;
        bool mine = false;
        if (w < total) {
            int qq;
            mine = ion_wave_slot(w, w_old, len, P.n_new, P.w_new, P.cap, lane, slot, qq, t);
            q0 = qq;
        }
        double x = 0.0, y = 0.0, vx = 0.0, vy = 0.0;
        if (mine) {
            if (w < w_old) {
                x = X[slot]; y = Y[slot]; vx = VX[slot]; vy = VY[slot];
            } else {
                ion_generate(P.seed, turn, (uint32_t)(P.j_begin + q0), p, (uint32_t)t, sbv[q0][0], sbv[q0][1],
                             sbv[q0][2], sbv[q0][3], x, y);
            }
            A[slot] = 0;  // survivors set it again at the end
        }
        int32_t myslot = mine ? (int32_t)slot : -1;
        int32_t first = wv;  // the wave the ion started in
        bool alive = false;  // walking and not lost
        if (tid == 0) s_qmin = nseg;
        for (int i = tid; i < kIonCohort * kIonSegment * kIonAcc; i += kIonBlock) (&wpart[0][0][0])[i] = 0.0;
        if (tid < kIonCohort) wkept[tid] = 0u;
        __syncthreads();
        const bool any_mine = __any(mine);
        if (lane == 0 && any_mine) atomicMin(&s_qmin, q0);
        __syncthreads();
        for (int q = s_qmin; q < nseg; q++) {
            // a wave's new ions join at their passage (wave-uniform)
            if (q == q0) alive = myslot >= 0;
            const bool walking = q0 <= q;
            if (__any(alive)) {
                const double xm = sbv[q][0], ym = sbv[q][1];
                const double dx = x - xm, dy = y - ym;
                double fx = 0.0, fy = 0.0;
                if (SUR) {
                    const bool ok = sur_field(L, P.sur_norm, alive, dx, dy, sbv[q][2], sbv[q][3], fx, fy);
                    if (alive && !ok) be_field(dx, dy, sF[q], fx, fy);
                } else if (alive) {
                    be_field(dx, dy, sF[q], fx, fy);
                }
                bool keep = false;
                if (alive) {
                    const double cI = scI[q];
                    vx += cI * fx;
                    vy += cI * fy;
                    x += vx * sdt[q];
                    y += vy * sdt[q];
                    keep = !(fabs(x) > bx || fabs(y) > by);
                }
                // the cloud as the bunch finds it (before the kick), and the
                // ions kept: counts by ballot, sums by DPP reductions
                const double c0 = (double)__popcll(__ballot(alive));
                double cs[4] = {alive ? dx : 0.0, alive ? dy : 0.0, alive ? dx * dx : 0.0, alive ? dy * dy : 0.0};
                wave_sum4_dpp(cs);
                const double c1 = cs[0], c2 = cs[1], c3 = cs[2], c4 = cs[3];
                const double c5 = (double)__popcll(__ballot(keep));
                if (lane == 0) {
                    double *o = wpart[wv][q];
                    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3; o[4] = c4; o[5] = c5;
                }
                alive = keep;
            }
            // pack the walking waves' ions into as few of them as hold them
            const uint64_t am = __ballot(alive);
            if (lane == 0) wn[wv] = walking ? __popcll(am) : -1;
            __syncthreads();
            int nalive = 0, busy = 0, base = 0, rank = 0;
            for (int k = 0; k < kIonCohort; k++) {
                const int n = wn[k];
                if (n < 0) continue;
                if (k < wv) { base += n; rank++; }
                nalive += n;
                busy += n > 0;
            }
            if ((nalive + 63) / 64 < busy) {  // block-uniform
                if (walking && alive) {
                    const int d = base + __popcll(am & below);
                    rp[0][d] = x; rp[1][d] = y; rp[2][d] = vx; rp[3][d] = vy;
                    rs[d] = myslot;
                    ro[d] = first;
                }
                __syncthreads();
                if (walking) {
                    const int d = rank * 64 + lane;
                    alive = d < nalive;
                    if (alive) {
                        x = rp[0][d]; y = rp[1][d]; vx = rp[2][d]; vy = rp[3][d];
                        myslot = rs[d];
                        first = ro[d];
                    } else {
                        myslot = -1;
                    }
                }
            }
            __syncthreads();
        }
        if (alive) {
            X[myslot] = x; Y[myslot] = y; VX[myslot] = vx; VY[myslot] = vy;
            A[myslot] = 1;
            atomicAdd(&wkept[first], 1u);
        }
        __syncthreads();
        if (tid < kIonCohort && c * kIonCohort + tid < total) P.wcnt[p * P.wcnt_stride + c * kIonCohort + tid] = wkept[tid];
        // the cohort's passage sums: over its waves in wave order, then to
        // fixed point (limb k of value v of passage q per thread)
        for (int i = tid; i < nseg * kIonAcc * 3; i += kIonBlock) {
            const int q = i / (kIonAcc * 3), v = (i / 3) % kIonAcc, kk = i % 3;
            double sv = 0.0;
            for (int k = 0; k < kIonCohort; k++) sv += wpart[k][q][v];
            const long long l = fx_limb(sv, kk);
            if (l != 0) atomicAdd(&sacc[(q * kIonAcc + v) * 3 + kk], (unsigned long long)l);
        }
        __syncthreads();
    }
    unsigned long long *acc = P.acc + (int64_t)(blockIdx.x % kIonReplicas) * P.acc_stride +
                              ((int64_t)p * P.nb + P.j_begin) * kIonAcc * 3;
    for (int i = tid; i < nseg * kIonAcc * 3; i += kIonBlock) {
        const unsigned long long v = sacc[i];
        if (v != 0ull) atomicAdd(acc + i, v);
    }
}

// Cloud moments and electron-side parameters of every passage, one block
// per species, a thread per passage: the cloud moments with the
// reference's one-ion and empty-cloud cases
// (src/LatticeInterActionPoint.cpp:1113-1155; with fewer than two ions the
// mean keeps the last passage's value, from the latest passage with two or
// more, or from the previous turn), the ion number (src/MPBunch.cpp:1172),
// the electron-side kick parameters, and the capacity test of IonsUpdate
// (src/LatticeInterActionPoint.cpp:1041-1049) against the ions present
// before each passage.
__device__ __forceinline__ double ion_sum(const unsigned long long *acc, int j, int v)
{
    return fx_value(acc + ((int64_t)j * kIonAcc + v) * 3);
}

__global__ void __launch_bounds__(256)
k_ion_finish(IonFinishParams P)
{
    __shared__ double last[2];
    const int p = blockIdx.x;
    const unsigned long long *acc = P.acc + (int64_t)p * P.nb * kIonAcc * 3;
    const double mx0 = P.mean_state[2 * p], my0 = P.mean_state[2 * p + 1];
    for (int j = threadIdx.x; j < P.nb; j += blockDim.x) {
        const double n = ion_sum(acc, j, 0), s1 = ion_sum(acc, j, 1), s2 = ion_sum(acc, j, 2);
        const double q1 = ion_sum(acc, j, 3), q2 = ion_sum(acc, j, 4);
        const double xm = P.bview[4 * j], ym = P.bview[4 * j + 1];
        double mx = mx0, my = my0;
        for (int k = j; k >= 0; k--) {
            const double nk = k == j ? n : ion_sum(acc, k, 0);
            if (nk >= 2) {
                const double sx = k == j ? s1 : ion_sum(acc, k, 1), sy = k == j ? s2 : ion_sum(acc, k, 2);
                mx = P.bview[4 * k] + sx / nk;
                my = P.bview[4 * k + 1] + sy / nk;
                break;
            }
        }
        double rx = 0.0, ry = 0.0;
        if (n >= 2) {
            const double ax = s1 / n, ay = s2 / n;
            rx = sqrt(fmax(q1 / n - ax * ax, 0.0));
            ry = sqrt(fmax(q2 / n - ay * ay, 0.0));
        } else if (n == 1) {
            // one ion: uncentred rms
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
        const double before = j == 0 ? (double)P.len0[p] : ion_sum(acc, j - 1, 5);
        if (before + P.n_new > (double)P.max_ions) atomicOr(P.status, 2u);
        if (j == P.nb - 1) {
            last[0] = mx;
            last[1] = my;
            P.n_alive[p] = (int64_t)ion_sum(acc, j, 5);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && P.nb > 0) {
        P.mean_state[2 * p] = last[0];
        P.mean_state[2 * p + 1] = last[1];
    }
}

// Bunch views for the ion phase: centroid and rms sizes of every bunch
// from the last moment record (rmsRx = sqrt(emitX * betaX[k]),
// src/MPBunch.cpp:481-482), and their maxima (MPGetBeamInfo,
// src/MPBeam.cpp:2203-2258).
__global__ void __launch_bounds__(256)
k_ion_prep(const cts_moments *cur, int nb, double beta_x, double beta_y, double *bview, double *size_max,
           uint32_t *fb_count, const int64_t *len, int64_t *len0, int nspec)
{
    if (threadIdx.x == 0 && fb_count) *fb_count = 0u;
    if ((int)threadIdx.x < nspec) len0[threadIdx.x] = len[threadIdx.x];
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

__global__ void __launch_bounds__(256)
k_ion_reduce(unsigned long long *acc, int64_t n, unsigned long long *out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    unsigned long long s = 0ull;  // modulo 2^64, as the atomics add
    for (int r = 0; r < kIonReplicas; r++) {
        s += acc[r * n + i];
        acc[r * n + i] = 0ull;
    }
    out[i] = s;
}

// Stable compaction of one species' cloud into the other buffer after a
// segment.  Pass 1 (one block): exclusive prefix sums of the survivors of
// the segment's waves.  Pass 2: every wave moves its survivors to their
// places, in the waves' order (old ions, then each passage's new ones), so
// the slot order is kept.
__global__ void __launch_bounds__(1024)
k_compact_scan(IonCompactParams P)
{
    __shared__ uint32_t s[1024];
    const int64_t len = *P.len;
    const int w_old = (int)((len + 63) / 64);
    const int total = w_old + P.nseg * P.w_new;
    const int per = (total + 1023) / 1024;
    const int a = min((int)threadIdx.x * per, total), b = min(a + per, total);
    uint32_t c = 0;
    for (int w = a; w < b; w++) c += P.wcnt[w];
    s[threadIdx.x] = c;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t v = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - c;
    for (int w = a; w < b; w++) {
        const uint32_t n = P.wcnt[w];
        P.wofs[w] = run;
        run += n;
    }
    if (threadIdx.x == 1023) *P.new_len = (int64_t)s[1023];
}

__global__ void __launch_bounds__(256)
k_compact_scatter(IonCompactParams P)
{
    const int64_t len = *P.len;
    const int w_old = (int)((len + 63) / 64);
    const int total = w_old + P.nseg * P.w_new;
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int w = blockIdx.x * 4 + (threadIdx.x >> 6); w < total; w += gridDim.x * 4) {
        int q0, t;
        int64_t i;
        const bool mine = ion_wave_slot(w, w_old, len, P.n_new, P.w_new, P.cap, lane, i, q0, t);
        const bool on = mine && P.alive[i];
        const uint64_t m = __ballot(on);
        if (on) {
            const uint64_t d = P.wofs[w] + (uint64_t)__popcll(m & below);
            P.ox[d] = P.x[i]; P.oy[d] = P.y[i]; P.ovx[d] = P.vx[i]; P.ovy[d] = P.vy[i];
            P.oalive[d] = 1;
        }
    }
}

// this translation unit's copy of the Faddeeva tables
hipError_t upload_faddeeva_tables_ions(const FadTables &t)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(c_weid), t.weid, (kWeidN + 1) * sizeof(double));
}

// wave_bound: a host bound of the segment's waves; the kernel reads the
// actual count from the device and strides over it.
hipError_t launch_ion_turn(const IonTurnParams &p, int nspec, int64_t wave_bound, hipStream_t st)
{
    int64_t blocks = (wave_bound + kIonCohort - 1) / kIonCohort;
    if (blocks > kIonTurnMaxBlocks) blocks = kIonTurnMaxBlocks;
    if (blocks < 1) blocks = 1;
    const dim3 grid((unsigned)blocks, (unsigned)nspec);
    if (p.sur_params)
        hipLaunchKernelGGL(k_ion_turn<true>, grid, dim3(kIonBlock), 0, st, p);
    else
        hipLaunchKernelGGL(k_ion_turn<false>, grid, dim3(kIonBlock), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_ion_finish(const IonFinishParams &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_ion_finish, dim3(p.nspec), dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_ion_prep(const cts_moments *cur, int nb, double bx, double by, double *bview, double *size_max,
                           uint32_t *fb_count, const int64_t *len, int64_t *len0, int nspec, hipStream_t st)
{
    hipLaunchKernelGGL(k_ion_prep, dim3(1), dim3(256), 0, st, cur, nb, bx, by, bview, size_max, fb_count, len,
                       len0, nspec);
    return hipGetLastError();
}

hipError_t launch_ion_reduce(unsigned long long *acc, int64_t n, unsigned long long *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_ion_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, acc, n, out);
    return hipGetLastError();
}

hipError_t launch_ion_compact(const IonCompactParams &p, int64_t wave_bound, hipStream_t st)
{
    int64_t blocks = (wave_bound + 3) / 4;
    if (blocks > kIonTurnMaxBlocks) blocks = kIonTurnMaxBlocks;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(1024), 0, st, p);
    hipLaunchKernelGGL(k_compact_scatter, dim3((unsigned)blocks), dim3(256), 0, st, p);
    return hipGetLastError();
}

}  // namespace cts
