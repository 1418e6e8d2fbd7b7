// cts_kernels.hip -- HIP kernels of the multi-particle tracking turn (gfx950).
//
// Data layout in HBM: FP64 structure of arrays x, px, y, py, z, dp for all
// bunches concatenated, one uint8 flag per macroparticle (bits 0-1: the
// reference's loss code eSurive, bit 7: "lost during the last marking"),
// and a segment table that cuts the particle range into contiguous
// per-workgroup chunks intersected with bunch boundaries.  Every workgroup
// owns one chunk; per-segment partial moments are written without atomics
// and reduced per bunch in a fixed order, so moments are reproducible run to
// run.  Longitudinal histograms are LDS-private per workgroup and merged
// into the per-bunch histogram with integer atomics (exact, order-free).
#include "cts_internal.hpp"

namespace cts {

__constant__ double c_ea2n2[64];
__constant__ double c_fad[3];

hipError_t upload_faddeeva_tables(const double ea2n2[64], const double fad[3])
{
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_ea2n2), ea2n2, 64 * sizeof(double));
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_fad), fad, 3 * sizeof(double));
}

// --------------------------------------------------------------- the pass
// One streaming pass over the particles; `ops` selects the stages, in the
// order of MPBeam::Run (src/MPBeam.cpp:1059-1102):
//   KICK   dp += wake kick of the previous binning (deferred SR wake kick,
//          src/MPBunch.cpp:2015-2025), survivors at kick time only
//   MOMB   moments of z/dp after the kick (completes the previous turn)
//   ION    electron kick by the ion cloud (src/Bunch.cpp:213-220)
//   MAP    section map (src/Bunch.cpp:625-755)
//   ZSTAT  post-map z moments and survivors' z range (SR wake edges)
//   SKEW   thin skew quad (src/Bunch.cpp:757-795)
//   RF/DRIFT/U0  (src/Bunch.cpp:453-501)
//   SYNRAD (src/Bunch.cpp:940-1044)
//   BIN    longitudinal histogram (src/MPBunch.cpp:1915-1919)
//   LOSS   loss marking (src/Bunch.cpp:172-200)
//   MOMA   moments after marking (src/MPBunch.cpp:342-498)
// M is the compile-time set of stages a variant may run; stages outside it
// are compiled out, which keeps the register budget of the fused turn small.
__host__ __device__ constexpr bool acc_used(uint32_t M, int j)
{
    return (j <= A_ZMAX && j != A_LOSS) ? (M & OP_MOMA) != 0
         : j == A_LOSS                  ? (M & OP_LOSS) != 0
         : j <= B_QZD                   ? (M & (OP_MOMB | OP_MOMD)) != 0
                                        : (M & OP_ZSTAT) != 0;
}

template <uint32_t M>
__global__ void __launch_bounds__(kBlock)
k_pass(PassParams P)
{
    extern __shared__ double smem_d[];
    double *kick_lds = smem_d;                                       // nbins
    uint32_t *hist_lds = reinterpret_cast<uint32_t *>(smem_d + P.nbins);  // nbins
    __shared__ double red[kBlock / 64][kNPart];

    const uint32_t ops = P.ops & M;
    const int tid = threadIdx.x;
    const int s_first = P.wg_seg[blockIdx.x], s_last = P.wg_seg[blockIdx.x + 1];

    for (int s = s_first; s < s_last; s++) {
        const int b = P.seg_bunch[s];
        const int64_t i0 = P.seg_begin[s], i1 = P.seg_end[s];
        const int64_t bunch_first = P.bunch_off[b];
        double sh[6];
#pragma unroll
        for (int j = 0; j < 6; j++) sh[j] = P.shift[b * 6 + j];

        const bool use_hist = (ops & OP_BIN) && P.nbins > 0;
        const bool use_kick = (ops & OP_KICK) && P.nbins > 0;
        if (use_kick)
            for (int j = tid; j < P.nbins; j += kBlock) kick_lds[j] = P.kick_prev[(int64_t)b * P.nbins + j];
        if (use_hist)
            for (int j = tid; j < P.nbins; j += kBlock) hist_lds[j] = 0u;
        double klo = 0, kdz = 1, blo = 0, bdz = 1;
        if (use_kick) { klo = P.edges_prev[2 * b]; kdz = P.edges_prev[2 * b + 1]; }
        if (use_hist) { blo = P.edges_now[2 * b]; bdz = P.edges_now[2 * b + 1]; }
        __syncthreads();

        double acc[kNPart];
#pragma unroll
        for (int j = 0; j < kNPart; j++) acc[j] = 0.0;
        acc[A_ZMIN] = INFINITY; acc[A_ZMAX] = -INFINITY;
        acc[C_ZMIN] = INFINITY; acc[C_ZMAX] = -INFINITY;

        for (int64_t i = i0 + tid; i < i1; i += kBlock) {
            double X[6];
            X[0] = P.x[i]; X[1] = P.px[i]; X[2] = P.y[i];
            X[3] = P.py[i]; X[4] = P.z[i]; X[5] = P.dp[i];
            uint32_t f = P.flag[i];

            if (use_kick) {
                // survivors when the kick was computed: code 0, or lost at
                // the marking that followed it
                if ((f & 3u) == 0u || (f & kNewLost)) {
                    const double q = (X[4] - klo) / kdz;
                    if (q > -1.0 && q < (double)P.nbins) X[5] += kick_lds[(int)q];
                }
            }
            f &= 3u;
            const bool alive = f == 0u;
            if ((ops & OP_MOMB) && alive) {
                const double dz = X[4] - sh[4], dd = X[5] - sh[5];
                acc[B_N] += 1.0;
                acc[B_SZ] += dz;
                acc[B_SD] += dd;
                acc[B_QDD] += dd * dd;
                acc[B_QZD] += dz * dd;
            }
            if ((ops & OP_MAP) && alive) {
                const double d = X[5];
                const double xx = X[0] - P.eta_x * d, pxx = X[1] - P.etap_x * d;
                const double yy = X[2] - P.eta_y * d, pyy = X[3] - P.etap_y * d;
                const double tx = P.alpha_x * xx + P.beta_x * pxx;
                const double ty = P.alpha_y * yy + P.beta_y * pyy;
                const double ampX = (xx * xx + tx * tx) / P.beta_x;
                const double ampY = (yy * yy + ty * ty) / P.beta_y;
                const double dnux = (P.chrom_x * d + P.adtx[0] * ampX + P.adtx[1] * ampY +
                                     P.adtx[2] * (ampX * ampX) / 2 + P.adtx[3] * (ampY * ampY) / 2 +
                                     P.adtx[4] * ampX * ampY) * P.weigh_x;
                const double dnuy = (P.chrom_y * d + P.adty[0] * ampX + P.adty[1] * ampY +
                                     P.adty[2] * (ampX * ampX) / 2 + P.adty[3] * (ampY * ampY) / 2 +
                                     P.adty[4] * ampX * ampY) * P.weigh_y;
                const double phiX = P.phase_x + 2 * kPi * dnux;
                const double phiY = P.phase_y + 2 * kPi * dnuy;
                double sx, cx, sy, cy;
                sincos(phiX, &sx, &cx);
                sincos(phiY, &sy, &cy);
                double Y[6];
                spmv(P.b1h1, X, Y);
                double R[6];
                R[0] = cx * Y[0] + sx * Y[1];
                R[1] = -sx * Y[0] + cx * Y[1];
                R[2] = cy * Y[2] + sy * Y[3];
                R[3] = -sy * Y[2] + cy * Y[3];
                R[4] = Y[4];
                R[5] = Y[5];
                spmv(P.h2b2, R, X);
            }
            if ((ops & OP_ZSTAT) && alive) {
                const double dz = X[4] - sh[4];
                acc[C_N] += 1.0;
                acc[C_SZ] += dz;
                acc[C_QZZ] += dz * dz;
            }
            if ((ops & OP_SKEW) && alive) {
                // identity with (1,2) = (3,0) = K, summed in column order
                const double npx = X[1] + P.skew_k * X[2];
                const double npy = P.skew_k * X[0] + X[3];
                X[1] = npx;
                X[3] = npy;
            }
            if (ops & OP_RF) {
                for (int r = 0; r < P.n_rf; r++) {
                    const double th = X[4] * P.rf_k[r];
                    double s, c;
                    sincos(th, &s, &c);
                    const double re = P.rf_vr[r] * c - P.rf_vi[r] * s;
                    X[5] += re * P.rf_scale;
                }
            }
            if (ops & OP_DRIFT) {
                const double d = X[5];
                X[4] -= P.drift_c * (P.alphac[0] * d + P.alphac[1] * (d * d) + P.alphac[2] * (d * d * d));
            }
            if (ops & OP_U0) X[5] -= P.u0_kick;
            if ((ops & OP_SYNRAD) && alive) {
                double N[6];
                spmv(P.sr_b1h1, X, N);
                N[0] *= P.sr_lambda[0];
                N[1] *= P.sr_lambda[0];
                N[2] *= P.sr_lambda[1];
                N[3] *= P.sr_lambda[1];
                N[5] *= P.sr_lz2;
                if (P.noise_mode != 0) {
                    double r[5];
                    if (P.noise_mode == 1) {
#pragma unroll
                        for (int j = 0; j < 5; j++) r[j] = P.noise[j * P.ntot + i];
                    } else {
                        qe_normals(P.seed, P.turn, (uint64_t)i, r);
                    }
                    N[0] += P.sr_coeff[0] * r[0];
                    N[1] += P.sr_coeff[0] * r[1];
                    N[2] += P.sr_coeff[1] * r[2];
                    N[3] += P.sr_coeff[1] * r[3];
                    N[5] += P.sr_coeff[2] * r[4];
                }
                spmv(P.sr_h1b1, N, X);
            }
            if ((ops & OP_ZSTAT) && (alive || i == bunch_first)) {
                // GetZMinMax starts from particle 0 of the bunch even when it
                // is lost (src/MPBunch.cpp:1862-1863)
                acc[C_ZMIN] = fmin(acc[C_ZMIN], X[4]);
                acc[C_ZMAX] = fmax(acc[C_ZMAX], X[4]);
            }
            if (use_hist) {
                const double q = (X[4] - blo) / bdz;
                if (q > -1.0 && q < (double)P.nbins) atomicAdd(&hist_lds[(int)q], 1u);
            }
            uint32_t fnew = f;
            if (ops & OP_LOSS) {
                double cnt = 0.0;
                if (fabs(X[4]) > P.zlim) { fnew = 2u; cnt += 1.0; }
                const double ax = X[0] / P.ap_x, ay = X[2] / P.ap_y;
                if (ax * ax + ay * ay > 1) { fnew = 1u; cnt += 1.0; }
                acc[A_LOSS] += cnt;
                if (f == 0u && fnew != 0u) fnew |= kNewLost;
            }
            if ((ops & OP_MOMA) && (fnew & 3u) == 0u) {
                const double d0 = X[0] - sh[0], d1 = X[1] - sh[1], d2 = X[2] - sh[2];
                const double d3 = X[3] - sh[3], d4 = X[4] - sh[4];
                acc[A_N] += 1.0;
                acc[A_S + 0] += d0; acc[A_S + 1] += d1; acc[A_S + 2] += d2;
                acc[A_S + 3] += d3; acc[A_S + 4] += d4;
                acc[A_Q + 0] += d0 * d0; acc[A_Q + 1] += d0 * d1; acc[A_Q + 2] += d1 * d1;
                acc[A_Q + 3] += d2 * d2; acc[A_Q + 4] += d2 * d3; acc[A_Q + 5] += d3 * d3;
                acc[A_Q + 6] += d4 * d4;
                acc[A_ZMIN] = fmin(acc[A_ZMIN], X[4]);
                acc[A_ZMAX] = fmax(acc[A_ZMAX], X[4]);
            }
            if ((ops & OP_MOMD) && (fnew & 3u) == 0u) {
                const double dz = X[4] - sh[4], dd = X[5] - sh[5];
                acc[B_N] += 1.0;
                acc[B_SZ] += dz;
                acc[B_SD] += dd;
                acc[B_QDD] += dd * dd;
                acc[B_QZD] += dz * dd;
            }
            if (ops & OP_WRITE) {
                P.x[i] = X[0]; P.px[i] = X[1]; P.y[i] = X[2];
                P.py[i] = X[3]; P.z[i] = X[4]; P.dp[i] = X[5];
                P.flag[i] = (uint8_t)fnew;
            }
        }

        // block reduction of the accumulators -> segment partials
        const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
        for (int j = 0; j < kNPart; j++) {
            if (!acc_used(M, j)) continue;
            double v;
            if (j == A_ZMIN || j == C_ZMIN) v = wave_min(acc[j]);
            else if (j == A_ZMAX || j == C_ZMAX) v = wave_max(acc[j]);
            else v = wave_sum(acc[j]);
            if (lane == 0) red[wave][j] = v;
        }
        __syncthreads();
        if (tid < kNPart) {
            // identity of the unused slots (0, +inf for minima, -inf for maxima)
            double v = (tid == A_ZMIN || tid == C_ZMIN) ? INFINITY
                     : (tid == A_ZMAX || tid == C_ZMAX) ? -INFINITY : 0.0;
            if (acc_used(M, tid)) {
                v = red[0][tid];
                for (int w = 1; w < kBlock / 64; w++) {
                    if (tid == A_ZMIN || tid == C_ZMIN) v = fmin(v, red[w][tid]);
                    else if (tid == A_ZMAX || tid == C_ZMAX) v = fmax(v, red[w][tid]);
                    else v = v + red[w][tid];
                }
            }
            P.seg_part[(int64_t)s * kNPart + tid] = v;
        }
        if (use_hist) {
            for (int j = tid; j < P.nbins; j += kBlock) {
                const uint32_t c = hist_lds[j];
                if (c) atomicAdd(&P.bunch_hist[(int64_t)b * P.nbins + j], c);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ binning
// Histogram pass for the reference's bin edges, which depend on the
// survivors' z range after the pass (src/MPBunch.cpp:1900-1919).
__global__ void __launch_bounds__(kBlock)
k_bin(BinParams P)
{
    extern __shared__ double smem_d[];
    uint32_t *hist_lds = reinterpret_cast<uint32_t *>(smem_d);
    const int tid = threadIdx.x;
    const int s_first = P.wg_seg[blockIdx.x], s_last = P.wg_seg[blockIdx.x + 1];
    for (int s = s_first; s < s_last; s++) {
        const int b = P.seg_bunch[s];
        const int64_t i0 = P.seg_begin[s], i1 = P.seg_end[s];
        for (int j = tid; j < P.nbins; j += kBlock) hist_lds[j] = 0u;
        const double lo = P.edges[2 * b], dz = P.edges[2 * b + 1];
        __syncthreads();
        for (int64_t i = i0 + tid; i < i1; i += kBlock) {
            const double q = (P.z[i] - lo) / dz;
            if (q > -1.0 && q < (double)P.nbins) atomicAdd(&hist_lds[(int)q], 1u);
        }
        __syncthreads();
        for (int j = tid; j < P.nbins; j += kBlock) {
            const uint32_t c = hist_lds[j];
            if (c) atomicAdd(&P.bunch_hist[(int64_t)b * P.nbins + j], c);
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------- finalize
// Per bunch: reduce segment partials in segment order, write the moment
// record, update the shift and (reference-edge mode) the bin edges.
__global__ void __launch_bounds__(kBlock)
k_finalize(FinParams P)
{
    __shared__ double red[kBlock / 64][kNPart];
    __shared__ double tot[kNPart];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int s0 = P.bunch_seg[b], s1 = P.bunch_seg[b + 1];
    double acc[kNPart];
#pragma unroll
    for (int j = 0; j < kNPart; j++) acc[j] = 0.0;
    acc[A_ZMIN] = INFINITY; acc[A_ZMAX] = -INFINITY;
    acc[C_ZMIN] = INFINITY; acc[C_ZMAX] = -INFINITY;
    for (int s = s0 + tid; s < s1; s += kBlock) {
        const double *p = P.seg_part + (int64_t)s * kNPart;
#pragma unroll
        for (int j = 0; j < kNPart; j++) {
            if (j == A_ZMIN || j == C_ZMIN) acc[j] = fmin(acc[j], p[j]);
            else if (j == A_ZMAX || j == C_ZMAX) acc[j] = fmax(acc[j], p[j]);
            else acc[j] += p[j];
        }
    }
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int j = 0; j < kNPart; j++) {
        double v;
        if (j == A_ZMIN || j == C_ZMIN) v = wave_min(acc[j]);
        else if (j == A_ZMAX || j == C_ZMAX) v = wave_max(acc[j]);
        else v = wave_sum(acc[j]);
        if (lane == 0) red[wave][j] = v;
    }
    __syncthreads();
    if (tid < kNPart) {
        double v = red[0][tid];
        for (int w = 1; w < kBlock / 64; w++) {
            if (tid == A_ZMIN || tid == C_ZMIN) v = fmin(v, red[w][tid]);
            else if (tid == A_ZMAX || tid == C_ZMAX) v = fmax(v, red[w][tid]);
            else v = v + red[w][tid];
        }
        tot[tid] = v;
    }
    __syncthreads();
    if (tid != 0) return;

    double *sh = P.shift + 6 * b;
    const double ncount = (double)P.bunch_count[b];
    cts_moments *cur = P.cur + b;
    double dmean = 0, dzd = 0, ddd = 0;
    if (P.ops & (OP_MOMB | OP_MOMD)) {
        const double n = tot[B_N];
        const double mz = tot[B_SZ] / n, md = tot[B_SD] / n;
        dmean = sh[5] + md;
        dzd = tot[B_QZD] / n - mz * md;
        ddd = tot[B_QDD] / n - md * md;
    }
    if (P.ops & OP_MOMB) {
        // completes the dp part of the record that was current
        cur->mean[5] = dmean;
        cur->c2[7] = dzd;
        cur->c2[8] = ddd;
        if (P.hist && P.prev_slot >= 0) P.hist[(int64_t)P.prev_slot * P.nb + b] = *cur;
        sh[5] = dmean;
    }
    if (P.ops & OP_MOMA) {
        const double n = tot[A_N];
        double m[5];
        for (int j = 0; j < 5; j++) m[j] = tot[A_S + j] / n;
        cur->n_alive = n;
        for (int j = 0; j < 5; j++) cur->mean[j] = sh[j] + m[j];
        cur->c2[0] = tot[A_Q + 0] / n - m[0] * m[0];
        cur->c2[1] = tot[A_Q + 1] / n - m[0] * m[1];
        cur->c2[2] = tot[A_Q + 2] / n - m[1] * m[1];
        cur->c2[3] = tot[A_Q + 3] / n - m[2] * m[2];
        cur->c2[4] = tot[A_Q + 4] / n - m[2] * m[3];
        cur->c2[5] = tot[A_Q + 5] / n - m[3] * m[3];
        cur->c2[6] = tot[A_Q + 6] / n - m[4] * m[4];
        cur->zmin = tot[A_ZMIN];
        cur->zmax = tot[A_ZMAX];
        for (int j = 0; j < 5; j++) sh[j] = cur->mean[j];
    }
    if (P.ops & OP_MOMD) {
        cur->mean[5] = dmean;
        cur->c2[7] = dzd;
        cur->c2[8] = ddd;
        sh[5] = dmean;
    }
    if (P.ops & OP_LOSS) {
        cur->loss_count = tot[A_LOSS];
        cur->transmission = (ncount - tot[A_LOSS]) / ncount;
        if (cur->transmission < 0.5) atomicOr(P.status, 1u);
    }
    if ((P.ops & (OP_MOMA | OP_MOMD | OP_LOSS)) && P.hist && P.slot >= 0)
        P.hist[(int64_t)P.slot * P.nb + b] = *cur;
    if (P.ops & OP_ZSTAT) {
        const double n = tot[C_N];
        const double mz = tot[C_SZ] / n;
        const double sig = sqrt(tot[C_QZZ] / n - mz * mz);
        P.sigz_stale[b] = sig;
        P.zrange[2 * b] = tot[C_ZMIN];
        P.zrange[2 * b + 1] = tot[C_ZMAX];
        if (P.nbins > 0 && P.ref_edges) {
            const double zMin = tot[C_ZMIN] - 1.e-6, zMax = tot[C_ZMAX] + 1.e-6;
            const double lo = zMin - 2 * sig, hi = zMax + 2 * sig;
            P.edges_now[2 * b] = lo;
            P.edges_now[2 * b + 1] = (hi - lo) / P.nbins;
        }
    }
}

// ---------------------------------------------------------- wake table
// W[l] = longitudinal BBR pseudo-Green function at tau = (l-(nbins-1))*dt
// (src/WakeFunction.cpp:558-652); one wave per lag, each lane sums a strided
// share of the 1000-step Riemann sum.  grid = (ceil(nlag*64/256), ntables).
__global__ void __launch_bounds__(kBlock)
k_wake_table(WakeTableParams P)
{
    const int nlag = 2 * P.nbins - 1;
    const int lag = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.y;
    if (lag >= nlag) return;
    const double dt = P.edges[2 * t + 1] / kClight;
    const double tau = (lag - (P.nbins - 1)) * dt;
    const double sigmat = 1.e-3 / kClight;
    const double u = tau / sigmat;
    const double sigmaz = sigmat * kClight;
    double total = 0.0;
    for (int r = 0; r < P.nres; r++) {
        const double Q = P.q[r];
        const double vz = P.omega[r] * sigmaz / kClight;
        const double zMax = 3 * vz / (2 * Q);
        const double dz = zMax / 1000;
        const double sq = sqrt(4 * Q * Q - 1);
        double sum = 0.0;
        for (int j = lane; j < 1000; j += 64) {
            const double z = j * dz;
            const double a = u + 2 * Q / vz * z;
            const double term0 = -1 / 2.0 * (a * a) - z;
            double s, c;
            sincos(z * sq, &s, &c);
            sum += exp(term0) * (c - s / sq) * dz;
        }
        sum = wave_sum(sum);
        total += (sqrt(2. / kPi) * P.rs[r] * kClight / sigmaz) * sum;
    }
    if (lane == 0) P.table[(int64_t)t * nlag + lag] = total;
}

// --------------------------------------------------------------- wake
// Per bunch: kick_i = (poten_i + sum_j W[i-j] * n_j) * (-e*q/E)
// (src/MPBunch.cpp:1983-1998); the histogram is cleared for the next pass.
__global__ void __launch_bounds__(kBlock)
k_wake(WakeParams P)
{
    extern __shared__ double smem_d[];
    double *n_lds = smem_d;
    const int b = blockIdx.x, tid = threadIdx.x;
    const int nb = P.nbins;
    uint32_t *hist = P.bunch_hist + (int64_t)b * nb;
    for (int j = tid; j < nb; j += kBlock) {
        const uint32_t c = hist[j];
        n_lds[j] = (double)c;
        P.counts_out[(int64_t)b * nb + j] = (int32_t)c;
        hist[j] = 0u;
    }
    __syncthreads();
    const double *W = P.table + (P.per_bunch_table ? (int64_t)b * (2 * nb - 1) : 0);
    const double scale = P.scale[b];
    double *poten = P.poten + (int64_t)b * nb;
    for (int i = tid; i < nb; i += kBlock) {
        double a = poten[i];
        const double *Wi = W + i + (nb - 1);
        for (int j = 0; j < nb; j++) a += Wi[-j] * n_lds[j];
        a *= scale;
        poten[i] = a;
        P.kick[(int64_t)b * nb + i] = a;
    }
    if (tid == 0) {
        P.edges_prev[2 * b] = P.edges_now[2 * b];
        P.edges_prev[2 * b + 1] = P.edges_now[2 * b + 1];
    }
}

// -------------------------------------------------------- test hooks
__global__ void k_qe_debug(uint64_t seed, uint32_t turn, uint64_t base, int64_t n, double *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r[5];
    qe_normals(seed, turn, base + (uint64_t)i, r);
#pragma unroll
    for (int j = 0; j < 5; j++) out[j * n + i] = r[j];
}

// ------------------------------------------------------- ion kick
// Electron side of MPBunch::SSIonBunchInteraction (src/MPBunch.cpp:1165-1206)
// followed by BunchTransferDueToIon (src/Bunch.cpp:213-220): survivors get
// px/py += sum over species of 2*N_ion*r_e/gamma * F(x - xbar, y - ybar).
__global__ void __launch_bounds__(kBlock)
k_ion_kick(IonKickParams P)
{
    const int tid = threadIdx.x;
    const int s_first = P.wg_seg[blockIdx.x], s_last = P.wg_seg[blockIdx.x + 1];
    for (int s = s_first; s < s_last; s++) {
        const int b = P.seg_bunch[s];
        const double *c = P.ion_e + (int64_t)b * CTS_MAX_SPECIES * 5;
        for (int64_t i = P.seg_begin[s] + tid; i < P.seg_end[s]; i += kBlock) {
            if (P.flag[i] & 3u) continue;
            const double x = P.x[i], y = P.y[i];
            double fxs = 0.0, fys = 0.0;
            for (int p = 0; p < P.nspec; p++) {
                double fx, fy;
                ion_field(x - c[5 * p + 1], y - c[5 * p + 2], c[5 * p + 3], c[5 * p + 4], fx, fy);
                fxs = fxs + c[5 * p] * fx;
                fys = fys + c[5 * p] * fy;
            }
            P.px[i] += fxs;
            P.py[i] += fys;
        }
    }
}

// ------------------------------------------------------------ launches
constexpr uint32_t V_KICK = OP_KICK | OP_MOMB | OP_WRITE;
constexpr uint32_t V_MOM = OP_MOMA | OP_MOMB | OP_MOMD | OP_ZSTAT | OP_LOSS;
constexpr uint32_t V_FIX = OP_KICK | OP_MOMB | OP_MAP | OP_SKEW | OP_RF | OP_DRIFT | OP_U0 | OP_SYNRAD |
                           OP_BIN | OP_LOSS | OP_MOMA | OP_MOMD | OP_WRITE;
constexpr uint32_t V_REF = OP_KICK | OP_MOMB | OP_MAP | OP_ZSTAT | OP_SKEW | OP_RF | OP_DRIFT | OP_U0 |
                           OP_SYNRAD | OP_LOSS | OP_MOMA | OP_MOMD | OP_WRITE;
constexpr uint32_t V_ALL = 0xffffffffu & ~OP_ION;

hipError_t launch_pass(const PassParams &p, int nwg, hipStream_t st)
{
    const size_t shm = (size_t)p.nbins * (sizeof(double) + sizeof(uint32_t));
    const uint32_t o = p.ops;
    if ((o & ~V_KICK) == 0) hipLaunchKernelGGL(k_pass<V_KICK>, dim3(nwg), dim3(kBlock), shm, st, p);
    else if ((o & ~V_MOM) == 0) hipLaunchKernelGGL(k_pass<V_MOM>, dim3(nwg), dim3(kBlock), shm, st, p);
    else if ((o & ~V_FIX) == 0) hipLaunchKernelGGL(k_pass<V_FIX>, dim3(nwg), dim3(kBlock), shm, st, p);
    else if ((o & ~V_REF) == 0) hipLaunchKernelGGL(k_pass<V_REF>, dim3(nwg), dim3(kBlock), shm, st, p);
    else hipLaunchKernelGGL(k_pass<V_ALL>, dim3(nwg), dim3(kBlock), shm, st, p);
    return hipGetLastError();
}
hipError_t launch_ion_kick(const IonKickParams &p, int nwg, hipStream_t st)
{
    hipLaunchKernelGGL(k_ion_kick, dim3(nwg), dim3(kBlock), 0, st, p);
    return hipGetLastError();
}
hipError_t launch_bin(const BinParams &p, int nwg, hipStream_t st)
{
    hipLaunchKernelGGL(k_bin, dim3(nwg), dim3(kBlock), (size_t)p.nbins * sizeof(uint32_t), st, p);
    return hipGetLastError();
}
hipError_t launch_finalize(const FinParams &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_finalize, dim3(p.nb), dim3(kBlock), 0, st, p);
    return hipGetLastError();
}
hipError_t launch_wake_table(const WakeTableParams &p, int ntables, hipStream_t st)
{
    const int nlag = 2 * p.nbins - 1;
    const int blocks = (nlag * 64 + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_wake_table, dim3(blocks, ntables), dim3(kBlock), 0, st, p);
    return hipGetLastError();
}
hipError_t launch_wake(const WakeParams &p, int nb, hipStream_t st)
{
    hipLaunchKernelGGL(k_wake, dim3(nb), dim3(kBlock), (size_t)p.nbins * sizeof(double), st, p);
    return hipGetLastError();
}
hipError_t launch_qe_debug(uint64_t seed, uint32_t turn, uint64_t base, int64_t n, double *out,
                           hipStream_t st)
{
    const int blocks = (int)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_qe_debug, dim3(blocks), dim3(kBlock), 0, st, seed, turn, base, n, out);
    return hipGetLastError();
}

}  // namespace cts
