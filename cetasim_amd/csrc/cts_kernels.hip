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
#include <algorithm>

#include "cts_internal.hpp"

namespace cts {

hipError_t upload_faddeeva_tables(const FadTables &t)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(c_weid), t.weid, (kWeidN + 1) * sizeof(double));
}

// Streaming (nontemporal) global accesses.
template <class T>
__device__ __forceinline__ T ntload(const T *p)
{
    return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void ntstore(T *p, T v)
{
    __builtin_nontemporal_store(v, p);
}

// Bin of z for edges (lo, dz) as the reference computes it,
// int((z - lo)/dz) (src/MPBunch.cpp:1917), or -1 outside [0, nbins).  The
// quotient is formed with the reciprocal and redone with the division when
// it lies within 1e-9 of an integer, where the two could truncate apart.
__device__ __forceinline__ int bin_of(double z, double lo, double dz, double inv_dz, int nbins)
{
    const double d = z - lo;
    double q = d * inv_dz;
    const double t = floor(q);
    if (q - t < 1e-9 || t + 1.0 - q < 1e-9) q = d / dz;
    return (q > -1.0 && q < (double)nbins) ? (int)q : -1;
}

// (x/ax)^2 + (y/ay)^2 > 1 (src/Bunch.cpp:186-187); the division form is
// used when the reciprocal form is within 1e-9 of the boundary.
__device__ __forceinline__ bool outside_aperture(double x, double y, cdptr K)
{
    const double fx = x * K[CI_IAPX], fy = y * K[CI_IAPY];
    double t = fx * fx + fy * fy;
    if (fabs(t - 1.0) < 1e-9) {
        const double ex = x / K[CI_APX], ey = y / K[CI_APY];
        t = ex * ex + ey * ey;
    }
    return t > 1;
}

// sin and cos for |d| <= 0.1: Taylor polynomials in d^2 (coefficients
// K[CI_PS..], K[CI_PC..]: 1/9!, -1/7!, 1/5!, -1/3! and -1/10!, 1/8!,
// -1/6!, 1/4!).
__device__ __forceinline__ void sincos_small_k(double d, double &s, double &c, cdptr K)
{
#pragma clang fp contract(fast)
    const double d2 = d * d;
    double sp = K[CI_PS];
    sp = fma_vvs(d2, sp, K[CI_PS + 1]);
    sp = fma_vvs(d2, sp, K[CI_PS + 2]);
    sp = fma_vvs(d2, sp, K[CI_PS + 3]);
    s = d + d * (d2 * sp);
    double cp = K[CI_PC];
    cp = fma_vvs(d2, cp, K[CI_PC + 1]);
    cp = fma_vvs(d2, cp, K[CI_PC + 2]);
    cp = fma_vvs(d2, cp, K[CI_PC + 3]);
    c = 1.0 - 0.5 * d2 + (d2 * d2) * cp;
}

// sin and cos of any x: reduction by pi/2 in three parts (Cody-Waite,
// K[CI_FR..]: 2/pi, pi/2 split), polynomials of degree 15/14 on the
// reduced angle (K[CI_FS..], K[CI_FC..]), quadrant selection.
__device__ __forceinline__ void sincos_fast_k(double x, double &s, double &c, cdptr K)
{
#pragma clang fp contract(fast)
    const double n = rint(x * K[CI_FR]);
    double r = x - n * K[CI_FR + 1];
    r = r - n * K[CI_FR + 2];
    r = r - n * K[CI_FR + 3];
    const double r2 = r * r;
    double sp = K[CI_FS];
    for (int i = 1; i < 7; i++) sp = fma_vvs(r2, sp, K[CI_FS + i]);
    const double sn = r + r * (r2 * sp);
    double cp = K[CI_FC];
    for (int i = 1; i < 7; i++) cp = fma_vvs(r2, cp, K[CI_FC + i]);
    const double cs = 1.0 - 0.5 * r2 - r2 * (r2 * cp);
    const int q = (int)n & 3;
    s = (q == 0) ? sn : (q == 1) ? cs : (q == 2) ? -sn : -cs;
    c = (q == 0) ? cs : (q == 1) ? -sn : (q == 2) ? -cs : sn;
}

// Section map, src/Bunch.cpp:625-755.
__device__ __forceinline__ void stage_map(double X[6], cdptr K)
{
#pragma clang fp contract(fast)
    const double d = X[5];
    const double xx = X[0] - K[CI_ETAX] * d, pxx = X[1] - K[CI_ETAPX] * d;
    const double yy = X[2] - K[CI_ETAY] * d, pyy = X[3] - K[CI_ETAPY] * d;
    const double tx = K[CI_ALPHAX] * xx + K[CI_BETAX] * pxx;
    const double ty = K[CI_ALPHAY] * yy + K[CI_BETAY] * pyy;
    const double ampX = (xx * xx + tx * tx) * K[CI_IBETAX];
    const double ampY = (yy * yy + ty * ty) * K[CI_IBETAY];
    const double dnux = (K[CI_CHX] * d + K[CI_ADTX + 0] * ampX + K[CI_ADTX + 1] * ampY +
                         K[CI_ADTX + 2] * (ampX * ampX) * 0.5 + K[CI_ADTX + 3] * (ampY * ampY) * 0.5 +
                         K[CI_ADTX + 4] * ampX * ampY) * K[CI_WX];
    const double dnuy = (K[CI_CHY] * d + K[CI_ADTY + 0] * ampX + K[CI_ADTY + 1] * ampY +
                         K[CI_ADTY + 2] * (ampX * ampX) * 0.5 + K[CI_ADTY + 3] * (ampY * ampY) * 0.5 +
                         K[CI_ADTY + 4] * ampX * ampY) * K[CI_WY];
    // phi = mu + 2*pi*dnu: for the usual small tune shift, angle addition
    // onto the section's cos/sin(mu); otherwise the full reduction
    const double ddx = K[CI_TWOPI] * dnux, ddy = K[CI_TWOPI] * dnuy;
    double sx, cx, sy, cy;
    if (fabs(ddx) <= K[CI_SMALL]) {
        double sd, cd;
        sincos_small_k(ddx, sd, cd, K);
        sx = K[CI_SMUX] * cd + K[CI_CMUX] * sd;
        cx = K[CI_CMUX] * cd - K[CI_SMUX] * sd;
    } else {
        sincos_fast_k(K[CI_PHX] + ddx, sx, cx, K);
    }
    if (fabs(ddy) <= K[CI_SMALL]) {
        double sd, cd;
        sincos_small_k(ddy, sd, cd, K);
        sy = K[CI_SMUY] * cd + K[CI_CMUY] * sd;
        cy = K[CI_CMUY] * cd - K[CI_SMUY] * sd;
    } else {
        sincos_fast_k(K[CI_PHY] + ddy, sy, cy, K);
    }
    double Y[6], R[6];
    spmv_k(K + CI_B1H1, X, Y);
    R[0] = cx * Y[0] + sx * Y[1];
    R[1] = -sx * Y[0] + cx * Y[1];
    R[2] = cy * Y[2] + sy * Y[3];
    R[3] = -sy * Y[2] + cy * Y[3];
    R[4] = Y[4];
    R[5] = Y[5];
    spmv_k(K + CI_H2B2, R, X);
}

// Radiation damping and quantum excitation, src/Bunch.cpp:940-1044.
__device__ __forceinline__ void stage_synrad(double X[6], cdptr K, const double r[5], bool noise)
{
#pragma clang fp contract(fast)
    double N[6];
    spmv_k(K + CI_SRB, X, N);
    N[0] *= K[CI_LAM + 0];
    N[1] *= K[CI_LAM + 0];
    N[2] *= K[CI_LAM + 1];
    N[3] *= K[CI_LAM + 1];
    N[5] *= K[CI_LZ2];
    if (noise) {
        N[0] += K[CI_COEF + 0] * r[0];
        N[1] += K[CI_COEF + 0] * r[1];
        N[2] += K[CI_COEF + 1] * r[2];
        N[3] += K[CI_COEF + 1] * r[3];
        N[5] += K[CI_COEF + 2] * r[4];
    }
    spmv_k(K + CI_SRH, N, X);
}

// --------------------------------------------------------------- the pass
// One streaming pass over the particles; `ops` selects the stages, in the
// order of MPBeam::Run (src/MPBeam.cpp:1059-1102):
//   KICK   dp += wake kick of the previous binning (deferred SR wake kick,
//          src/MPBunch.cpp:2015-2025), survivors at kick time only
//   MOMB   moments of z/dp after the kick (completes the previous turn)
//   MAP    section map (src/Bunch.cpp:625-755)
//   ZSTAT  post-map z moments and survivors' z range (SR wake edges)
//   SKEW   thin skew quad (src/Bunch.cpp:757-795)
//   RF/DRIFT/U0  (src/Bunch.cpp:453-501)
//   SYNRAD (src/Bunch.cpp:940-1044)
//   BIN    longitudinal histogram (src/MPBunch.cpp:1915-1919)
//   LOSS   loss marking (src/Bunch.cpp:172-200)
//   MOMA   moments after marking (src/MPBunch.cpp:342-498)
//   MOMD   z/dp moments after marking (no wake to defer)
// M is the compile-time set of stages a variant may run; stages outside it
// are compiled out.
__host__ __device__ constexpr bool acc_used(uint32_t M, int j)
{
    return (j <= A_ZMAX && j != A_LOSS) ? (M & OP_MOMA) != 0
         : j == A_LOSS                  ? (M & OP_LOSS) != 0
         : j <= B_QZD                   ? (M & (OP_MOMB | OP_MOMD)) != 0
         : j <= C_ZMAX                  ? (M & OP_ZSTAT) != 0
                                        : (M & OP_LRC) != 0;
}

// particles each thread has in flight ahead of the one it computes (two or
// three steps ahead measured 2 % and 4 % slower on the 480-bunch train: the
// pass is not short of loads in flight)
constexpr int kPassDepth = 1;

template <uint32_t M, bool EXACT, int NOISE, int NRF>
__global__ void __launch_bounds__(kBlock)
k_pass(PassParams P)
{
    extern __shared__ double smem_d[];
    double *kick_lds = smem_d;                                            // nbins
    uint32_t *hist_lds = reinterpret_cast<uint32_t *>(smem_d + P.nbins);  // nbins
    double *bl_lds = smem_d + P.nbins + (P.nbins + 1) / 2;                // bl_nbl * (bl_nbins + 1)
    __shared__ double red[kBlock / 64][kNPart];

    // EXACT variants run exactly the stages in M (no runtime stage tests);
    // NOISE >= 0 fixes the noise source at compile time
    const uint32_t ops = EXACT ? M : (P.ops & M);
    const int noise_mode = NOISE >= 0 ? NOISE : P.noise_mode;
    const int n_rf = NRF >= 0 ? NRF : P.n_rf;
    const int tid = threadIdx.x;
    const int s_first = P.wg_seg[blockIdx.x], s_last = P.wg_seg[blockIdx.x + 1];
    const uint32_t turn = P.turn_dev ? *P.turn_dev + P.turn : P.turn;
    __shared__ NormalTab ntab;
    if ((ops & OP_SYNRAD) && noise_mode == 2) {
        load_normal_tab(ntab, P.ntab, tid, kBlock);
        __syncthreads();
    }

    for (int s = s_first; s < s_last; s++) {
        const int b = P.seg_bunch[s];
        const int64_t i0 = P.seg_begin[s], i1 = P.seg_end[s];
        const int64_t bunch_first = P.bunch_off[b];
        // the bunch's moment shift, re-read as scalars where it is used
        const uint64_t sh_addr = (uint64_t)(uintptr_t)(P.shift + 6 * b);

        const bool use_hist = (ops & OP_BIN) && P.nbins > 0;
        const bool use_kick = (ops & OP_KICK) && P.nbins > 0;
        if (use_kick)
            for (int j = tid; j < P.nbins; j += kBlock) kick_lds[j] = P.kick_prev[(int64_t)b * P.nbins + j];
        if (use_hist)
            for (int j = tid; j < P.nbins; j += kBlock) hist_lds[j] = 0u;
        // beam loading: the cavities' voltages this bunch sees this turn and
        // its bins over the z range after the map (GetZMinMax's padding)
        // (MPBunch::BunchMomentumUpdateDueToRFModeStable,
        // src/MPBunch.cpp:1338-1375); the kick of every bin centre, head
        // first, goes to LDS: bl_lds[q][k] for cavity q, then the self-loss
        // term of each cavity
        double bl_zmax = 0, bl_dz = 1;
        if (ops & (OP_BL | OP_BLH)) {
            bl_zmax = P.zrange[2 * b + 1] + 1.e-6;
            const double zmin = P.zrange[2 * b] - 1.e-6;
            bl_dz = (bl_zmax - zmin) / P.bl_nbins;
        }
        if ((ops & OP_BL) && P.bl_walk) {
            // walked cavities: each bin's cavity and self-loss kicks, then
            // each cavity's feedback kick, as the walk left them
            const int nk = P.bl_nbl * P.bl_nbins;
            const double *w = P.bl_walk + (int64_t)b * nk * 2;
            for (int j = tid; j < nk; j += kBlock) {
                bl_lds[j] = w[2 * j];
                bl_lds[nk + j] = w[2 * j + 1];
            }
            for (int q = tid; q < P.bl_nbl; q += kBlock)
                bl_lds[2 * nk + q] = P.bl_fbk ? P.bl_fbk[(int64_t)b * P.bl_nbl + q] : 0.0;
        } else if (ops & OP_BL) {
            const int bt = (int)turn < P.bl_nturns ? (int)turn : P.bl_nturns - 1;
            const double *blv = P.bl_tab + ((int64_t)bt * P.nb + b) * P.bl_nbl * 3;
            const int nk = P.bl_nbl * P.bl_nbins;
            for (int j = tid; j < nk; j += kBlock) {
                const int q = j / P.bl_nbins, kb = j - q * P.bl_nbins;
                const double zc = bl_zmax - kb * bl_dz - bl_dz / 2.0;
                double sn, cs;
                sincos_fast_k(-zc * P.bl_kfac[q], sn, cs, reload(P.cst));
                bl_lds[j] = (blv[3 * q] * cs - blv[3 * q + 1] * sn) / P.bl_energy / P.bl_beta2;
            }
            for (int q = tid; q < P.bl_nbl; q += kBlock)
                bl_lds[nk + q] = blv[3 * q + 2] / 2.0 / P.bl_energy / P.bl_beta2;
        }
        if (ops & OP_BLH)
            for (int j = tid; j < P.bl_nbins; j += kBlock) reinterpret_cast<uint32_t *>(bl_lds)[j] = 0u;
        // the bunch's long-range wake kick of the previous turn (uniform)
        double lrk0 = 0, lrk1 = 0, lrk2 = 0;
        if (ops & OP_LRK) {
            lrk0 = P.lr_kick[3 * b];
            lrk1 = P.lr_kick[3 * b + 1];
            lrk2 = P.lr_kick[3 * b + 2];
        }
        double klo = 0, kdz = 1, blo = 0, bdz = 1;
        if (use_kick) { klo = P.edges_prev[2 * b]; kdz = P.edges_prev[2 * b + 1]; }
        if (use_hist) { blo = P.edges_now[2 * b]; bdz = P.edges_now[2 * b + 1]; }
        const double kinv = 1.0 / kdz, binv = 1.0 / bdz;
        __syncthreads();

        double acc[kNPart];
#pragma unroll
        for (int j = 0; j < kNPart; j++) acc[j] = 0.0;
        acc[A_ZMIN] = INFINITY; acc[A_ZMAX] = -INFINITY;
        acc[C_ZMIN] = INFINITY; acc[C_ZMAX] = -INFINITY;

        // per-segment base pointers (uniform) and a 32-bit lane offset, so the
        // loads and stores use scalar-base + vector-offset addressing
        double *__restrict__ sx_ = P.x + i0, *__restrict__ spx_ = P.px + i0, *__restrict__ sy_ = P.y + i0;
        double *__restrict__ spy_ = P.py + i0, *__restrict__ sz_ = P.z + i0, *__restrict__ sdp_ = P.dp + i0;
        uint8_t *__restrict__ sfl_ = P.flag + i0;
        const uint32_t nseg_i = (uint32_t)(i1 - i0);
        // software pipeline: the coordinates of the particles kPassDepth
        // steps ahead are loaded while the current one is computed, so HBM
        // latency (longer than one step's VALU work) overlaps the compute
        double Xq[kPassDepth][6];
        uint32_t fq[kPassDepth];
        // the coordinates are streamed once per pass: nontemporal loads and
        // stores keep them from displacing the tables in the caches
#pragma unroll
        for (int d = 0; d < kPassDepth; d++) {
            const uint32_t kd = (uint32_t)tid + (uint32_t)d * kBlock;
#pragma unroll
            for (int j = 0; j < 6; j++) Xq[d][j] = 0.0;
            fq[d] = 0u;
            if (kd < nseg_i) {
                Xq[d][0] = ntload(sx_ + kd); Xq[d][1] = ntload(spx_ + kd); Xq[d][2] = ntload(sy_ + kd);
                Xq[d][3] = ntload(spy_ + kd); Xq[d][4] = ntload(sz_ + kd); Xq[d][5] = ntload(sdp_ + kd);
                fq[d] = ntload(sfl_ + kd);
            }
        }
        for (uint32_t k = tid; k < nseg_i; k += kBlock) {
            const int64_t i = i0 + k;
            double X[6];
#pragma unroll
            for (int j = 0; j < 6; j++) X[j] = Xq[0][j];
            uint32_t f = fq[0];
#pragma unroll
            for (int d = 0; d + 1 < kPassDepth; d++) {
#pragma unroll
                for (int j = 0; j < 6; j++) Xq[d][j] = Xq[d + 1][j];
                fq[d] = fq[d + 1];
            }
            const uint32_t kn = k + (uint32_t)kPassDepth * kBlock;
            if (kn < nseg_i) {
                double *Xl = Xq[kPassDepth - 1];
                Xl[0] = ntload(sx_ + kn); Xl[1] = ntload(spx_ + kn); Xl[2] = ntload(sy_ + kn);
                Xl[3] = ntload(spy_ + kn); Xl[4] = ntload(sz_ + kn); Xl[5] = ntload(sdp_ + kn);
                fq[kPassDepth - 1] = ntload(sfl_ + kn);
            }

            // long-range wake kick (src/Bunch.cpp:1091-1101), applied before
            // the short-range one as the reference orders them
            // (src/MPBeam.cpp:1085-1086); survivors at kick time
            if ((ops & OP_LRK) && ((f & 3u) == 0u || (f & kNewLost))) {
                X[1] += lrk0;
                X[3] += lrk1;
                X[5] += lrk2;
            }
            if (use_kick) {
                // survivors when the kick was computed: code 0, or lost at
                // the marking that followed it
                if ((f & 3u) == 0u || (f & kNewLost)) {
                    const int kb = bin_of(X[4], klo, kdz, kinv, P.nbins);
                    if (kb >= 0) X[5] += kick_lds[kb];
                }
            }
            f &= 3u;
            const bool alive = f == 0u;
            if ((ops & OP_MOMB) && alive) {
#pragma clang fp contract(fast)
                cdptr sh = reload(sh_addr);
                const double dz = X[4] - sh[4], dd = X[5] - sh[5];
                acc[B_N] += 1.0;
                acc[B_SZ] += dz;
                acc[B_SD] += dd;
                acc[B_QDD] += dd * dd;
                acc[B_QZD] += dz * dd;
            }
            if ((ops & OP_MAP) && alive) stage_map(X, reload(P.cst));
            if ((ops & OP_LRC) && alive) {
                // bunch centroid after the section map (MPBeamRMSCal,
                // src/MPBeam.cpp:1068), the long-range wake's source
                cdptr sh = reload(sh_addr);
                acc[L_N] += 1.0;
                acc[L_SX] += X[0] - sh[0];
                acc[L_SY] += X[2] - sh[2];
                acc[L_SZ] += X[4] - sh[4];
            }
            if ((ops & OP_ZSTAT) && alive) {
                const double dz = X[4] - reload(sh_addr)[4];
                acc[C_N] += 1.0;
                acc[C_SZ] += dz;
                acc[C_QZZ] += dz * dz;
            }
            if ((ops & OP_SKEW) && alive) {
                // identity with (1,2) = (3,0) = K, summed in column order
                cdptr K = reload(P.cst);
                const double npx = X[1] + K[CI_SKEW] * X[2];
                const double npy = K[CI_SKEW] * X[0] + X[3];
                X[1] = npx;
                X[3] = npy;
            }
            if (ops & OP_RF) {
                cdptr K = reload(P.cst);
                for (int r = 0; r < n_rf; r++) {
                    double s, c;
                    const double th = X[4] * K[CI_RFK + r];
                    if (fabs(th) <= K[CI_SMALL]) sincos_small_k(th, s, c, K);
                    else sincos_fast_k(th, s, c, K);
                    const double re = K[CI_RFVR + r] * c - K[CI_RFVI + r] * s;
                    X[5] += re * K[CI_RFSCALE];
                }
            }
            if ((ops & OP_BL) && alive) {
                // the particle's bin, head first (clamped: the padded range
                // holds every survivor)
                int kb = (int)floor((bl_zmax - X[4]) / bl_dz);
                kb = kb < 0 ? 0 : (kb >= P.bl_nbins ? P.bl_nbins - 1 : kb);
                const int nk = P.bl_nbl * P.bl_nbins;
                for (int q = 0; q < P.bl_nbl; q++) {
                    X[5] += bl_lds[q * P.bl_nbins + kb];
                    X[5] += P.bl_walk ? bl_lds[nk + q * P.bl_nbins + kb] : bl_lds[nk + q];
                }
            }
            if ((ops & OP_BL) && P.bl_fbmask) {
                // the cavities' direct feedback, on every particle
                // (Bunch::GetLongiKickDueToCavFB, src/Bunch.cpp:431-449)
                const int nk = P.bl_nbl * P.bl_nbins;
                for (int q = 0; q < P.bl_nbl; q++)
                    if ((P.bl_fbmask >> q) & 1u) X[5] += bl_lds[2 * nk + q];
            }
            if ((ops & OP_BLH) && alive) {
                int kb = (int)floor((bl_zmax - X[4]) / bl_dz);
                kb = kb < 0 ? 0 : (kb >= P.bl_nbins ? P.bl_nbins - 1 : kb);
                atomicAdd(&reinterpret_cast<uint32_t *>(bl_lds)[kb], 1u);
            }
            if (ops & OP_DRIFT) {
                cdptr K = reload(P.cst);
                const double d = X[5];
                X[4] -= K[CI_DRIFTC] * (K[CI_ALPHAC] * d + K[CI_ALPHAC + 1] * (d * d) +
                                         K[CI_ALPHAC + 2] * (d * d * d));
            }
            if (ops & OP_U0) X[5] -= reload(P.cst)[CI_U0];
            if ((ops & OP_SYNRAD) && alive) {
                double r[5] = {0, 0, 0, 0, 0};
                if (noise_mode == 1) {
#pragma unroll
                    for (int j = 0; j < 5; j++) r[j] = P.noise[j * P.ntot + i];
                } else if (noise_mode == 2) {
                    qe_normals(ntab, P.seed, turn, (uint64_t)i, r);
                }
                stage_synrad(X, reload(P.cst), r, noise_mode != 0);
            }
            if ((ops & OP_ZSTAT) && (alive || i0 + (int64_t)k == bunch_first)) {
                // GetZMinMax starts from particle 0 of the bunch even when it
                // is lost (src/MPBunch.cpp:1862-1863)
                acc[C_ZMIN] = fmin(acc[C_ZMIN], X[4]);
                acc[C_ZMAX] = fmax(acc[C_ZMAX], X[4]);
            }
            if (use_hist) {
                const int hb = bin_of(X[4], blo, bdz, binv, P.nbins);
                if (hb >= 0) atomicAdd(&hist_lds[hb], 1u);
            }
            uint32_t fnew = f;
            if (ops & OP_LOSS) {
                cdptr K = reload(P.cst);
                double cnt = 0.0;
                if (fabs(X[4]) > K[CI_ZLIM]) { fnew = 2u; cnt += 1.0; }
                if (outside_aperture(X[0], X[2], K)) { fnew = 1u; cnt += 1.0; }
                acc[A_LOSS] += cnt;
                if (f == 0u && fnew != 0u) fnew |= kNewLost;
            }
            if ((ops & OP_MOMA) && (fnew & 3u) == 0u) {
#pragma clang fp contract(fast)
                cdptr sh = reload(sh_addr);
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
#pragma clang fp contract(fast)
                cdptr sh = reload(sh_addr);
                const double dz = X[4] - sh[4], dd = X[5] - sh[5];
                acc[B_N] += 1.0;
                acc[B_SZ] += dz;
                acc[B_SD] += dd;
                acc[B_QDD] += dd * dd;
                acc[B_QZD] += dz * dd;
            }
            if (ops & OP_WRITE) {
                ntstore(sx_ + k, X[0]); ntstore(spx_ + k, X[1]); ntstore(sy_ + k, X[2]);
                ntstore(spy_ + k, X[3]); ntstore(sz_ + k, X[4]); ntstore(sdp_ + k, X[5]);
                ntstore(sfl_ + k, (uint8_t)fnew);
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
            P.seg_part[(int64_t)tid * P.nseg + (P.seg_rank ? P.seg_rank[s] : s)] = v;
        }
        if (use_hist) {
            for (int j = tid; j < P.nbins; j += kBlock) {
                const uint32_t c = hist_lds[j];
                if (c) atomicAdd(&P.bunch_hist[(int64_t)b * P.nbins + j], c);
            }
        }
        if (ops & OP_BLH) {
            for (int j = tid; j < P.bl_nbins; j += kBlock) {
                const uint32_t c = reinterpret_cast<uint32_t *>(bl_lds)[j];
                if (c) atomicAdd(&P.bl_hist[(int64_t)b * P.bl_nbins + j], c);
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
        const double lo = P.edges[2 * b], dz = P.edges[2 * b + 1], inv = 1.0 / dz;
        __syncthreads();
        for (int64_t i = i0 + tid; i < i1; i += kBlock) {
            const int hb = bin_of(P.z[i], lo, dz, inv, P.nbins);
            if (hb >= 0) atomicAdd(&hist_lds[hb], 1u);
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
// Per bunch: reduce the segment partials and write the moment record.
// grid = (chunks, nb), one wave per chunk of 64 segments: every lane folds
// one segment, the wave reduces, and the last chunk of a bunch to finish
// (device-scope counter) folds the chunk results in chunk order, so the
// result does not depend on scheduling.
__device__ __forceinline__ double part_identity(int j)
{
    return (j == A_ZMIN || j == C_ZMIN) ? INFINITY : (j == A_ZMAX || j == C_ZMAX) ? -INFINITY : 0.0;
}
__device__ __forceinline__ double part_fold(int j, double a, double v)
{
    return (j == A_ZMIN || j == C_ZMIN) ? fmin(a, v) : (j == A_ZMAX || j == C_ZMAX) ? fmax(a, v) : a + v;
}
__device__ __forceinline__ double part_wave(int j, double a)
{
    return (j == A_ZMIN || j == C_ZMIN) ? wave_min(a) : (j == A_ZMAX || j == C_ZMAX) ? wave_max(a) : wave_sum(a);
}

__device__ __forceinline__ void finalize_body(const FinParams &P, int c, int b)
{
    __shared__ double tot[kNPart];
    __shared__ uint32_t last;
    const int lane = threadIdx.x;
    const int s0 = P.bunch_seg[b], s1 = P.bunch_seg[b + 1];
    const int nchunk = (s1 - s0 + 63) / 64;
    if (c >= nchunk && !(c == 0 && nchunk == 0)) return;
    {
        const int s = s0 + c * 64 + lane;
        double v[kNPart];
#pragma unroll
        for (int j = 0; j < kNPart; j++) v[j] = (s < s1) ? P.seg_part[(int64_t)j * P.nseg + s] : part_identity(j);
        double *dst = P.chunk_part + ((int64_t)b * P.max_chunks + c) * kNPart;
#pragma unroll
        for (int j = 0; j < kNPart; j++) {
            const double r = part_wave(j, v[j]);
            if (lane == 0) dst[j] = r;
        }
    }
    if (nchunk > 1) {
        __threadfence();
        if (lane == 0) last = atomicAdd(&P.chunk_cnt[b], 1u) == (uint32_t)(nchunk - 1);
        __syncthreads();
        if (!last) return;
        __threadfence();
        if (lane == 0) P.chunk_cnt[b] = 0u;
    }
    {
        // lane q folds chunks q, q + 64, ...; then the fixed wave tree
        const double *src = P.chunk_part + (int64_t)b * P.max_chunks * kNPart;
        const int nc = nchunk > 0 ? nchunk : 1;
        double v[kNPart];
#pragma unroll
        for (int j = 0; j < kNPart; j++) v[j] = part_identity(j);
        for (int q0 = 0; q0 < nc; q0 += 64) {
            const int q = q0 + lane;
#pragma unroll
            for (int j = 0; j < kNPart; j++)
                v[j] = part_fold(j, v[j], q < nc ? src[(int64_t)q * kNPart + j] : part_identity(j));
        }
#pragma unroll
        for (int j = 0; j < kNPart; j++) {
            const double r = part_wave(j, v[j]);
            if (lane == 0) tot[j] = r;
        }
    }
    __syncthreads();
    if (lane != 0) return;

    double *sh = P.shift + 6 * b;
    const double ncount = (double)P.bunch_count[b];
    cts_moments *cur = P.cur + b;
    int32_t slot = P.slot, prev_slot = P.prev_slot;
    if (P.turn_dev && P.hist_cap > 0) {
        const uint32_t t = *P.turn_dev + (uint32_t)P.turn_off;
        slot = (int32_t)(t % (uint32_t)P.hist_cap);
        prev_slot = t >= 1u ? (int32_t)((t - 1u) % (uint32_t)P.hist_cap) : -1;
    }
    if ((P.ops & OP_LRC) && P.lr_cen) {
        // post-map centroids (with the shift of the pass, before MOMA moves it)
        const double n = tot[L_N];
        P.lr_cen[3 * b] = sh[0] + tot[L_SX] / n;
        P.lr_cen[3 * b + 1] = sh[2] + tot[L_SY] / n;
        P.lr_cen[3 * b + 2] = sh[4] + tot[L_SZ] / n;
    }
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
        if (P.hist && prev_slot >= 0) P.hist[(int64_t)prev_slot * P.nb + b] = *cur;
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
    if ((P.ops & (OP_MOMA | OP_MOMD | OP_LOSS)) && P.hist && slot >= 0)
        P.hist[(int64_t)slot * P.nb + b] = *cur;
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
// (src/MPBunch.cpp:1983-1998), summed over j in order.  grid = (bin groups
// of 64, nb): each wave stages the table and the counts in LDS and owns 64
// bins.  The histogram is double-buffered: this turn's buffer is read, the
// other one (next turn's) is cleared.
__device__ __forceinline__ void wake_body(const WakeParams &P, int group, int b)
{
    extern __shared__ double smem_d[];
    const int nb = P.nbins, nlag = 2 * nb - 1;
    double *n_lds = smem_d;
    double *w_lds = smem_d + nb;
    const int lane = threadIdx.x;
    const int i = group * 64 + lane;
    const uint32_t *hist = P.bunch_hist + (int64_t)b * nb;
    for (int j = lane; j < nb; j += 64) n_lds[j] = (double)hist[j];
    const double *W = P.table + (P.per_bunch_table ? (int64_t)b * nlag : 0);
    for (int j = lane; j < nlag; j += 64) w_lds[j] = W[j];
    if (i < nb) {
        P.counts_out[(int64_t)b * nb + i] = (int32_t)hist[i];
        P.hist_clear[(int64_t)b * nb + i] = 0u;
    }
    __syncthreads();
    if (i < nb) {
        double a = P.poten[(int64_t)b * nb + i];
        const double *Wi = w_lds + i + (nb - 1);
#pragma unroll 16
        for (int j = 0; j < nb; j++) a += Wi[-j] * n_lds[j];
        a *= P.scale[b];
        P.poten[(int64_t)b * nb + i] = a;
        P.kick[(int64_t)b * nb + i] = a;
    }
    if (group == 0 && lane == 0) {
        P.edges_prev[2 * b] = P.edges_now[2 * b];
        P.edges_prev[2 * b + 1] = P.edges_now[2 * b + 1];
    }
}

__global__ void __launch_bounds__(64) k_finalize(FinParams P) { finalize_body(P, blockIdx.x, blockIdx.y); }
__global__ void __launch_bounds__(64) k_wake(WakeParams P)
{
    if ((int)blockIdx.x * 64 < P.nbins) wake_body(P, blockIdx.x, blockIdx.y);
}
// finalize (z = 0) and wake (z = 1) of one turn in a single launch: the
// fixed-grid wake needs only the histogram, not the moments
__global__ void __launch_bounds__(64) k_post(FinParams F, WakeParams W)
{
    if (blockIdx.z == 0) finalize_body(F, blockIdx.x, blockIdx.y);
    else if ((int)blockIdx.x * 64 < W.nbins) wake_body(W, blockIdx.x, blockIdx.y);
}

// ------------------------------------------------------- ion kick
// Electron side of MPBunch::SSIonBunchInteraction (src/MPBunch.cpp:1165-1206)
// followed by BunchTransferDueToIon (src/Bunch.cpp:213-220): survivors get
// px/py += sum over species of 2*N_ion*r_e/gamma * F(x - xbar, y - ybar).
// Four waves per SIMD (128 VGPRs): the field is compute-bound and the
// extra waves hide its latency better than the registers they cost
// (a few spills, in the rarely taken far-field branch).
__global__ void __launch_bounds__(kBlock, 4)
k_ion_kick(IonKickParams P)
{
    __shared__ BeField sF[CTS_MAX_SPECIES];
    const int tid = threadIdx.x;
    // the segments in particle order dealt round-robin to this launch's own
    // grid (four workgroups per CU, the occupancy the kernel is compiled
    // for): the workgroups running at one time stream one compact window
    for (int r = blockIdx.x; r < P.nseg; r += gridDim.x) {
        const int s = P.seg_order ? P.seg_order[r] : r;
        const int b = P.seg_bunch[s];
        const double *c = P.ion_e + (int64_t)b * CTS_MAX_SPECIES * 5;
        // the clouds' constants, once per segment
        if (tid < P.nspec) sF[tid] = be_plan(c[5 * tid + 3], c[5 * tid + 4]);
        __syncthreads();
        // the momenta are read with the positions, so that one memory
        // latency per electron, not two, sits in front of the field
        for (int64_t i = P.seg_begin[s] + tid; i < P.seg_end[s]; i += kBlock) {
            const uint8_t f = P.flag[i];
            const double x = P.x[i], y = P.y[i], px0 = P.px[i], py0 = P.py[i];
            if (f & 3u) continue;
            double fxs = 0.0, fys = 0.0;
            for (int p = 0; p < P.nspec; p++) {
                double fx, fy;
                be_field(x - c[5 * p + 1], y - c[5 * p + 2], sF[p], fx, fy);
                fxs = fxs + c[5 * p] * fx;
                fys = fys + c[5 * p] * fy;
            }
            P.px[i] = px0 + fxs;
            P.py[i] = py0 + fys;
        }
        __syncthreads();
    }
}

// The same kick with the field surrogate inside its domain and the exact
// field elsewhere.  The loop runs in whole blocks so that every lane of a
// wave reaches the network together.
__global__ void __launch_bounds__(kBlock)
k_ion_kick_sur(IonKickParams P)
{
    __shared__ __attribute__((aligned(16))) MlpLds L;
    const int tid = threadIdx.x;
    load_mlp_lds(L, P.sur_params, tid, kBlock);
    __syncthreads();
    for (int r = blockIdx.x; r < P.nseg; r += gridDim.x) {
        const int s = P.seg_order ? P.seg_order[r] : r;
        const int b = P.seg_bunch[s];
        const double *c = P.ion_e + (int64_t)b * CTS_MAX_SPECIES * 5;
        for (int64_t base = P.seg_begin[s]; base < P.seg_end[s]; base += kBlock) {
            const int64_t i = base + tid;
            const bool active = i < P.seg_end[s] && (P.flag[i] & 3u) == 0;
            const double x = active ? P.x[i] : 0.0, y = active ? P.y[i] : 0.0;
            double fxs = 0.0, fys = 0.0;
            uint32_t mask = 0;
            for (int p = 0; p < P.nspec; p++) {
                double fx = 0.0, fy = 0.0;
                const double dx = x - c[5 * p + 1], dy = y - c[5 * p + 2];
                const bool ok = sur_field(L, P.sur_norm, active, dx, dy, c[5 * p + 3], c[5 * p + 4], fx, fy);
                if (ok) {
                    fxs = fxs + c[5 * p] * fx;
                    fys = fys + c[5 * p] * fy;
                } else {
                    mask |= 1u << p;
                }
            }
            if (active) {
                P.px[i] += fxs;
                P.py[i] += fys;
            }
            // points outside the surrogate's domain: queued for the exact
            // kernel, one counter update per wave
            const bool q = active && mask != 0u;
            const uint64_t bal = __ballot(q);
            if (bal) {
                const int lane = tid & 63;
                const uint32_t nq = (uint32_t)__popcll(bal);
                uint32_t first = 0;
                const int leader = __ffsll((long long)bal) - 1;
                if (lane == leader) first = atomicAdd(P.fb_count, nq);
                first = __shfl(first, leader, 64);
                if (q) {
                    const uint32_t pos = first + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                    P.fb_list[pos] = ((uint64_t)i << 4) | mask;
                }
            }
        }
    }
}

// Exact kicks of the queued points (the species in each entry's mask).
__global__ void __launch_bounds__(kBlock)
k_ion_kick_fb(IonKickParams P)
{
    const uint32_t n = *P.fb_count;
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < n; e += gridDim.x * kBlock) {
        const uint64_t ent = P.fb_list[e];
        const int64_t i = (int64_t)(ent >> 4);
        const uint32_t mask = (uint32_t)(ent & 15u);
        int lo = 0, hi = P.nb;  // bunch of particle i: bunch_off[b] <= i < bunch_off[b + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (P.bunch_off[mid] <= i) lo = mid;
            else hi = mid;
        }
        const double *c = P.ion_e + (int64_t)lo * CTS_MAX_SPECIES * 5;
        const double x = P.x[i], y = P.y[i];
        double fxs = 0.0, fys = 0.0;
        for (int p = 0; p < P.nspec; p++) {
            if (!(mask & (1u << p))) continue;
            double fx, fy;
            ion_field(x - c[5 * p + 1], y - c[5 * p + 2], c[5 * p + 3], c[5 * p + 4], fx, fy);
            fxs = fxs + c[5 * p] * fx;
            fys = fys + c[5 * p] * fy;
        }
        P.px[i] += fxs;
        P.py[i] += fys;
    }
}

// -------------------------------------------------------- test hooks
__global__ void __launch_bounds__(kBlock)
k_qe_debug(const float2 *tab, uint64_t seed, uint32_t turn, uint64_t base, int64_t n, double *out)
{
    __shared__ NormalTab T;
    load_normal_tab(T, tab, threadIdx.x, kBlock);
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r[5];
    qe_normals(T, seed, turn, base + (uint64_t)i, r);
#pragma unroll
    for (int j = 0; j < 5; j++) out[j * n + i] = r[j];
}

// Test hook: w(z) at given points.
__global__ void __launch_bounds__(kBlock)
k_w_debug(const double *re, const double *im, int64_t n, double *wr, double *wi)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a, b;
    faddeeva_w(re[i], im[i], a, b);
    wr[i] = a;
    wi[i] = b;
}
// the kick field (be_plan + be_field) at n points: in = {x, y, sx, sy}[n], out = {fx, fy}[n]
__global__ void __launch_bounds__(kBlock)
k_kick_field_debug(const double *in, int64_t n, double *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BeField F = be_plan(in[2 * n + i], in[3 * n + i]);
    double fx, fy;
    be_field(in[i], in[n + i], F, fx, fy);
    out[i] = fx;
    out[n + i] = fy;
}
hipError_t launch_kick_field_debug(const double *in, int64_t n, double *out, hipStream_t st)
{
    const int blocks = (int)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_kick_field_debug, dim3(blocks), dim3(kBlock), 0, st, in, n, out);
    return hipGetLastError();
}
hipError_t launch_w_debug(const double *re, const double *im, int64_t n, double *wr, double *wi, hipStream_t st)
{
    const int blocks = (int)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_w_debug, dim3(blocks), dim3(kBlock), 0, st, re, im, n, wr, wi);
    return hipGetLastError();
}

// ---------------------------------------------------- eigen-emittances
// The second moments of MPBunch::GetEigenEmit (src/MPBunch.cpp:500-547):
// every particle of a bunch, lost ones included, with x, px, y, py less
// the dispersion of section 0 times delta.  Block (c, b) sums chunk c of
// bunch b: out[(b * nchunk + c) * 15] = the ten products x x, x px, x y,
// x py, px px, px y, px py, y y, y py, py py, the four coordinate sums and
// the largest |coordinate| (the reference's integer-valued std::accumulate
// leaves its "means" at zero only while every coordinate is below 1 in
// magnitude).  Threads add in a fixed order and the block folds its lanes
// in a fixed tree, so the partial sums do not depend on timing.
__global__ void __launch_bounds__(kBlock)
k_eigen_sums(const double *x, const double *px, const double *y, const double *py, const double *dp,
             const int64_t *bunch_off, int nchunk, double ex, double exp_, double ey, double eyp, double *out)
{
    __shared__ double red[15][kBlock];
    const int c = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
    const int64_t b0 = bunch_off[b], n = bunch_off[b + 1] - b0;
    const int64_t per = (n + nchunk - 1) / nchunk;
    const int64_t i0 = b0 + (int64_t)c * per, i1 = b0 + std::min(n, (int64_t)(c + 1) * per);
    double a[15];
#pragma unroll
    for (int k = 0; k < 15; k++) a[k] = 0.0;
    for (int64_t i = i0 + tid; i < i1; i += kBlock) {
        const double d = dp[i];
        const double u0 = x[i] - ex * d, u1 = px[i] - exp_ * d, u2 = y[i] - ey * d, u3 = py[i] - eyp * d;
        a[0] += u0 * u0; a[1] += u0 * u1; a[2] += u0 * u2; a[3] += u0 * u3; a[4] += u1 * u1;
        a[5] += u1 * u2; a[6] += u1 * u3; a[7] += u2 * u2; a[8] += u2 * u3; a[9] += u3 * u3;
        a[10] += u0; a[11] += u1; a[12] += u2; a[13] += u3;
        a[14] = fmax(a[14], fmax(fmax(fabs(u0), fabs(u1)), fmax(fabs(u2), fabs(u3))));
    }
#pragma unroll
    for (int k = 0; k < 15; k++) red[k][tid] = a[k];
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (tid < o) {
            for (int k = 0; k < 14; k++) red[k][tid] += red[k][tid + o];
            red[14][tid] = fmax(red[14][tid], red[14][tid + o]);
        }
        __syncthreads();
    }
    if (tid < 15) out[((int64_t)b * nchunk + c) * 15 + tid] = red[tid][0];
}
hipError_t launch_eigen_sums(const double *x, const double *px, const double *y, const double *py, const double *dp,
                             const int64_t *bunch_off, int nb, int nchunk, const double eta[4], double *out,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_eigen_sums, dim3(nchunk, nb), dim3(kBlock), 0, st, x, px, y, py, dp, bunch_off, nchunk,
                       eta[0], eta[1], eta[2], eta[3], out);
    return hipGetLastError();
}

// ------------------------------------------------------ long-range wake
// MPBeam::LRWakeBeamIntaction (src/MPBeam.cpp:3232-3351) for one turn,
// one block per receiving bunch j: over the last nturns turns n (n = 0 the
// present one, bunches ahead of j only; n = nturns - 1 bunches from j on),
// tau = (h_i - h_j - n h) T_rf + (z_j - z_i[n]) / (c beta), the wake table
// interpolated linearly at -tau (GetTotLRWakeLinearFit,
// src/WakeFunction.cpp:204-229), F -= W N_i {x_i[n], y_i[n], 1}; scaled by
// e / E / beta^2.  The px, py (and, without a short-range wake, dp) means of
// the turn's moment record move by the kick, which every survivor gets with
// the next pass.  The threads' partial sums are folded in a fixed order.
__global__ void __launch_bounds__(256)
k_lrwake(LrwParams P)
{
    __shared__ double red[3][256];
    const int j = blockIdx.x, tid = threadIdx.x;
    const uint32_t t = P.turn_dev ? *P.turn_dev + P.turn : P.turn;
    const int T = P.nturns, nb = P.nb;
    const double zj = P.cen[3 * j + 2];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    bool ahead = false;
    for (int n = 0; n < T; n++) {
        int i0 = 0, i1 = nb - 1;
        if (n == 0) i1 = j - 1;
        else if (n == T - 1) i0 = j;
        const int64_t slot = (((int64_t)t - n) % T + T) % T;
        const double *pn = n == 0 ? P.cen : P.ring + slot * nb * 3;
        for (int i = i0 + tid; i <= i1; i += 256) {
            const int ntau = P.harm[i] - P.harm[j] - n * P.harmonic;
            const double tau = ntau * P.trf + (zj - pn[3 * i + 2]) / kClight / P.rbeta;
            if (tau > 0) {
                ahead = true;
                continue;
            }
            const double u = -tau / P.dt;
            int idx = (int)floor(u);
            idx = idx < 0 ? 0 : (idx > P.np - 2 ? P.np - 2 : idx);
            const double v1 = u - idx, v0 = 1 - v1;
            const double w0 = P.table[idx] * v0 + P.table[idx + 1] * v1;
            const double w1 = P.table[P.np + idx] * v0 + P.table[P.np + idx + 1] * v1;
            const double w2 = P.table[2 * P.np + idx] * v0 + P.table[2 * P.np + idx + 1] * v1;
            f0 -= w0 * P.ne[i] * pn[3 * i];
            f1 -= w1 * P.ne[i] * pn[3 * i + 1];
            f2 -= w2 * P.ne[i];
        }
    }
    if (ahead) atomicOr(P.status, 4u);
    red[0][tid] = f0;
    red[1][tid] = f1;
    red[2][tid] = f2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o)
            for (int k = 0; k < 3; k++) red[k][tid] += red[k][tid + o];
        __syncthreads();
    }
    if (tid < 3) P.ring[(int64_t)(t % (uint32_t)T) * nb * 3 + 3 * j + tid] = P.cen[3 * j + tid];
    if (tid == 0) {
        const double F[3] = {red[0][0] * P.scale, red[1][0] * P.scale, red[2][0] * P.scale};
        P.kick[3 * j] = F[0];
        P.kick[3 * j + 1] = F[1];
        P.kick[3 * j + 2] = F[2];
        cts_moments *m = P.cur + j;
        m->mean[1] += F[0];
        m->mean[3] += F[1];
        if (P.fix_dp) m->mean[5] += F[2];
        if (P.hist && P.hist_cap > 0) {
            cts_moments *hm = P.hist + (int64_t)(t % (uint32_t)P.hist_cap) * nb + j;
            hm->mean[1] = m->mean[1];
            hm->mean[3] = m->mean[3];
            if (P.fix_dp) hm->mean[5] = m->mean[5];
        }
    }
}
hipError_t launch_lrwake(const LrwParams &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_lrwake, dim3(p.nb), dim3(256), 0, st, p);
    return hipGetLastError();
}

// ------------------------------------------------------------ launches
constexpr uint32_t V_KICK = OP_KICK | OP_MOMB | OP_LRK | OP_WRITE;
constexpr uint32_t V_MOM = OP_MOMA | OP_MOMB | OP_MOMD | OP_ZSTAT | OP_LOSS;
constexpr uint32_t V_FIX = OP_KICK | OP_MOMB | OP_MAP | OP_SKEW | OP_RF | OP_DRIFT | OP_U0 | OP_SYNRAD |
                           OP_BIN | OP_LOSS | OP_MOMA | OP_MOMD | OP_WRITE | OP_LRC | OP_LRK;
constexpr uint32_t V_REF = OP_KICK | OP_MOMB | OP_MAP | OP_ZSTAT | OP_SKEW | OP_RF | OP_DRIFT | OP_U0 |
                           OP_SYNRAD | OP_LOSS | OP_MOMA | OP_MOMD | OP_WRITE | OP_LRC | OP_LRK;
// a turn with beam-loaded cavities: the read-only pass up to the section
// map (z range of the RF stage), then the turn's pass with the kick
constexpr uint32_t V_BLA = OP_KICK | OP_MAP | OP_ZSTAT | OP_SKEW | OP_LRK;
constexpr uint32_t V_BLB = V_FIX | V_REF | OP_BL;
constexpr uint32_t V_BLH = OP_KICK | OP_MAP | OP_SKEW | OP_LRK | OP_BLH;  // bins for the cavity walk
constexpr uint32_t V_ALL = 0xffffffffu;

// stage sets of the fused turn in its steady state
constexpr uint32_t T_QE = OP_KICK | OP_MOMB | OP_MAP | OP_RF | OP_DRIFT | OP_U0 | OP_SYNRAD | OP_BIN |
                          OP_LOSS | OP_MOMA | OP_WRITE;
constexpr uint32_t T_NOSR = OP_KICK | OP_MOMB | OP_MAP | OP_RF | OP_DRIFT | OP_U0 | OP_BIN | OP_LOSS |
                            OP_MOMA | OP_WRITE;
constexpr uint32_t T_QE_LR = T_QE | OP_LRC | OP_LRK;  // with the long-range wake
constexpr uint32_t T_QE_BL = T_QE | OP_BL;            // a beam-loaded cavity in place of the ideal one

// ---------------------------------------------------------- cavity walk
// One turn of MPBeam::BeamMomtumUpdateDueToRFTest (src/MPBeam.cpp:2624-2706)
// for cavities whose state depends on the particles: one wave per cavity
// walks the train bunch by bunch.  At each bunch the lanes take its bins
// (head first, over the survivors' z range after the map, padded as
// GetZMinMax pads it): the stable variant's kick Re(V e^{-i k z}) with the
// rigid self-induced V_b0 of the whole bunch
// (MPBunch::BunchMomentumUpdateDueToRFModeStable, src/MPBunch.cpp:1323-1417),
// or the bin-by-bin variant's (MPBunch::BunchMomentumUpdateDueToRFMode,
// src/MPBunch.cpp:1419-1514): bin k sees the beam-induced voltage
// V_k = D^k (V_start + sum_{m<k} V0_m D^-m), D = exp(-dt/t_F + 2 pi i f_r dt)
// (the recurrence V <- (V + V0_m) D in closed form, the sum a prefix sum over
// the lanes), plus the generator's voltage turned to the bin, and its own
// charge's V0_k / 2.  Then the cavity sample at the bunch's bucket for the
// direct feedback (Resonator::GetResonatorInfoAtNTrf, src/Resonator.cpp:
// 263-275; Bunch::GetLongiKickDueToCavFB, src/Bunch.cpp:431-449), the
// generator's dynamics bucket by bucket over the gap
// (Resonator::ResonatorDynamics, src/Resonator.cpp:235-261) and the decay of
// the beam-induced voltage to the next bunch (Resonator::GetBeamInducedVol,
// src/Resonator.cpp:205-232): over bunchGap tRF in the stable variant, over
// the time from this bunch's tail to the next one's head in the other
// (MPBeam::GetTimeDisToNextBunch, src/MPBeam.cpp:2716-2752).
namespace {

struct cplx_d { double r, i; };
__device__ __forceinline__ cplx_d cmul(cplx_d a, cplx_d b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cplx_d cadd(cplx_d a, cplx_d b) { return {a.r + b.r, a.i + b.i}; }
// exp(-x) exp(i y), formed as the reference's double times complex exponential
__device__ __forceinline__ cplx_d decay_turn(double x, double y)
{
    double sn, cs;
    sincos(y, &sn, &cs);
    const double e = exp(-x);
    return {e * cs, e * sn};
}

}  // namespace

__global__ void __launch_bounds__(64) k_bl_walk(BlWalkParams P)
{
    const int q = blockIdx.x, lane = threadIdx.x;
    double *st = P.state + (int64_t)q * kBlS;
    const bool exc = st[BS_EXC] != 0.0, dfb = st[BS_DFB] != 0.0;
    const double harm = st[BS_HARM], rs = st[BS_RS], q0 = st[BS_Q0], ql = st[BS_QL], det = st[BS_DET];
    const double fre = st[BS_FRE], tf = st[BS_TF];
    const cplx_d req = {st[BS_REQ], st[BS_REQ + 1]}, ig = {st[BS_IG], st[BS_IG + 1]};
    cplx_d vb = {st[BS_VB], st[BS_VB + 1]}, gen = {st[BS_GEN], st[BS_GEN + 1]};
    const double pi = 3.14159265358979323846;
    const double frf = 1.0 / P.trf;
    const double c0 = -1 * 2 * pi * fre * rs / q0;  // V0 per electron charge
    // the generator's step over one bucket (the same each time)
    const double sigma = 2.0 * pi * fre / (2.0 * ql), dom = 2.0 * pi * det, tB = P.trf;
    const cplx_d gstep = decay_turn(sigma * tB, dom * tB);
    const double es = exp(-sigma * tB), sn_d = sin(dom * tB), cs_d = cos(dom * tB);
    const double alpha = dom * es * sn_d - sigma * es * cs_d + sigma;
    const double beta = sigma * es * sn_d + dom * es * cs_d - dom;
    const double coefB = 2 * pi * fre / 2.0 * rs / q0 / (sigma * sigma + dom * dom);
    const cplx_d gadd = {coefB * (alpha * ig.r + beta * ig.i), coefB * (-beta * ig.r + alpha * ig.i)};
    const int ci = P.bl_index[q];
    const int nbins = P.nbins;
    const uint32_t turn = P.turn_dev ? *P.turn_dev + P.turn : P.turn;
    double *info_turn = (int64_t)turn < P.nturns ? P.info + (int64_t)turn * P.nb * P.nres * 8 : nullptr;
    for (int b = 0; b < P.nb; b++) {
        const double zmin = P.zrange[2 * b] - 1.e-6, zmax = P.zrange[2 * b + 1] + 1.e-6;
        const double dz = (zmax - zmin) / nbins, dtb = dz / kClight / P.rbeta;
        uint32_t *hist = P.hist + (int64_t)b * nbins;
        double *kick = P.kick + ((int64_t)b * P.nbl + q) * nbins * 2;
        double sn_[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // n, cav re/im, gen re/im, induced re/im (sum n), self re
        if (!exc) {
            const double vb0 = c0 * P.ne[b] * kEcharge;
            const cplx_d cav = cadd(vb, gen);
            for (int k = lane; k < nbins; k += 64) {
                const double zc = zmax - k * dz - dz / 2.0;
                double sn, cs;
                sincos(-zc * P.kfac[q], &sn, &cs);
                kick[2 * k] = (cav.r * cs - cav.i * sn) / P.energy / P.beta2;
                kick[2 * k + 1] = vb0 / 2.0 / P.energy / P.beta2;
                const double n = (double)hist[k];
                sn_[0] += n;
                sn_[1] += cav.r * n; sn_[2] += cav.i * n;
                sn_[3] += gen.r * n; sn_[4] += gen.i * n;
                sn_[5] += (vb.r + vb0) * n; sn_[6] += vb.i * n;
                sn_[7] += vb0 / 2.0 * n;
            }
            vb.r += vb0;
        } else {
            const double psi = 2.0 * pi * fre * dtb, lam = dtb / tf;
            cplx_d carry = vb;  // V_start + the sum over the earlier chunks' bins
            for (int k0 = 0; k0 < nbins; k0 += 64) {
                const int k = k0 + lane;
                const bool in = k < nbins;
                const double n = in ? (double)hist[k] : 0.0;
                const double v0 = c0 * P.mq[b] * n * kEcharge;
                // t_k = V0_k D^-k, its exclusive prefix sum over the lanes
                const cplx_d dinv = decay_turn(-(double)k * lam, -(double)k * psi);
                cplx_d t = {v0 * dinv.r, v0 * dinv.i};
                cplx_d inc = t;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const double ur = __shfl_up(inc.r, o, 64), ui = __shfl_up(inc.i, o, 64);
                    if (lane >= o) { inc.r += ur; inc.i += ui; }
                }
                const cplx_d excl = {carry.r + inc.r - t.r, carry.i + inc.i - t.i};
                const cplx_d dk = decay_turn((double)k * lam, (double)k * psi);
                const cplx_d vbk = cmul(dk, excl);
                if (in) {
                    const double zc = zmax - k * dz - dz / 2.0;
                    double sn, cs;
                    sincos(-zc * P.kfac[q], &sn, &cs);
                    const cplx_d g = {gen.r * cs - gen.i * sn, gen.r * sn + gen.i * cs};
                    const cplx_d cav = cadd(vbk, g);
                    kick[2 * k] = cav.r / P.energy / P.beta2;
                    kick[2 * k + 1] = v0 / 2.0 / P.energy / P.beta2;
                    sn_[0] += n;
                    sn_[1] += cav.r * n; sn_[2] += cav.i * n;
                    sn_[3] += g.r * n; sn_[4] += g.i * n;
                    sn_[5] += (vbk.r + v0) * n; sn_[6] += vbk.i * n;
                    sn_[7] += v0 / 2.0 * n;
                }
                carry.r += __shfl(inc.r, 63, 64);
                carry.i += __shfl(inc.i, 63, 64);
            }
            // after the last bin: D^nbins (V_start + the whole sum)
            vb = cmul(decay_turn((double)nbins * lam, (double)nbins * psi), carry);
        }
#pragma unroll
        for (int j = 0; j < 8; j++) sn_[j] = wave_sum(sn_[j]);
        if (lane == 0 && info_turn) {
            const double N = sn_[0];
            double *f = info_turn + ((int64_t)b * P.nres + ci) * 8;
            const double selfr = sn_[7] / N;
            f[0] = sn_[1] / N; f[1] = sn_[2] / N;
            f[2] = sn_[5] / N + (exc ? selfr : 0.0); f[3] = sn_[6] / N;
            f[4] = sn_[3] / N; f[5] = sn_[4] / N;
            f[6] = selfr; f[7] = 0.0;
        }
        // the cavity at the bunch's own bucket, read by the direct feedback
        if (dfb) {
            const double dt = -zmin / kClight;
            const cplx_d vB = cmul(vb, decay_turn(dt / tf, 2.0 * pi * fre * dt));
            const cplx_d dV = {-(vB.r + gen.r - req.r), -(vB.i + gen.i - req.i)};
            const double a = sqrt(dV.r * dV.r + dV.i * dV.i) * st[BS_FBG];
            const double ph = atan2(dV.i, dV.r) + st[BS_FBP] + st[BS_FBD] * fre * 2 * pi;
            if (lane == 0) P.fb[(int64_t)b * P.nbl + q] = a * cos(ph) / P.energy / P.beta2;
        }
        // the generator over the bunch's gap, bucket by bucket
        for (int k = 0; k < P.gap[b]; k++) gen = cadd(cmul(gen, gstep), gadd);
        // the beam-induced voltage decayed and turned to the next bunch
        const int nx = b + 1 < P.nb ? b + 1 : 0;
        const double T = exc ? P.gap[b] * P.trf + (zmin - (P.zrange[2 * nx + 1] + 1.e-6)) / kClight / P.rbeta
                             : P.gap[b] * P.trf;
        vb = cmul(vb, decay_turn(T / tf, 2.0 * pi * fre * T));
        // the bins are counted again next turn
        for (int k = lane; k < nbins; k += 64) hist[k] = 0u;
    }
    if (lane == 0) {
        st[BS_VB] = vb.r; st[BS_VB + 1] = vb.i;
        st[BS_GEN] = gen.r; st[BS_GEN + 1] = gen.i;
    }
    (void)harm;
    (void)frf;
}

hipError_t launch_bl_walk(const BlWalkParams &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_bl_walk, dim3(p.nbl), dim3(64), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_pass(const PassParams &p, int nwg, hipStream_t st)
{
    const uint32_t o = p.ops;
    // LDS: the wake's kicks and bins, then the cavities' per-bin kicks
    // (walked: cavity and self-loss parts per bin, and the feedback kicks) or
    // the walk's bin counts
    size_t bl = 0;
    if (o & OP_BL) bl = p.bl_walk ? (size_t)p.bl_nbl * (2 * p.bl_nbins + 1) : (size_t)p.bl_nbl * (p.bl_nbins + 1);
    if (o & OP_BLH) bl = std::max(bl, (size_t)(p.bl_nbins + 1) / 2);
    const size_t shm = (o & (OP_BL | OP_BLH)) ? ((size_t)p.nbins + (p.nbins + 1) / 2 + bl) * sizeof(double)
                                              : (size_t)p.nbins * (sizeof(double) + sizeof(uint32_t));
    const dim3 g(nwg), b(kBlock);
    const bool one_rf = p.n_rf == 1;
    if (one_rf && o == T_QE && p.noise_mode == 2)
        hipLaunchKernelGGL((k_pass<T_QE, true, 2, 1>), g, b, shm, st, p);
    else if (one_rf && o == T_QE_LR && p.noise_mode == 2)
        hipLaunchKernelGGL((k_pass<T_QE_LR, true, 2, 1>), g, b, shm, st, p);
    else if (p.n_rf == 0 && o == T_QE_BL && p.noise_mode == 2)
        hipLaunchKernelGGL((k_pass<T_QE_BL, true, 2, 0>), g, b, shm, st, p);
    else if (one_rf && o == T_QE && p.noise_mode == 1)
        hipLaunchKernelGGL((k_pass<T_QE, true, 1, 1>), g, b, shm, st, p);
    else if (one_rf && o == T_NOSR)
        hipLaunchKernelGGL((k_pass<T_NOSR, true, 0, 1>), g, b, shm, st, p);
    else if ((o & ~V_KICK) == 0) hipLaunchKernelGGL((k_pass<V_KICK, false, -1, -1>), g, b, shm, st, p);
    else if ((o & OP_BL) && (o & ~V_BLB) == 0) hipLaunchKernelGGL((k_pass<V_BLB, false, -1, -1>), g, b, shm, st, p);
    else if ((o & OP_BLH) && (o & ~V_BLH) == 0) hipLaunchKernelGGL((k_pass<V_BLH, false, -1, -1>), g, b, shm, st, p);
    else if ((o & OP_MAP) && (o & ~V_BLA) == 0)
        hipLaunchKernelGGL((k_pass<V_BLA, false, -1, -1>), g, b, shm, st, p);
    else if ((o & ~V_MOM) == 0) hipLaunchKernelGGL((k_pass<V_MOM, false, -1, -1>), g, b, shm, st, p);
    else if ((o & ~V_FIX) == 0) hipLaunchKernelGGL((k_pass<V_FIX, false, -1, -1>), g, b, shm, st, p);
    else if ((o & ~V_REF) == 0) hipLaunchKernelGGL((k_pass<V_REF, false, -1, -1>), g, b, shm, st, p);
    else hipLaunchKernelGGL((k_pass<V_ALL, false, -1, -1>), g, b, shm, st, p);
    return hipGetLastError();
}
hipError_t launch_ion_kick(const IonKickParams &p, int nwg, hipStream_t st)
{
    if (p.sur_params) {
        hipLaunchKernelGGL(k_ion_kick_sur, dim3(nwg), dim3(kBlock), 0, st, p);
        hipLaunchKernelGGL(k_ion_kick_fb, dim3(512), dim3(kBlock), 0, st, p);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_ion_kick, dim3(nwg), dim3(kBlock), 0, st, p);
    return hipGetLastError();
}
__global__ void k_set_turn(uint32_t *t, uint32_t v)
{
    if (threadIdx.x == 0) *t = v;
}
__global__ void k_advance_turn(uint32_t *t, uint32_t by)
{
    if (threadIdx.x == 0) *t += by;
}
hipError_t launch_set_turn(uint32_t *t, uint32_t v, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_turn, dim3(1), dim3(64), 0, st, t, v);
    return hipGetLastError();
}
hipError_t launch_advance_turn(uint32_t *t, uint32_t by, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance_turn, dim3(1), dim3(64), 0, st, t, by);
    return hipGetLastError();
}
hipError_t launch_bin(const BinParams &p, int nwg, hipStream_t st)
{
    hipLaunchKernelGGL(k_bin, dim3(nwg), dim3(kBlock), (size_t)p.nbins * sizeof(uint32_t), st, p);
    return hipGetLastError();
}
hipError_t launch_finalize(const FinParams &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_finalize, dim3(p.max_chunks, p.nb), dim3(64), 0, st, p);
    return hipGetLastError();
}
hipError_t launch_post(const FinParams &f, const WakeParams &w, hipStream_t st)
{
    const int gx = std::max(f.max_chunks, (w.nbins + 63) / 64);
    const size_t shm = (size_t)(3 * w.nbins - 1) * sizeof(double);
    hipLaunchKernelGGL(k_post, dim3(gx, f.nb, 2), dim3(64), shm, st, f, w);
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
    const size_t shm = (size_t)(3 * p.nbins - 1) * sizeof(double);
    hipLaunchKernelGGL(k_wake, dim3((p.nbins + 63) / 64, nb), dim3(64), shm, st, p);
    return hipGetLastError();
}
hipError_t launch_qe_debug(const float2 *tab, uint64_t seed, uint32_t turn, uint64_t base, int64_t n, double *out,
                           hipStream_t st)
{
    const int blocks = (int)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_qe_debug, dim3(blocks), dim3(kBlock), 0, st, tab, seed, turn, base, n, out);
    return hipGetLastError();
}

}  // namespace cts
