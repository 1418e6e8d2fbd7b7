// cts_internal.hpp -- kernel parameter blocks and launch wrappers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/cetasim_hip.h"
#include "cts_device.hpp"
#include "cts_faddeeva.hpp"
#include "cts_mlp.hpp"

namespace cts {

constexpr int kBlock = 256;
constexpr uint32_t kNewLost = 0x80u;

enum : uint32_t {
    OP_KICK = 1u << 0,
    OP_MOMB = 1u << 1,
    OP_MAP = 1u << 3,
    OP_ZSTAT = 1u << 4,
    OP_SKEW = 1u << 5,
    OP_RF = 1u << 6,
    OP_DRIFT = 1u << 7,
    OP_U0 = 1u << 8,
    OP_SYNRAD = 1u << 9,
    OP_BIN = 1u << 10,
    OP_LOSS = 1u << 11,
    OP_MOMA = 1u << 12,
    OP_WRITE = 1u << 13,
    OP_MOMD = 1u << 14,  // z/dp moments at the end of the pass (no wake)
    OP_LRC = 1u << 15,   // post-map centroids of the long-range wake
    OP_LRK = 1u << 16,   // the previous turn's long-range wake kick (deferred)
    OP_BL = 1u << 17,    // beam-loaded cavities (rfMode 1), after the ideal ones
    OP_BLH = 1u << 18,   // per-bunch bins over the z range for the cavity walk (read-only pass)
};

// per-segment partial sums, stored value-major: seg_part[j * nseg + s]
enum {
    A_N = 0, A_S = 1, A_Q = 6, A_LOSS = 13, A_ZMIN = 14, A_ZMAX = 15,
    B_N = 16, B_SZ = 17, B_SD = 18, B_QDD = 19, B_QZD = 20,
    C_N = 21, C_SZ = 22, C_QZZ = 23, C_ZMIN = 24, C_ZMAX = 25,
    L_N = 26, L_SX = 27, L_SY = 28, L_SZ = 29,
    kNPart = 30
};

// Layout of the per-section constant block in device memory (doubles).
// The pass reads it through scalar loads inside the particle loop, so the
// ~120 lattice constants never have to stay resident in SGPRs.
enum : int {
    CI_B1H1 = 0, CI_H2B2 = 16, CI_SRB = 32, CI_SRH = 48,
    CI_ETAX = 64, CI_ETAPX = 65, CI_ETAY = 66, CI_ETAPY = 67,
    CI_ALPHAX = 68, CI_BETAX = 69, CI_ALPHAY = 70, CI_BETAY = 71,
    CI_IBETAX = 72, CI_IBETAY = 73, CI_PHX = 74, CI_PHY = 75, CI_WX = 76, CI_WY = 77,
    CI_CHX = 78, CI_CHY = 79, CI_ADTX = 80, CI_ADTY = 85, CI_SKEW = 90,
    CI_RFVR = 91, CI_RFVI = 95, CI_RFK = 99, CI_RFSCALE = 103,
    CI_DRIFTC = 104, CI_ALPHAC = 105, CI_U0 = 108, CI_LAM = 109, CI_LZ2 = 112, CI_COEF = 113,
    CI_ZLIM = 116, CI_APX = 117, CI_APY = 118, CI_IAPX = 119, CI_IAPY = 120,
    CI_CMUX = 121, CI_SMUX = 122, CI_CMUY = 123, CI_SMUY = 124,
    // polynomial coefficients and thresholds of the angle functions: kept
    // in the block so they reach the ALU as scalar operands instead of
    // occupying vector registers for the whole kernel
    CI_TWOPI = 125, CI_SMALL = 126,
    CI_PS = 128, CI_PC = 132,               // small-angle sin, cos (4 each)
    CI_FR = 136,                            // 2/pi and the three parts of pi/2
    CI_FS = 140, CI_FC = 147,               // reduced-angle sin, cos (7 each)
    kNConst = 160
};

struct PassParams {
    double *x, *px, *y, *py, *z, *dp;
    uint8_t *flag;
    const int64_t *seg_begin, *seg_end;
    const int32_t *seg_bunch, *wg_seg;
    const int64_t *bunch_off;
    double *seg_part;
    const int32_t *seg_rank;  // segment's place in particle order (null: identity)
    uint32_t *bunch_hist;
    const double *shift;
    const double *kick_prev, *edges_prev, *edges_now;
    const double *noise;
    uint64_t cst;  // device address of the section's constant block
    int64_t ntot;
    uint32_t ops;
    int32_t nseg, nbins, n_rf, noise_mode;
    // turn number keying the noise stream: turn, or *turn_dev + turn when
    // turn_dev is set (graph replays read the turn from device memory)
    uint32_t turn;
    const uint32_t *turn_dev;
    uint64_t seed;
    const float2 *ntab;  // normal tables (NormalTab layout: sc then lg)
    const double *lr_kick;  // [nb][3] long-range wake kick (px, py, dp) of OP_LRK
    // beam loading (OP_BL): per turn, bunch and cavity {Re V, Im V, Re V_b0};
    // the bunch's z range after the map (zrange, unpadded); kfac = 2 pi h
    // f_rf / (c beta) per cavity
    const double *bl_tab;
    const double *zrange;
    // walked cavities: per bunch the bins' kicks [nb][nbl][bl_nbins][2] and
    // the feedback kicks [nb][nbl] (bit q of bl_fbmask: cavity q feeds back);
    // OP_BLH counts survivors per bin into bl_hist [nb][bl_nbins]
    const double *bl_walk, *bl_fbk;
    uint32_t *bl_hist;
    uint32_t bl_fbmask;
    int32_t bl_nbl, bl_nbins, bl_nturns, nb;
    double bl_kfac[CTS_MAX_RF];
    double bl_energy, bl_beta2;
};

struct IonKickParams {
    const double *x, *y;
    double *px, *py;
    const uint8_t *flag;
    const int64_t *seg_begin, *seg_end;
    const int32_t *seg_bunch, *wg_seg;
    const double *ion_e;  // per bunch CTS_MAX_SPECIES x {coeffE, xbar, ybar, sx, sy}
    const float *sur_params;  // field surrogate (null: exact field)
    SurNorm sur_norm;
    // exact-field fallback list of the surrogate variant: entries
    // (particle << 4) | species mask; fb_count is zeroed before each phase
    uint32_t *fb_count;
    uint64_t *fb_list;
    const int64_t *bunch_off;
    int32_t nb;
    int32_t nspec;
    int32_t nseg;  // segments of the particle layout (seg_begin/seg_end/seg_bunch)
    const int32_t *seg_order;  // the segment at each place in particle order (null: identity)
};

// One segment of the ion turn: passages j_begin .. j_end - 1, all species.
// An ion's motion depends only on the bunches' moments, which are fixed for
// the turn, so every ion walks through the segment's passages on its own
// thread; the per-passage cloud moments are accumulated in 128-bit fixed
// point (order-independent, hence deterministic) and finished passage by
// passage after the last segment.  Between segments the cloud is compacted,
// so the waves of a segment carry few lost ions.
struct IonTurnParams {
    double *x[CTS_MAX_SPECIES], *y[CTS_MAX_SPECIES], *vx[CTS_MAX_SPECIES], *vy[CTS_MAX_SPECIES];
    uint8_t *alive[CTS_MAX_SPECIES];
    const int64_t *len;      // [nspec] ions of the current buffer (compacted, all alive)
    const double *bview;     // [nb][4] bunch xbar, ybar, rmsRx, rmsRy
    const double *size_max;  // [2] largest rmsRx, rmsRy over the bunches
    const double *coeffI;    // [nb][CTS_MAX_SPECIES] 2 N_e r_e (m_e/m_u) / A c
    const double *dt;        // [nb] C / h * gap / c
    // fixed-point sums (three limbs), kIonReplicas copies of
    // [nspec][nb][kIonAcc][3] acc_stride apart: wave w adds into copy
    // w % kIonReplicas, which spreads the atomics of one passage over as
    // many addresses; k_ion_reduce adds the copies up and clears them
    unsigned long long *acc;
    int64_t acc_stride;
    uint32_t *wcnt;          // [nspec][wcnt_stride]: survivors of each wave
    int64_t wcnt_stride;
    const float *sur_params; // field surrogate (null: exact field)
    SurNorm sur_norm;
    double loss_boundary, ap_x, ap_y;
    int64_t cap;             // slots per species buffer
    uint64_t seed;
    int32_t nb, n_new, w_new;         // w_new: waves per passage of new ions
    int32_t j_begin, j_end;           // the segment's passages
    uint32_t turn;                    // or *turn_dev + turn when turn_dev is set
    const uint32_t *turn_dev;
};

// Cloud moments and electron-side parameters of every passage.
struct IonFinishParams {
    const unsigned long long *acc;  // the reduced sums, [nspec][nb][kIonAcc][3]
    const double *bview;
    const double *ion_q;     // [nb][nspec]
    const int64_t *len0;     // [nspec] ions at the start of the turn
    int64_t *n_alive;        // [nspec] out: after the last passage
    double *mean_state;      // [2 * nspec]
    double *ion_e;           // [nb][CTS_MAX_SPECIES][5]
    double *cloud;           // [nb][CTS_MAX_SPECIES][6]
    uint32_t *status;
    double gamma;
    int64_t max_ions;
    int32_t nb, nspec, n_new;
};
constexpr int kIonAcc = 6;  // n, sum dx, sum dy, sum dx^2, sum dy^2, n kept

// Stable compaction of one species after a segment, wave by wave in the
// segment's wave order (old ions, then each passage's new ones: slot order).
struct IonCompactParams {
    const double *x, *y, *vx, *vy;
    const uint8_t *alive;
    double *ox, *oy, *ovx, *ovy;
    uint8_t *oalive;
    const int64_t *len;      // compacted ions of the source buffer before the segment
    int64_t *new_len;        // ions of the destination buffer
    const uint32_t *wcnt;    // survivors of every wave of the segment
    uint32_t *wofs;          // their exclusive prefix sums
    int64_t cap;
    int32_t n_new, w_new, nseg;
};

struct BinParams {
    const double *z;
    const int64_t *seg_begin, *seg_end;
    const int32_t *seg_bunch, *wg_seg;
    const double *edges;
    uint32_t *bunch_hist;
    int32_t nbins;
};

struct FinParams {
    const double *seg_part;
    const int32_t *bunch_seg;
    const int64_t *bunch_count;
    double *shift;
    cts_moments *cur;
    cts_moments *hist;  // history ring or null
    int32_t slot, prev_slot, nb, nbins, ref_edges, nseg;
    uint32_t ops;
    double *sigz_stale;
    double *zrange;     // survivors' z range (GetZMinMax, without the pad)
    double *edges_now;
    uint32_t *status;
    double *chunk_part;   // [nb][max_chunks][kNPart]
    uint32_t *chunk_cnt;  // [nb], left at zero after each launch
    int32_t max_chunks;
    // history slots from the device turn counter (graph replays):
    // slot = t % hist_cap, prev_slot = (t - 1) % hist_cap, t = *turn_dev + turn_off
    const uint32_t *turn_dev;
    int32_t turn_off, hist_cap;
    double *lr_cen;       // [nb][3] post-map centroids (OP_LRC)
};

// MPBeam::LRWakeBeamIntaction for one turn: one block per receiving bunch.
struct LrwParams {
    const double *table;   // [3][np] W_x, W_y, W_z at tau = -i dt
    int32_t np, nturns, nb, harmonic;
    double dt, trf, rbeta, scale;  // scale = e / E / rbeta^2
    const int32_t *harm;   // bunch harmonic numbers
    const double *ne;      // electrons per bunch
    const double *cen;     // [nb][3] this turn's post-map centroids
    double *ring;          // [nturns][nb][3] centroids by turn mod nturns
    double *kick;          // [nb][3] out
    cts_moments *cur, *hist;
    int32_t hist_cap, fix_dp;
    uint32_t turn;         // or *turn_dev + turn when turn_dev is set
    const uint32_t *turn_dev;
    uint32_t *status;
};
hipError_t launch_lrwake(const LrwParams &p, hipStream_t st);

// Host: the long-range wake table (w, rw, bbr: [3][np]; time[np] = -tau).
struct LrwTable {
    int np = 0;
    double dt = 0, trf = 0;
    std::vector<double> time, w, rw, bbr;
};
int lrwake_table(const cts_lrwake_config &c, double t0, int harmonic, LrwTable &tab, std::string &err);

// Broadband impedance (cts_bbi.hip): the grid, the real circular kernel of
// Z (frequency domain) or the wake table (time domain), and the smoothing
// filter, set up on the host.
constexpr int kBbiTaps = 51;          // the Gaussian filter's window
constexpr int kBbiWin = 8192;         // bins of one LDS histogram window (32 KB of counts)
constexpr int kBbiWinCopies = 2048;   // narrower ranges: one copy per wave in the same 32 KB
constexpr int kBbiDipWin = 1535;      // with the dipolar planes: counts and x, y limb sums (55 KB)
constexpr int kBbiMaxGridFD = 16383;  // 2 nimp - 1 (the host forms the kernel in O(n^2))
constexpr int kBbiMaxGridTD = 65535;  // (int)(bucket / dz)
struct BbiTable {
    int n = 0, is = 0, ie = 0;   // bins; smoothed range [is, ie)
    int td = 0, nw = 0;          // time domain; rows of its wake table
    int planes[5] = {0, 0, 0, 0, 0};  // the reference's flags: Z, Dx, Dy, Qx, Qy
    double zmin = 0, dz = 0, half = 0, bx = 1, by = 1;  // bin (int)((z - zmin + half)/dz)
    std::vector<double> g;       // frequency domain [5][n] kernels: V = -(profile (*) g_z), transverse
                                 // V = +(profile (*) g); time domain [5][nw] the wake table
    double taps[kBbiTaps];       // normalised Gaussian weights
};
int bbimp_table(const cts_bbimp_config &c, const cts_ring &ring, BbiTable &t, std::string &err);
struct BbiParams {
    int n, is, ie, nb, td, nw;
    int planes[5];
    double zmin, dz, half, energy, beta2, bx, by;
    const int64_t *seg_begin, *seg_end;
    const int32_t *seg_bunch, *wg_seg;
    const double *x, *y, *z;
    double *px, *py, *dp;
    const uint8_t *flag;
    uint32_t *hist;        // [nb][n], zero between turns
    unsigned long long *dsum;  // [nb][n][4] x, y limb sums (dipolar planes; else null), zero between turns
    double *prof;          // [3][nb][n]: charge, x-weighted, y-weighted (time domain: charge, mean x, mean y)
    double *wake;          // [5][nb][n], flag order
    const double *g;       // [5][n] kernels, or [5][nw] wake table
    const double *scale;   // [nb] charge per macroparticle (frequency domain: times c / dz; 0: skipped)
    double taps[kBbiTaps];
};
hipError_t launch_bbi_bin(const BbiParams &p, int nwg, hipStream_t st);
hipError_t launch_bbi_wake(const BbiParams &p, hipStream_t st);
hipError_t launch_bbi_kick(const BbiParams &p, int nwg, hipStream_t st);

// Host: beam-loaded cavities, tabulated per turn (cts_bl.cpp).
// the state of a walked cavity (kBlS doubles): bin-by-bin and feedback
// switches, harmonic, R_s, Q_0, Q_L, detuning, f_r, t_F, required voltage,
// beam-induced voltage, generator voltage and current (re, im), feedback
// gain, phase shift and delay
enum { BS_EXC, BS_DFB, BS_HARM, BS_RS, BS_Q0, BS_QL, BS_DET, BS_FRE, BS_TF, BS_REQ, BS_VB = BS_REQ + 2,
       BS_GEN = BS_VB + 2, BS_IG = BS_GEN + 2, BS_FBG = BS_IG + 2, BS_FBP, BS_FBD, kBlS };
struct BlTable {
    int nturns = 0, nbl = 0, nbins = 0, nres = 0;
    bool walk = false;           // state walked on the device every turn
    std::vector<int> bl_index;   // cavity index of each beam-loaded cavity
    std::vector<double> kfac;    // 2 pi h f_rf / (c beta) of each
    std::vector<double> dev;     // [nturns][nb][nbl][3]: Re V, Im V, Re V_b0
    std::vector<double> info;    // [nturns][nb][nres][8]
    std::vector<double> state;   // walk: [nbl][kBlS]
};
int beam_loading_tables(const cts_beam_loading_config &c, const cts_ring &ring, const std::vector<double> &ne,
                        const std::vector<int32_t> &gap, BlTable &t, std::string &err);
int beam_loading_walk_state(const cts_beam_loading_config &c, const cts_ring &ring, const std::vector<double> &ne,
                            const std::vector<int32_t> &gap, BlTable &t, std::string &err);
// One turn of the cavities' walk over the train (one wave per cavity).
struct BlWalkParams {
    double *state;          // [nbl][kBlS], carried from turn to turn
    uint32_t *hist;         // [nb][nbins] survivors per bin (cleared by the walk)
    const double *zrange;   // [nb][2] survivors' z range after the map (unpadded)
    const int32_t *gap;     // [nb] buckets to the next bunch
    const double *ne;       // [nb] electrons per bunch
    const double *mq;       // [nb] electrons per macroparticle
    double *kick;           // out [nb][nbl][nbins][2]: cavity and self-loss parts of each bin's kick
    double *fb;             // out [nb][nbl]: the feedback kick
    double *info;           // out [nturns][nb][nres][8] (turns past nturns are not kept)
    uint32_t turn;          // the turn, or *turn_dev + turn when turn_dev is set (graph replays)
    const uint32_t *turn_dev;
    int32_t nturns;
    int32_t nb, nbins, nbl, nres;
    int32_t bl_index[CTS_MAX_RF];
    double kfac[CTS_MAX_RF];
    double trf, rbeta, energy, beta2;
};
hipError_t launch_bl_walk(const BlWalkParams &p, hipStream_t st);
void lrwake_yokoya(double rx, double ry, double out[5]);

struct WakeTableParams {
    const double *edges;  // per table: {lo, dz}
    double *table;
    int32_t nbins, nres;
    double rs[CTS_MAX_RES], q[CTS_MAX_RES], omega[CTS_MAX_RES];
};

struct WakeParams {
    const uint32_t *bunch_hist;  // this turn's histogram
    uint32_t *hist_clear;        // the other buffer, cleared for the next turn
    int32_t *counts_out;
    const double *table;
    int32_t per_bunch_table;
    const double *scale;
    double *poten, *kick;
    const double *edges_now;
    double *edges_prev;
    int32_t nbins;
};

hipError_t launch_pass(const PassParams &p, int nwg, hipStream_t st);
// Host: the NormalTab contents (2 * kNormTab float2, sc then lg).
void build_normal_tables(float2 *out);
// device turn counter of graph replays: *t = v, and *t += by
hipError_t launch_set_turn(uint32_t *t, uint32_t v, hipStream_t st);
hipError_t launch_advance_turn(uint32_t *t, uint32_t by, hipStream_t st);
hipError_t launch_bin(const BinParams &p, int nwg, hipStream_t st);
hipError_t launch_ion_kick(const IonKickParams &p, int nwg, hipStream_t st);
hipError_t launch_finalize(const FinParams &p, hipStream_t st);
hipError_t launch_post(const FinParams &f, const WakeParams &w, hipStream_t st);
hipError_t launch_wake_table(const WakeTableParams &p, int ntables, hipStream_t st);
hipError_t launch_wake(const WakeParams &p, int nb, hipStream_t st);
hipError_t launch_ion_prep(const cts_moments *cur, int nb, double beta_x, double beta_y, double *bview,
                           double *size_max, uint32_t *fb_count, const int64_t *len, int64_t *len0, int nspec,
                           hipStream_t st);
// out[i] = sum over the kIonReplicas copies (stride n) of acc[i]; the
// copies are left at zero
hipError_t launch_ion_reduce(unsigned long long *acc, int64_t n, unsigned long long *out, hipStream_t st);
hipError_t launch_ion_turn(const IonTurnParams &p, int nspec, int64_t wave_bound, hipStream_t st);
hipError_t launch_ion_finish(const IonFinishParams &p, hipStream_t st);
// wave_bound: a host bound of the segment's waves
hipError_t launch_ion_compact(const IonCompactParams &p, int64_t wave_bound, hipStream_t st);
constexpr int kBlMaxLds = 4096;  // beam-loaded cavities x (bins + 1): the per-bin kicks in LDS (32 KB)
constexpr int kIonTurnMaxBlocks = 4096;  // grid bound of an ion segment or compaction (blocks stride over the waves)
#ifndef CTS_ION_REPLICAS
#define CTS_ION_REPLICAS 8
#endif
constexpr int kIonReplicas = CTS_ION_REPLICAS;  // copies of the passage sums
#ifndef CTS_ION_SEGMENT
#define CTS_ION_SEGMENT 32
#endif
constexpr int kIonSegment = CTS_ION_SEGMENT;  // bunch passages per ion segment
#ifndef CTS_ION_COHORT
#define CTS_ION_COHORT 4
#endif
#ifndef CTS_ION_WAVES_PER_EU
#define CTS_ION_WAVES_PER_EU 4
#endif
constexpr int kIonCohort = CTS_ION_COHORT;      // waves of an ion-turn block, walking the passages together
constexpr int kIonBlock = 64 * kIonCohort;
constexpr int kIonWavesPerEU = CTS_ION_WAVES_PER_EU;  // occupancy the ion-turn kernel is compiled for
hipError_t launch_qe_debug(const float2 *tab, uint64_t seed, uint32_t turn, uint64_t base, int64_t n, double *out,
                           hipStream_t st);
hipError_t upload_faddeeva_tables(const FadTables &t);
hipError_t launch_eigen_sums(const double *x, const double *px, const double *y, const double *py, const double *dp,
                             const int64_t *bunch_off, int nb, int nchunk, const double eta[4], double *out,
                             hipStream_t st);
hipError_t launch_w_debug(const double *re, const double *im, int64_t n, double *wr, double *wi, hipStream_t st);
hipError_t launch_kick_field_debug(const double *in, int64_t n, double *out, hipStream_t st);
hipError_t upload_faddeeva_tables_ions(const FadTables &t);
void sur_export(const cts_surrogate *s, const float **params, SurNorm *norm);

}  // namespace cts
