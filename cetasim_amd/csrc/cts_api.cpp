// cts_api.cpp -- C ABI of the MI355X tracking turn (include/cetasim_hip.h).
//
// The handle owns the device copy of every bunch, the segment table that
// assigns contiguous particle chunks to workgroups, the per-bunch wake
// state and the turn-by-turn moment records.  Each entry point maps to one
// MPBeam seam of the reference (src/MPBeam.cpp:1010-1173).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "cts_internal.hpp"

using namespace cts;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(CTS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

SparseMat compress(const double m[36])
{
    // structural pattern of B1H1 / InvH1InvB1 / InvH2InvB2
    static const int idx[16] = {0 * 6 + 0, 0 * 6 + 5, 1 * 6 + 0, 1 * 6 + 1, 1 * 6 + 5, 2 * 6 + 2,
                                2 * 6 + 5, 3 * 6 + 2, 3 * 6 + 3, 3 * 6 + 5, 4 * 6 + 0, 4 * 6 + 1,
                                4 * 6 + 2, 4 * 6 + 3, 4 * 6 + 4, 5 * 6 + 5};
    SparseMat s;
    for (int i = 0; i < 16; i++) s.m[i] = m[idx[i]];
    return s;
}

bool pattern_ok(const double m[36])
{
    static const int allowed[6][6] = {{1, 0, 0, 0, 0, 1}, {1, 1, 0, 0, 0, 1}, {0, 0, 1, 0, 0, 1},
                                      {0, 0, 1, 1, 0, 1}, {1, 1, 1, 1, 1, 0}, {0, 0, 0, 0, 0, 1}};
    for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++)
            if (!allowed[r][c] && m[r * 6 + c] != 0.0) return false;
    return true;
}

struct Timer {
    double total_ms = 0;
    int64_t n = 0;
};

}  // namespace

// The ion cloud of one interaction point: two buffers of cap slots per
// species (x, y, vx, vy, alive); the compaction after each segment of
// passages swaps them.
struct IonIp {
    cts_ion_config cfg{};
    int nspec = 0;
    int buf = 0;
    int64_t cap = 0;
    int64_t len_ub = 0;  // host upper bound of the slots in use at the phase's start
    int64_t wstride = 0;
    double *d_ion[2][CTS_MAX_SPECIES][4] = {};
    uint8_t *d_alive[2][CTS_MAX_SPECIES] = {};
    int64_t *d_len = nullptr;     // [3][CTS_MAX_SPECIES]: the two buffers, ions at the phase's start
    int64_t *d_nalive = nullptr;  // [CTS_MAX_SPECIES]
    double *d_mean = nullptr;     // [2 * CTS_MAX_SPECIES]
    double *d_smax = nullptr, *d_q = nullptr;
    double *d_cloud = nullptr;
    double *d_coeff = nullptr, *d_dt = nullptr;  // [nb][CTS_MAX_SPECIES], [nb]
    unsigned long long *d_acc = nullptr;  // kIonReplicas + 1 copies of [nspec][nb][kIonAcc][3]
    uint32_t *d_wcnt = nullptr;  // [nspec][wstride]: survivors per wave of a segment
    uint32_t *d_wofs = nullptr;  // [wstride]: their prefix sums
    std::vector<double> q;       // host copy of d_q: [nb][nspec] macro-ion charge per passage
};

struct cts_handle {
    cts_ring ring{};
    std::vector<cts_section> secs;
    int dev = 0;
    int ncu = 1;  // compute units of the device
    hipStream_t stream = nullptr;

    // bunches
    int nb = 0;
    int64_t ntot = 0;
    std::vector<int64_t> counts, offsets;
    std::vector<double> macro_charge, n_e;
    std::vector<int32_t> gap;
    double *d_x = nullptr, *d_px = nullptr, *d_y = nullptr, *d_py = nullptr, *d_z = nullptr, *d_dp = nullptr;
    uint8_t *d_flag = nullptr;

    // segments
    int nwg = 0, nseg = 0, wg_target = 1024;
    int tile = 0;  // 0: contiguous chunk per workgroup; > 0: tiles dealt round-robin
    int64_t *d_seg_begin = nullptr, *d_seg_end = nullptr, *d_bunch_off = nullptr, *d_bunch_count = nullptr;
    int32_t *d_seg_bunch = nullptr, *d_wg_seg = nullptr, *d_bunch_seg = nullptr, *d_seg_rank = nullptr;
    int32_t *d_seg_order = nullptr;  // inverse of d_seg_rank: the segment at each place in particle order
    double *d_seg_part = nullptr;
    double *d_chunk_part = nullptr;
    uint32_t *d_chunk_cnt = nullptr;
    int max_chunks = 1;

    // moments
    double *d_shift = nullptr;
    cts_moments *d_cur = nullptr;
    cts_moments *d_hist = nullptr;
    int hist_cap = 0;
    int last_turn = -1;  // last turn whose non-dp moments are in d_cur
    bool dp_pending = false;  // d_cur's dp part belongs to an older state
    double *d_sigz = nullptr, *d_zrange = nullptr;
    uint32_t *d_status = nullptr;
    uint32_t *h_status = nullptr;      // pinned copy, written by the stream
    hipEvent_t status_ev = nullptr;    // recorded after each copy
    bool status_inflight = false;
    int low_turn = -1;                 // first turn seen with transmission < 0.5

    // per-section constant blocks (kNConst doubles each)
    double *d_cst = nullptr;
    // tables of the excitation normals (NormalTab layout)
    float2 *d_ntab = nullptr;

    // wake
    int nbins = 0, wake_mode = 1, nres = 0;
    double grid_lo = 0, grid_hi = 0;
    std::vector<double> rs, q, omega;
    uint32_t *d_bhist = nullptr;   // two buffers of nb*nbins: turn parity
    int hist_parity = 0;
    int32_t *d_counts = nullptr;
    double *d_kick = nullptr, *d_poten = nullptr, *d_edges_prev = nullptr, *d_edges_now = nullptr;
    double *d_table = nullptr, *d_scale = nullptr;
    bool kick_pending = false;

    // synrad
    int synrad_on = 0, noise_mode = 0;
    uint64_t seed = 0;
    double *d_noise = nullptr;

    // ions: the electron-side kick parameters of the phase being run
    // ([nb][species][5]) and the bunch views it reads ([nb][4]), shared by
    // the interaction points (their phases run one after the other on the
    // stream); one IonIp per section, nspec 0 where there are no ions
    double *d_ion_e = nullptr;
    double *d_ion_bview = nullptr;
    std::vector<IonIp> ips;
    // d_cur holds the moments of the present coordinates
    bool cur_valid = false;
    // field model of the beam-ion kicks: surrogate parameters (null: exact)
    const cts_surrogate *sur = nullptr;
    int32_t sur_mode = 0;  // bit 0: ions kicked by bunches, bit 1: electrons kicked by ions
    uint32_t *d_fb_count = nullptr;  // exact-field fallback queue of the electron-side kick
    uint64_t *d_fb_list = nullptr;

    // long-range wake (cts_set_lrwake); its kick is deferred to the next
    // pass like the short-range one
    int lrw_T = 0;                      // nturns, 0: off
    LrwTable lrw;                       // host copy of the table
    double *d_lrw_table = nullptr;      // [3][np]
    double *d_lrw_ring = nullptr;       // [T][nb][3] post-map centroids by turn mod T
    double *d_lrw_cen = nullptr;        // [nb][3] this turn's
    double *d_lrw_kick = nullptr;       // [nb][3]
    double *d_lrw_ne = nullptr;         // [nb]
    int32_t *d_lrw_harm = nullptr;      // [nb]
    bool lr_pending = false;

    // broadband impedance (cts_set_bbimp)
    bool bbi_on = false;
    BbiTable bbi;
    uint32_t *d_bbi_hist = nullptr;                 // [nb][n]
    unsigned long long *d_bbi_dsum = nullptr;       // [nb][n][4] (dipolar planes)
    double *d_bbi_prof = nullptr, *d_bbi_wake = nullptr;  // [3][nb][n], [5][nb][n]
    double *d_bbi_g = nullptr, *d_bbi_scale = nullptr;    // [n], [nb]

    // beam-loaded cavities (cts_set_beam_loading)
    bool bl_on = false;
    BlTable bl;
    double *d_bl_tab = nullptr;
    // walked on the device: state, bins, kicks, feedback kicks, last turn's
    // info, and the bunches' populations and gaps
    double *d_bl_state = nullptr, *d_bl_kick = nullptr, *d_bl_fb = nullptr, *d_bl_info = nullptr;
    double *d_bl_ne = nullptr, *d_bl_mq = nullptr;
    uint32_t *d_bl_hist = nullptr;
    int32_t *d_bl_gap = nullptr;
    uint32_t bl_fbmask = 0;
    int bl_info_turn = -1;

    // partial sums of the eigen-emittance moments (cts_eigen_emit)
    double *d_eig = nullptr;
    size_t eig_cap = 0;

    // the settings of the per-bunch effects above, deep copies kept so that
    // cts_set_bunches can set them up again for a new bunch layout
    struct KeepLrw { bool on = false; cts_lrwake_config cfg{}; std::vector<double> sectors, res; } keep_lrw;
    struct KeepBbi { bool on = false; cts_bbimp_config cfg{}; std::vector<double> arr[16]; } keep_bbi;
    struct KeepBl { bool on = false; cts_beam_loading_config cfg{}; } keep_bl;

    // graph replays of two consecutive turns (cts_run_turns): the turn is
    // read from d_turn, which the graph advances by 2; one instance per
    // (wake buffer parity, ion buffer) at capture
    uint32_t *d_turn = nullptr;
    bool capturing = false;
    int cap_off = 0;  // turn offset (0, 1) inside the graph being captured
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        int parity = -1, ion_buf = -1;
        const float *sur_params = nullptr;
        SurNorm sur_norm{};
    };
    GraphSlot graphs[4];
    int64_t turn_dev_next = -1;  // value d_turn holds for the next replay (-1: unknown)

    // timing
    bool timing = false;
    std::map<std::string, Timer> timers;
    struct Pending { std::string name; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
};

namespace {

void free_lrw_state(cts_handle *h)
{
    void *ptrs[] = {h->d_lrw_table, h->d_lrw_ring, h->d_lrw_cen, h->d_lrw_kick, h->d_lrw_ne, h->d_lrw_harm};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    h->d_lrw_table = h->d_lrw_ring = h->d_lrw_cen = h->d_lrw_kick = h->d_lrw_ne = nullptr;
    h->d_lrw_harm = nullptr;
    h->lrw_T = 0;
    h->lrw = LrwTable{};
    h->lr_pending = false;
}

void free_bl_state(cts_handle *h)
{
    void *ptrs[] = {h->d_bl_tab, h->d_bl_state, h->d_bl_kick, h->d_bl_fb, h->d_bl_info,
                    h->d_bl_ne,  h->d_bl_mq,    h->d_bl_hist, h->d_bl_gap};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    h->d_bl_tab = h->d_bl_state = h->d_bl_kick = h->d_bl_fb = h->d_bl_info = h->d_bl_ne = h->d_bl_mq = nullptr;
    h->d_bl_hist = nullptr;
    h->d_bl_gap = nullptr;
    h->bl_fbmask = 0;
    h->bl_info_turn = -1;
    h->bl_on = false;
    h->bl = BlTable{};
}

void free_bbi_state(cts_handle *h)
{
    void *ptrs[] = {h->d_bbi_hist, h->d_bbi_prof, h->d_bbi_wake, h->d_bbi_g, h->d_bbi_scale, h->d_bbi_dsum};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    h->d_bbi_hist = nullptr;
    h->d_bbi_dsum = nullptr;
    h->d_bbi_prof = h->d_bbi_wake = h->d_bbi_g = h->d_bbi_scale = nullptr;
    h->bbi_on = false;
    h->bbi = BbiTable{};
}

BbiParams bbi_params(cts_handle *h)
{
    BbiParams p{};
    p.n = h->bbi.n; p.is = h->bbi.is; p.ie = h->bbi.ie; p.nb = h->nb; p.td = h->bbi.td; p.nw = h->bbi.nw;
    p.zmin = h->bbi.zmin; p.dz = h->bbi.dz; p.half = h->bbi.half; p.bx = h->bbi.bx; p.by = h->bbi.by;
    for (int q = 0; q < 5; q++) p.planes[q] = h->bbi.planes[q];
    p.energy = h->ring.energy; p.beta2 = std::pow(h->ring.rbeta, 2);
    p.seg_begin = h->d_seg_begin; p.seg_end = h->d_seg_end;
    p.seg_bunch = h->d_seg_bunch; p.wg_seg = h->d_wg_seg;
    p.x = h->d_x; p.y = h->d_y; p.z = h->d_z;
    p.px = h->d_px; p.py = h->d_py; p.dp = h->d_dp; p.flag = h->d_flag;
    p.hist = h->d_bbi_hist; p.dsum = h->d_bbi_dsum; p.prof = h->d_bbi_prof; p.wake = h->d_bbi_wake;
    p.g = h->d_bbi_g; p.scale = h->d_bbi_scale;
    for (int j = 0; j < kBbiTaps; j++) p.taps[j] = h->bbi.taps[j];
    return p;
}

void free_bunch_state(cts_handle *h)
{
    free_lrw_state(h);
    free_bl_state(h);
    free_bbi_state(h);
    void *ptrs[] = {h->d_x, h->d_px, h->d_y, h->d_py, h->d_z, h->d_dp, h->d_flag,
                    h->d_seg_begin, h->d_seg_end, h->d_bunch_off, h->d_bunch_count,
                    h->d_seg_bunch, h->d_wg_seg, h->d_bunch_seg, h->d_seg_part, h->d_shift,
                    h->d_cur, h->d_hist, h->d_sigz, h->d_zrange, h->d_noise, h->d_ion_e, h->d_ion_bview,
                    h->d_bhist, h->d_counts, h->d_kick, h->d_poten, h->d_edges_prev,
                    h->d_edges_now, h->d_table, h->d_scale, h->d_chunk_part, h->d_chunk_cnt,
                    h->d_fb_count, h->d_fb_list, h->d_eig, h->d_seg_rank, h->d_seg_order};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    h->d_eig = nullptr;
    h->d_seg_rank = h->d_seg_order = nullptr;
    h->eig_cap = 0;
    h->d_x = h->d_px = h->d_y = h->d_py = h->d_z = h->d_dp = nullptr;
    h->d_flag = nullptr;
    h->d_seg_begin = h->d_seg_end = h->d_bunch_off = h->d_bunch_count = nullptr;
    h->d_seg_bunch = h->d_wg_seg = h->d_bunch_seg = nullptr;
    h->d_seg_part = h->d_shift = h->d_sigz = h->d_zrange = h->d_noise = h->d_ion_e = h->d_ion_bview = nullptr;
    h->d_cur = h->d_hist = nullptr;
    h->d_bhist = nullptr;
    h->d_counts = nullptr;
    h->d_kick = h->d_poten = h->d_edges_prev = h->d_edges_now = h->d_table = h->d_scale = nullptr;
    h->d_chunk_part = nullptr;
    h->d_chunk_cnt = nullptr;
    h->d_fb_count = nullptr;
    h->d_fb_list = nullptr;
}

template <typename T>
int dalloc(T **p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) return fail(CTS_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemset(*p, 0, n * sizeof(T));
    if (e != hipSuccess) return fail(CTS_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    return CTS_OK;
}

hipEvent_t get_event(cts_handle *h)
{
    if (!h->event_pool.empty()) {
        hipEvent_t e = h->event_pool.back();
        h->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void resolve_timers(cts_handle *h)
{
    for (auto &p : h->pending) {
        float ms = 0;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            Timer &t = h->timers[p.name];
            t.total_ms += ms;
            t.n += 1;
        }
        h->event_pool.push_back(p.a);
        h->event_pool.push_back(p.b);
    }
    h->pending.clear();
}

struct TimedScope {
    cts_handle *h;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    TimedScope(cts_handle *h_, const char *n) : h(h_), name(n)
    {
        if (h->timing) {
            if (h->pending.size() > 4096) resolve_timers(h);
            a = get_event(h);
            b = get_event(h);
            if (a) (void)hipEventRecord(a, h->stream);
        }
    }
    ~TimedScope()
    {
        if (h->timing && a && b) {
            (void)hipEventRecord(b, h->stream);
            h->pending.push_back({name, a, b});
        }
    }
};

}  // namespace

namespace cts {
// sc[j] = {sin, cos}(2 pi j / 256) and lg[j] = {ln(1 + j/256), 256/(256 + j)},
// each the float rounding of the double value
void build_normal_tables(float2 *out)
{
    for (int j = 0; j < kNormTab; j++) {
        const double a = 6.283185307179586 * (double)j / 256.0;
        out[j].x = (float)std::sin(a);
        out[j].y = (float)std::cos(a);
        out[kNormTab + j].x = (float)std::log1p((double)j / 256.0);
        out[kNormTab + j].y = (float)(256.0 / (256.0 + (double)j));
    }
}
}  // namespace cts

namespace {

// Constant block of section k (layout CI_* in cts_internal.hpp).
void build_consts(const cts_ring &R, const std::vector<cts_section> &secs, int k, double *c)
{
    std::memset(c, 0, kNConst * sizeof(double));
    const cts_section &S = secs[k];
    const SparseMat b1h1 = compress(S.b1h1), h2b2 = compress(S.h2b2);
    const SparseMat srb = compress(secs[0].b1h1), srh = compress(secs[0].h1b1);
    for (int i = 0; i < 16; i++) {
        c[CI_B1H1 + i] = b1h1.m[i];
        c[CI_H2B2 + i] = h2b2.m[i];
        c[CI_SRB + i] = srb.m[i];
        c[CI_SRH + i] = srh.m[i];
    }
    c[CI_ETAX] = S.eta_x; c[CI_ETAPX] = S.etap_x; c[CI_ETAY] = S.eta_y; c[CI_ETAPY] = S.etap_y;
    c[CI_ALPHAX] = S.alpha_x; c[CI_BETAX] = S.beta_x; c[CI_ALPHAY] = S.alpha_y; c[CI_BETAY] = S.beta_y;
    c[CI_IBETAX] = 1.0 / S.beta_x;
    c[CI_IBETAY] = 1.0 / S.beta_y;
    c[CI_PHX] = S.phase_adv_x;
    c[CI_PHY] = S.phase_adv_y;
    c[CI_CMUX] = std::cos(S.phase_adv_x);
    c[CI_SMUX] = std::sin(S.phase_adv_x);
    c[CI_CMUY] = std::cos(S.phase_adv_y);
    c[CI_SMUY] = std::sin(S.phase_adv_y);
    // src/Bunch.cpp:658-659
    c[CI_WX] = S.phase_adv_x / (2 * kPi * R.nux);
    c[CI_WY] = S.phase_adv_y / (2 * kPi * R.nuy);
    c[CI_CHX] = R.chrom[0];
    c[CI_CHY] = R.chrom[1];
    for (int i = 0; i < 5; i++) { c[CI_ADTX + i] = R.adtx[i]; c[CI_ADTY + i] = R.adty[i]; }
    c[CI_SKEW] = R.skew_k;
    // RF phase per unit z: the reference's -z/c/beta*2*pi*h*fRF
    const double f_rf = R.f0 * R.harmonic;
    for (int r = 0; r < R.n_rf && r < CTS_MAX_RF; r++) {
        c[CI_RFVR + r] = R.rf_vr[r];
        c[CI_RFVI + r] = R.rf_vi[r];
        c[CI_RFK + r] = (((-1.0 / kClight) / R.rbeta) * 2. * kPi) * (double)R.rf_harm[r] * f_rf;
    }
    c[CI_RFSCALE] = 1.0 / R.energy / (R.rbeta * R.rbeta);
    c[CI_DRIFTC] = R.circ;
    for (int i = 0; i < 3; i++) c[CI_ALPHAC + i] = R.alphac[i];
    c[CI_U0] = R.u0 / R.energy / std::pow(R.rbeta, 2);
    // synrad constants (src/Bunch.cpp:971-980), section 0 matrices
    double lam[3];
    for (int i = 0; i < 3; i++) lam[i] = R.synrad_tau[i] > 0 ? std::exp(-1.0 / R.synrad_tau[i]) : 0.0;
    for (int i = 0; i < 3; i++) c[CI_LAM + i] = lam[i];
    c[CI_LZ2] = lam[2] * lam[2];
    c[CI_COEF + 0] = std::sqrt(1 - std::pow(lam[0], 2)) * std::sqrt(R.emit_nat[0]);
    c[CI_COEF + 1] = std::sqrt(1 - std::pow(lam[1], 2)) * std::sqrt(R.emit_nat[1]);
    c[CI_COEF + 2] = std::sqrt(1 - std::pow(lam[2], 4)) * std::sqrt(R.emit_nat[2]);
    // loss limits (src/Bunch.cpp:180-186)
    c[CI_ZLIM] = R.t0 * kClight / R.harmonic / 2;
    c[CI_APX] = R.ap_x;
    c[CI_APY] = R.ap_y;
    c[CI_IAPX] = 1.0 / R.ap_x;
    c[CI_IAPY] = 1.0 / R.ap_y;
    // angle functions of the map and the RF kick (sincos_small_k and
    // sincos_fast_k in cts_kernels.hip)
    c[CI_TWOPI] = 2 * kPi;
    c[CI_SMALL] = 0.1;
    const double ps[4] = {2.7557319223985893e-06, -0.0001984126984126984, 0.008333333333333333,
                          -0.16666666666666666};
    const double pc[4] = {-2.755731922398589e-07, 2.48015873015873e-05, -0.001388888888888889,
                          0.041666666666666664};
    const double fr[4] = {0.6366197723675814, 0x1.921fb548p+0, -0x1.de973dc8p-31, -0x1.9d9cceba3f91fp-62};
    const double fs[7] = {-7.647163731819816e-13, 1.6059043836821613e-10, -2.505210838544172e-08,
                          2.7557319223985893e-06, -0.0001984126984126984, 0.008333333333333333,
                          -0.16666666666666666};
    const double fc[7] = {-4.779477332387385e-14, 1.1470745597729725e-11, -2.08767569878681e-09,
                          2.755731922398589e-07, -2.48015873015873e-05, 0.001388888888888889,
                          -0.041666666666666664};
    for (int i = 0; i < 4; i++) { c[CI_PS + i] = ps[i]; c[CI_PC + i] = pc[i]; c[CI_FR + i] = fr[i]; }
    for (int i = 0; i < 7; i++) { c[CI_FS + i] = fs[i]; c[CI_FC + i] = fc[i]; }
}

PassParams base_pass(cts_handle *h, uint32_t ops, int k)
{
    PassParams p;
    std::memset(&p, 0, sizeof p);
    p.x = h->d_x; p.px = h->d_px; p.y = h->d_y; p.py = h->d_py; p.z = h->d_z; p.dp = h->d_dp;
    p.flag = h->d_flag;
    p.seg_begin = h->d_seg_begin; p.seg_end = h->d_seg_end;
    p.seg_bunch = h->d_seg_bunch; p.wg_seg = h->d_wg_seg;
    p.bunch_off = h->d_bunch_off;
    p.seg_part = h->d_seg_part;
    p.seg_rank = h->d_seg_rank;
    p.bunch_hist = h->d_bhist ? h->d_bhist + (size_t)h->hist_parity * h->nb * h->nbins : nullptr;
    p.shift = h->d_shift;
    p.kick_prev = h->d_kick; p.edges_prev = h->d_edges_prev; p.edges_now = h->d_edges_now;
    p.noise = h->d_noise;
    p.cst = (uint64_t)(uintptr_t)(h->d_cst + (size_t)k * kNConst);
    p.ntot = h->ntot;
    p.ops = ops;
    p.nseg = h->nseg;
    p.nbins = h->nbins;
    p.n_rf = h->ring.n_rf;
    p.noise_mode = h->noise_mode;
    p.seed = h->seed;
    p.ntab = h->d_ntab;
    p.lr_kick = h->d_lrw_kick;
    if (h->bl_on) {
        p.bl_tab = h->d_bl_tab;
        p.zrange = h->d_zrange;
        p.bl_nbl = h->bl.nbl;
        p.bl_nbins = h->bl.nbins;
        p.bl_nturns = h->bl.nturns;
        p.nb = h->nb;
        for (int q = 0; q < h->bl.nbl; q++) p.bl_kfac[q] = h->bl.kfac[q];
        p.bl_energy = h->ring.energy;
        p.bl_beta2 = std::pow(h->ring.rbeta, 2);
        if (h->bl.walk) {
            p.bl_walk = h->d_bl_kick;
            p.bl_fbk = h->bl_fbmask ? h->d_bl_fb : nullptr;
            p.bl_fbmask = h->bl_fbmask;
            p.bl_hist = h->d_bl_hist;
        }
    }
    if (h->capturing) {
        p.turn_dev = h->d_turn;
        p.turn = (uint32_t)h->cap_off;
    }
    return p;
}

FinParams base_fin(cts_handle *h, uint32_t ops, int slot, int prev_slot)
{
    FinParams f;
    std::memset(&f, 0, sizeof f);
    f.seg_part = h->d_seg_part;
    f.bunch_seg = h->d_bunch_seg;
    f.bunch_count = h->d_bunch_count;
    f.shift = h->d_shift;
    f.cur = h->d_cur;
    f.hist = h->d_hist;
    f.slot = slot;
    f.prev_slot = prev_slot;
    f.nb = h->nb;
    f.nbins = h->nbins;
    f.ops = ops;
    f.sigz_stale = h->d_sigz;
    f.zrange = h->d_zrange;
    f.ref_edges = h->wake_mode == 0 ? 1 : 0;
    f.edges_now = h->d_edges_now;
    f.status = h->d_status;
    f.nseg = h->nseg;
    f.chunk_part = h->d_chunk_part;
    f.chunk_cnt = h->d_chunk_cnt;
    f.max_chunks = h->max_chunks;
    f.lr_cen = h->d_lrw_cen;
    if (h->capturing) {
        f.turn_dev = h->d_turn;
        f.turn_off = h->cap_off;
        f.hist_cap = h->hist_cap;
    }
    return f;
}

int run_pass(cts_handle *h, const PassParams &p, const char *name = "pass")
{
    // host-supplied normals are read through p.noise: refuse the launch
    // rather than hand the kernel a null buffer
    if ((p.ops & OP_SYNRAD) && p.noise_mode == 1 && !p.noise)
        return fail(CTS_ERR_STATE, "synrad with noise_mode 1 but no noise buffer (call cts_set_synrad again)");
    TimedScope ts(h, name);
    HIPCHK(launch_pass(p, h->nwg, h->stream));
    return CTS_OK;
}

int run_finalize(cts_handle *h, const FinParams &f)
{
    TimedScope ts(h, "finalize");
    HIPCHK(launch_finalize(f, h->stream));
    return CTS_OK;
}

int slot_of(cts_handle *h, int turn) { return (h->hist_cap > 0 && turn >= 0) ? turn % h->hist_cap : -1; }

// Wake stage after the histogram is complete: reference-edge mode needs
// the per-bunch table for this turn's bin width first.
WakeParams wake_params(cts_handle *h)
{
    WakeParams w;
    std::memset(&w, 0, sizeof w);
    w.bunch_hist = h->d_bhist + (size_t)h->hist_parity * h->nb * h->nbins;
    w.hist_clear = h->d_bhist + (size_t)(1 - h->hist_parity) * h->nb * h->nbins;
    w.counts_out = h->d_counts;
    w.table = h->d_table;
    w.per_bunch_table = h->wake_mode == 0 ? 1 : 0;
    w.scale = h->d_scale;
    w.poten = h->d_poten;
    w.kick = h->d_kick;
    w.edges_now = h->d_edges_now;
    w.edges_prev = h->d_edges_prev;
    w.nbins = h->nbins;
    return w;
}

int run_wake(cts_handle *h)
{
    if (h->wake_mode == 0) {
        WakeTableParams t;
        std::memset(&t, 0, sizeof t);
        t.edges = h->d_edges_now;
        t.table = h->d_table;
        t.nbins = h->nbins;
        t.nres = h->nres;
        for (int r = 0; r < h->nres; r++) { t.rs[r] = h->rs[r]; t.q[r] = h->q[r]; t.omega[r] = h->omega[r]; }
        TimedScope ts(h, "wake_table");
        HIPCHK(launch_wake_table(t, h->nb, h->stream));
    }
    WakeParams w = wake_params(h);
    TimedScope ts(h, "wake");
    HIPCHK(launch_wake(w, h->nb, h->stream));
    h->hist_parity = 1 - h->hist_parity;
    return CTS_OK;
}

int run_bin(cts_handle *h)
{
    BinParams b;
    b.z = h->d_z;
    b.seg_begin = h->d_seg_begin; b.seg_end = h->d_seg_end;
    b.seg_bunch = h->d_seg_bunch; b.wg_seg = h->d_wg_seg;
    b.edges = h->d_edges_now;
    b.bunch_hist = h->d_bhist + (size_t)h->hist_parity * h->nb * h->nbins;
    b.nbins = h->nbins;
    TimedScope ts(h, "bin");
    HIPCHK(launch_bin(b, h->nwg, h->stream));
    return CTS_OK;
}

// Apply a pending wake kick; the dp moments of the state it completes go
// into the current record (and its history slot).
int flush_kick(cts_handle *h)
{
    if (!h->kick_pending && !h->lr_pending) return CTS_OK;
    uint32_t ops = OP_WRITE;
    if (h->kick_pending) ops |= OP_KICK | OP_MOMB;
    if (h->lr_pending) ops |= OP_LRK;
    PassParams p = base_pass(h, ops, 0);
    int rc = run_pass(h, p, "kick");
    if (rc) return rc;
    if (h->kick_pending) {
        FinParams f = base_fin(h, OP_MOMB, -1, slot_of(h, h->last_turn));
        rc = run_finalize(h, f);
        if (rc) return rc;
    }
    h->kick_pending = false;
    h->lr_pending = false;
    h->dp_pending = false;
    return CTS_OK;
}

int check_ready(cts_handle *h)
{
    if (!h) return fail(CTS_ERR_ARG, "null handle");
    if (h->nb == 0) return fail(CTS_ERR_STATE, "no bunches: call cts_set_bunches first");
    return CTS_OK;
}

// Transmission status: the finalize kernel sets a sticky device flag; a
// pinned copy is refreshed asynchronously once per turn and polled without
// blocking, so the turn loop never waits on the host.
int enqueue_status_copy(cts_handle *h)
{
    if (h->status_inflight) return CTS_OK;
    HIPCHK(hipMemcpyAsync(h->h_status, h->d_status, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipEventRecord(h->status_ev, h->stream));
    h->status_inflight = true;
    return CTS_OK;
}

int check_status(cts_handle *h, bool block)
{
    if (h->status_inflight) {
        if (block) {
            HIPCHK(hipEventSynchronize(h->status_ev));
        } else {
            const hipError_t q = hipEventQuery(h->status_ev);
            if (q == hipErrorNotReady) return CTS_OK;
            if (q != hipSuccess) return fail(CTS_ERR_HIP, std::string("status event: ") + hipGetErrorString(q));
        }
        h->status_inflight = false;
    }
    if (*h->h_status & 1u)
        return fail(CTS_LOW_TRANSMISSION, "more than 50% particles are lost in a bunch (src/Bunch.cpp:193-198)");
    if (*h->h_status & 2u)
        return fail(CTS_ION_OVERFLOW, "the ion cloud would exceed its maximum macro-ion number "
                                      "(src/LatticeInterActionPoint.cpp:1041-1049)");
    if (*h->h_status & 4u)
        return fail(CTS_LRWAKE_TAU, "long-range wake: a source bunch lies ahead of its receiver in time "
                                    "(tau > 0, src/WakeFunction.cpp:206-210)");
    return CTS_OK;
}

int reset_status(cts_handle *h)
{
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemset(h->d_status, 0, sizeof(uint32_t)));
    *h->h_status = 0;
    h->status_inflight = false;
    return CTS_OK;
}

// Segment table.  tile == 0: the particle range cut into wg_target
// contiguous chunks (multiples of 64); tile > 0: cut into tiles of `tile`
// particles dealt round-robin to min(wg_target, tiles) workgroups, so the
// workgroups running at any moment stream one compact window of the arrays
// (as a grid-stride loop does).  Every piece is cut again at bunch
// boundaries.  Workgroup w owns the segments wg_seg[w] .. wg_seg[w+1]; the
// segments are numbered in workgroup order, and seg_rank[s] is segment s's
// place in particle order, where each bunch's segments are consecutive
// (bunch_seg) and its partial sums are folded (seg_part is indexed by rank).
// (Re)allocates everything sized by the number of segments.
int layout_segments(cts_handle *h, int wg_target, int tile)
{
    const int nb = h->nb;
    const int64_t n = h->ntot;
    int64_t piece;
    int nwg;
    if (tile > 0) {
        piece = ((int64_t)tile + kBlock - 1) / kBlock * kBlock;
        const int64_t ntile = std::max<int64_t>(1, (n + piece - 1) / piece);
        nwg = (int)std::min<int64_t>(std::max(wg_target, 1), ntile);
    } else {
        int64_t want = (n + 2047) / 2048;
        if (want < 1) want = 1;
        if (want > wg_target) want = wg_target;
        piece = (n + want - 1) / want;
        piece = (piece + 63) / 64 * 64;
        if (piece < 64) piece = 64;
        nwg = (int)((n + piece - 1) / piece) > 0 ? (int)((n + piece - 1) / piece) : 1;
    }
    const int64_t npiece = std::max<int64_t>(1, (n + piece - 1) / piece);
    // segments in particle order, with the piece each belongs to
    std::vector<int64_t> pb, pe;
    std::vector<int32_t> pbunch, ppiece;
    {
        int b = 0;
        for (int64_t t = 0; t < npiece; t++) {
            const int64_t lo = t * piece, hi = std::min<int64_t>(n, lo + piece);
            int64_t p = lo;
            while (p < hi) {
                while (b < nb && h->offsets[b + 1] <= p) b++;
                const int64_t e2 = std::min<int64_t>(hi, h->offsets[b + 1]);
                pb.push_back(p);
                pe.push_back(e2);
                pbunch.push_back(b);
                ppiece.push_back((int32_t)t);
                p = e2;
            }
        }
    }
    const int nseg = (int)pb.size();
    std::vector<int32_t> bseg(nb + 1, 0);
    {
        int s = 0;
        for (int bb = 0; bb < nb; bb++) {
            while (s < nseg && pbunch[s] < bb) s++;
            bseg[bb] = s;
        }
        bseg[nb] = nseg;
    }
    // workgroup order: piece t goes to workgroup t % nwg (chunked: t == w)
    std::vector<int32_t> first_seg(npiece + 1, nseg);
    for (int s = nseg - 1; s >= 0; s--) first_seg[ppiece[s]] = s;
    for (int64_t t = npiece - 1; t >= 0; t--)
        if (first_seg[t] == nseg) first_seg[t] = first_seg[t + 1];
    std::vector<int64_t> sb, se;
    std::vector<int32_t> sbunch, rank, wgseg(nwg + 1, 0);
    sb.reserve(nseg); se.reserve(nseg); sbunch.reserve(nseg); rank.reserve(nseg);
    for (int w = 0; w < nwg; w++) {
        wgseg[w] = (int32_t)sb.size();
        for (int64_t t = w; t < npiece; t += nwg)
            for (int s = first_seg[t]; s < first_seg[t + 1]; s++) {
                sb.push_back(pb[s]);
                se.push_back(pe[s]);
                sbunch.push_back(pbunch[s]);
                rank.push_back(s);
            }
    }
    wgseg[nwg] = (int32_t)sb.size();
    bool identity = true;
    for (int s = 0; s < nseg && identity; s++) identity = rank[s] == s;
    int max_chunks = 1;
    for (int bb = 0; bb < nb; bb++) max_chunks = std::max(max_chunks, (bseg[bb + 1] - bseg[bb] + 63) / 64);
    void *old[] = {h->d_seg_begin, h->d_seg_end, h->d_seg_bunch, h->d_wg_seg, h->d_bunch_seg, h->d_seg_part,
                   h->d_chunk_part, h->d_seg_rank, h->d_seg_order};
    for (void *p : old)
        if (p) (void)hipFree(p);
    h->d_seg_begin = h->d_seg_end = nullptr;
    h->d_seg_bunch = h->d_wg_seg = h->d_bunch_seg = h->d_seg_rank = h->d_seg_order = nullptr;
    h->d_seg_part = h->d_chunk_part = nullptr;
    int rc;
    if ((rc = dalloc(&h->d_seg_begin, nseg)) || (rc = dalloc(&h->d_seg_end, nseg)) ||
        (rc = dalloc(&h->d_seg_bunch, nseg)) || (rc = dalloc(&h->d_wg_seg, nwg + 1)) ||
        (rc = dalloc(&h->d_bunch_seg, nb + 1)) || (rc = dalloc(&h->d_seg_part, (size_t)nseg * kNPart)) ||
        (rc = dalloc(&h->d_chunk_part, (size_t)nb * max_chunks * kNPart)) ||
        (!identity && ((rc = dalloc(&h->d_seg_rank, nseg)) || (rc = dalloc(&h->d_seg_order, nseg)))))
        return rc;
    if (!identity) {
        std::vector<int32_t> order(nseg);
        for (int sg = 0; sg < nseg; sg++) order[rank[sg]] = sg;
        HIPCHK(hipMemcpy(h->d_seg_rank, rank.data(), rank.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->d_seg_order, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(h->d_seg_begin, sb.data(), sb.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_seg_end, se.data(), se.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_seg_bunch, sbunch.data(), sbunch.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_wg_seg, wgseg.data(), wgseg.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_bunch_seg, bseg.data(), bseg.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    h->nwg = nwg;
    h->nseg = nseg;
    h->max_chunks = max_chunks;
    h->tile = tile;
    return CTS_OK;
}

// Moments of the present coordinates into d_cur (MPBeamRMSCal): the first
// pass fixes the shift (means), the second takes the centred sums about it.
int enqueue_moments(cts_handle *h)
{
    int rc = flush_kick(h);
    if (rc) return rc;
    for (int it = 0; it < 2; it++) {
        PassParams p = base_pass(h, OP_MOMA | OP_MOMB | OP_ZSTAT, 0);
        if ((rc = run_pass(h, p, "moments"))) return rc;
        FinParams f = base_fin(h, OP_MOMB | OP_MOMA | OP_ZSTAT, -1, -1);
        if ((rc = run_finalize(h, f))) return rc;
    }
    h->cur_valid = true;
    return CTS_OK;
}

void free_ion_ip(IonIp &I)
{
    for (int b = 0; b < 2; b++)
        for (int p = 0; p < CTS_MAX_SPECIES; p++) {
            for (int c = 0; c < 4; c++)
                if (I.d_ion[b][p][c]) (void)hipFree(I.d_ion[b][p][c]);
            if (I.d_alive[b][p]) (void)hipFree(I.d_alive[b][p]);
        }
    void *ptrs[] = {I.d_len, I.d_nalive, I.d_mean, I.d_smax, I.d_q, I.d_cloud, I.d_coeff, I.d_dt, I.d_acc,
                    I.d_wcnt, I.d_wofs};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    I = IonIp{};
}

void free_ion_state(cts_handle *h)
{
    for (IonIp &I : h->ips) free_ion_ip(I);
}

bool ions_anywhere(const cts_handle *h)
{
    for (const IonIp &I : h->ips)
        if (I.nspec > 0) return true;
    return false;
}


// Drop the captured turn graphs (their kernel arguments name buffers and
// settings that are about to change).
void graph_reset(cts_handle *h)
{
    for (auto &g : h->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = cts_handle::GraphSlot{};
    }
    h->turn_dev_next = -1;
}

// The cloud buffer after one ion phase: one compaction per segment.
int ion_buf_after_turn(const cts_handle *h)
{
    const IonIp &I = h->ips[0];
    const int segs = (h->nb + kIonSegment - 1) / kIonSegment;
    return (I.nspec > 0 && segs % 2 == 1) ? 1 - I.buf : I.buf;
}

// MPBeam::SSBeamIonEffectOneInteractionPoint at interaction point k
// (src/MPBeam.cpp:2585-2611): bunch views from the moments of the present
// state with section k's beta functions (MPBeamRMSCal(k)), the bunch
// passages in train order (segments of kIonSegment passages, each followed
// by a compaction of the cloud), the passage moments, then the electron
// kicks of every bunch.
int ion_phase(cts_handle *h, int32_t turn, int k)
{
    int rc;
    if (!h->cur_valid && (rc = enqueue_moments(h))) return rc;
    IonIp &Ip = h->ips[k];
    const cts_ion_config &I = Ip.cfg;
    const cts_section &S = h->secs[k];
    const int64_t nacc = (int64_t)Ip.nspec * h->nb * kIonAcc * 3;
    int64_t *len0 = Ip.d_len + 2 * CTS_MAX_SPECIES;
    {
        TimedScope ts(h, "ion_prep");
        HIPCHK(launch_ion_prep(h->d_cur, h->nb, S.beta_x, S.beta_y, h->d_ion_bview, Ip.d_smax, h->d_fb_count,
                               Ip.d_len + Ip.buf * CTS_MAX_SPECIES, len0, Ip.nspec, h->stream));
    }
    IonTurnParams P;
    std::memset(&P, 0, sizeof P);
    P.bview = h->d_ion_bview;
    P.size_max = Ip.d_smax;
    P.coeffI = Ip.d_coeff;
    P.dt = Ip.d_dt;
    P.acc = Ip.d_acc;
    P.acc_stride = nacc;
    P.wcnt = Ip.d_wcnt;
    P.wcnt_stride = Ip.wstride;
    P.sur_params = nullptr;
    if (h->sur && (h->sur_mode & 1)) sur_export(h->sur, &P.sur_params, &P.sur_norm);
    P.loss_boundary = I.loss_boundary;
    P.ap_x = I.ap_x;
    P.ap_y = I.ap_y;
    P.cap = Ip.cap;
    P.seed = I.seed;
    P.nb = h->nb;
    P.n_new = I.n_new;
    P.w_new = (I.n_new + 63) / 64;
    P.turn = (uint32_t)turn;
    if (h->capturing) {
        P.turn_dev = h->d_turn;
        P.turn = (uint32_t)h->cap_off;
    }
    // segments of kIonSegment passages; after each, the cloud (survivors and
    // the segment's new ions, in slot order) is compacted into the other
    // buffer.  ub: the host's bound of the ions at the segment's start (the
    // capacity in a graph, which replays any turn).
    int64_t ub = h->capturing ? Ip.cap : Ip.len_ub;
    for (int ja = 0; ja < h->nb; ja += kIonSegment) {
        const int jb = std::min(h->nb, ja + kIonSegment);
        const int buf = Ip.buf;
        for (int p = 0; p < Ip.nspec; p++) {
            P.x[p] = Ip.d_ion[buf][p][0];
            P.y[p] = Ip.d_ion[buf][p][1];
            P.vx[p] = Ip.d_ion[buf][p][2];
            P.vy[p] = Ip.d_ion[buf][p][3];
            P.alive[p] = Ip.d_alive[buf][p];
        }
        P.len = Ip.d_len + buf * CTS_MAX_SPECIES;
        P.j_begin = ja;
        P.j_end = jb;
        const int64_t waves = (ub + 63) / 64 + (int64_t)(jb - ja) * P.w_new;
        {
            TimedScope ts(h, "ion_turn");
            HIPCHK(launch_ion_turn(P, Ip.nspec, waves, h->stream));
        }
        ub = std::min<int64_t>(ub + (int64_t)(jb - ja) * I.n_new, Ip.cap);
        for (int p = 0; p < Ip.nspec; p++) {
            IonCompactParams c;
            c.x = Ip.d_ion[buf][p][0]; c.y = Ip.d_ion[buf][p][1];
            c.vx = Ip.d_ion[buf][p][2]; c.vy = Ip.d_ion[buf][p][3];
            c.alive = Ip.d_alive[buf][p];
            c.ox = Ip.d_ion[1 - buf][p][0]; c.oy = Ip.d_ion[1 - buf][p][1];
            c.ovx = Ip.d_ion[1 - buf][p][2]; c.ovy = Ip.d_ion[1 - buf][p][3];
            c.oalive = Ip.d_alive[1 - buf][p];
            c.len = Ip.d_len + buf * CTS_MAX_SPECIES + p;
            c.new_len = Ip.d_len + (1 - buf) * CTS_MAX_SPECIES + p;
            c.wcnt = Ip.d_wcnt + p * Ip.wstride;
            c.wofs = Ip.d_wofs;
            c.cap = Ip.cap;
            c.n_new = I.n_new;
            c.w_new = P.w_new;
            c.nseg = jb - ja;
            TimedScope ts(h, "ion_compact");
            HIPCHK(launch_ion_compact(c, waves, h->stream));
        }
        Ip.buf = 1 - buf;
    }
    Ip.len_ub = ub;
    {
        TimedScope ts(h, "ion_reduce");
        HIPCHK(launch_ion_reduce(Ip.d_acc, nacc, Ip.d_acc + kIonReplicas * nacc, h->stream));
    }
    IonFinishParams F;
    F.acc = Ip.d_acc + kIonReplicas * nacc;
    F.bview = h->d_ion_bview;
    F.ion_q = Ip.d_q;
    F.len0 = len0;
    F.n_alive = Ip.d_nalive;
    F.mean_state = Ip.d_mean;
    F.ion_e = h->d_ion_e;
    F.cloud = Ip.d_cloud;
    F.status = h->d_status;
    F.gamma = h->ring.rgamma;
    F.max_ions = I.max_ions;
    F.nb = h->nb;
    F.nspec = Ip.nspec;
    F.n_new = I.n_new;
    {
        TimedScope ts(h, "ion_finish");
        HIPCHK(launch_ion_finish(F, h->stream));
    }
    // electron side (src/MPBunch.cpp:1165-1206, src/Bunch.cpp:213-220)
    IonKickParams ik;
    ik.x = h->d_x; ik.y = h->d_y; ik.px = h->d_px; ik.py = h->d_py; ik.flag = h->d_flag;
    ik.seg_begin = h->d_seg_begin; ik.seg_end = h->d_seg_end;
    ik.seg_bunch = h->d_seg_bunch; ik.wg_seg = h->d_wg_seg;
    ik.ion_e = h->d_ion_e;
    ik.nspec = Ip.nspec;
    ik.sur_params = nullptr;
    if (h->sur && (h->sur_mode & 2)) sur_export(h->sur, &ik.sur_params, &ik.sur_norm);
    ik.fb_count = h->d_fb_count;
    ik.fb_list = h->d_fb_list;
    ik.bunch_off = h->d_bunch_off;
    ik.nb = h->nb;
    ik.nseg = h->nseg;
    ik.seg_order = h->d_seg_order;
    TimedScope ts(h, "ion_kick");
    HIPCHK(launch_ion_kick(ik, 4 * h->ncu, h->stream));
    h->cur_valid = false;
    return CTS_OK;
}

}  // namespace

extern "C" {

const char *cts_last_error(void) { return g_err.c_str(); }

int cts_device_info(char *buf, int32_t len)
{
    int dev = 0;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, (size_t)len, "%s %s CUs=%d", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return CTS_OK;
}

int cts_create(const cts_ring *ring, const cts_section *secs, int32_t nsec, cts_handle **out)
{
    if (!ring || !secs || nsec < 1 || !out) return fail(CTS_ERR_ARG, "cts_create: bad arguments");
    if (ring->n_rf < 0 || ring->n_rf > CTS_MAX_RF) return fail(CTS_ERR_ARG, "cts_create: n_rf out of range");
    for (int k = 0; k < nsec; k++) {
        if (!pattern_ok(secs[k].b1h1) || !pattern_ok(secs[k].h1b1) || !pattern_ok(secs[k].h2b2))
            return fail(CTS_ERR_ARG, "cts_create: section matrix outside the B*H structure");
    }
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(CTS_ERR_HIP, "no HIP device");
    cts_handle *h = new cts_handle();
    h->ring = *ring;
    h->secs.assign(secs, secs + nsec);
    h->ips.resize(nsec);
    hipError_t e = hipGetDevice(&h->dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, h->dev);
    if (e == hipSuccess) e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) {
        delete h;
        return fail(CTS_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
    }
    // Weideman's coefficients of the kick field's w(z)
    FadTables ft;
    build_fad_tables(ft);
    e = upload_faddeeva_tables(ft);
    if (e == hipSuccess) e = upload_faddeeva_tables_ions(ft);
    if (e == hipSuccess) e = hipMalloc((void **)&h->d_status, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(h->d_status, 0, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&h->d_turn, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&h->d_ntab, 2 * kNormTab * sizeof(float2));
    if (e == hipSuccess) {
        std::vector<float2> tab(2 * kNormTab);
        build_normal_tables(tab.data());
        e = hipMemcpy(h->d_ntab, tab.data(), tab.size() * sizeof(float2), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemset(h->d_turn, 0, sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc((void **)&h->h_status, sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->status_ev, hipEventDisableTiming);
    if (e == hipSuccess) {
        *h->h_status = 0;
        std::vector<double> cst((size_t)nsec * kNConst);
        for (int k = 0; k < nsec; k++) build_consts(h->ring, h->secs, k, cst.data() + (size_t)k * kNConst);
        e = hipMalloc((void **)&h->d_cst, cst.size() * sizeof(double));
        if (e == hipSuccess) e = hipMemcpy(h->d_cst, cst.data(), cst.size() * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        (void)hipStreamDestroy(h->stream);
        delete h;
        return fail(CTS_ERR_HIP, std::string("init: ") + hipGetErrorString(e));
    }
    *out = h;
    return CTS_OK;
}

int cts_destroy(cts_handle *h)
{
    if (!h) return CTS_OK;
    (void)hipStreamSynchronize(h->stream);
    resolve_timers(h);
    graph_reset(h);
    for (hipEvent_t e : h->event_pool) (void)hipEventDestroy(e);
    if (h->d_turn) (void)hipFree(h->d_turn);
    if (h->d_ntab) (void)hipFree(h->d_ntab);
    free_bunch_state(h);
    free_ion_state(h);
    if (h->d_status) (void)hipFree(h->d_status);
    if (h->d_cst) (void)hipFree(h->d_cst);
    if (h->h_status) (void)hipHostFree(h->h_status);
    if (h->status_ev) (void)hipEventDestroy(h->status_ev);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return CTS_OK;
}

int cts_set_bunches(cts_handle *h, int32_t nb, const int64_t *counts, const double *macro_charge,
                    const double *n_e, const int32_t *gap)
{
    if (!h || nb < 1 || !counts || !macro_charge || !n_e) return fail(CTS_ERR_ARG, "cts_set_bunches: bad arguments");
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    free_bunch_state(h);
    // the per-bunch effects are set up again below for the new layout
    const cts_handle::KeepLrw keep_lrw = h->keep_lrw;
    const cts_handle::KeepBbi keep_bbi = h->keep_bbi;
    const cts_handle::KeepBl keep_bl = h->keep_bl;
    std::vector<cts_ion_config> ions;
    for (const IonIp &I : h->ips) ions.push_back(I.nspec > 0 ? I.cfg : cts_ion_config{});
    free_ion_state(h);
    h->nb = nb;
    h->counts.assign(counts, counts + nb);
    h->macro_charge.assign(macro_charge, macro_charge + nb);
    h->n_e.assign(n_e, n_e + nb);
    h->gap.assign(nb, 1);
    if (gap) h->gap.assign(gap, gap + nb);
    h->offsets.assign(nb + 1, 0);
    for (int b = 0; b < nb; b++) {
        if (counts[b] < 0) return fail(CTS_ERR_ARG, "cts_set_bunches: negative count");
        h->offsets[b + 1] = h->offsets[b] + counts[b];
    }
    h->ntot = h->offsets[nb];
    const int64_t n = h->ntot;

    int rc;
    if ((rc = dalloc(&h->d_x, n)) || (rc = dalloc(&h->d_px, n)) || (rc = dalloc(&h->d_y, n)) ||
        (rc = dalloc(&h->d_py, n)) || (rc = dalloc(&h->d_z, n)) || (rc = dalloc(&h->d_dp, n)) ||
        (rc = dalloc(&h->d_flag, n)) || (rc = dalloc(&h->d_bunch_off, nb + 1)) ||
        (rc = dalloc(&h->d_bunch_count, nb)) || (rc = dalloc(&h->d_shift, (size_t)nb * 6)) ||
        (rc = dalloc(&h->d_cur, nb)) || (rc = dalloc(&h->d_sigz, nb)) || (rc = dalloc(&h->d_zrange, (size_t)nb * 2)) ||
        (rc = dalloc(&h->d_ion_e, (size_t)nb * CTS_MAX_SPECIES * 5)) || (rc = dalloc(&h->d_ion_bview, (size_t)nb * 4)) ||
        (rc = dalloc(&h->d_chunk_cnt, nb)))
        return rc;
    if ((rc = layout_segments(h, h->wg_target, h->tile))) return rc;
    HIPCHK(hipMemcpy(h->d_bunch_off, h->offsets.data(), (nb + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_bunch_count, h->counts.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice));
    h->kick_pending = false;
    h->dp_pending = false;
    h->last_turn = -1;
    h->cur_valid = false;
    if (h->hist_cap > 0) {
        if ((rc = dalloc(&h->d_hist, (size_t)h->hist_cap * nb))) return rc;
    }
    if (h->nbins > 0) {
        const int nbins = h->nbins;
        h->nbins = 0;  // re-run allocation of wake state for the new layout
        std::vector<double> fr(h->omega.size());
        for (size_t i = 0; i < fr.size(); i++) fr[i] = h->omega[i] / (2 * kPi);
        std::vector<double> rs = h->rs, q = h->q;
        rc = cts_set_wake(h, nbins, h->wake_mode, h->grid_lo, h->grid_hi, (int32_t)rs.size(), rs.data(),
                          q.data(), fr.data());
        if (rc) return rc;
    }
    if (h->sur_mode & 2) {
        if ((rc = dalloc(&h->d_fb_count, 1)) || (rc = dalloc(&h->d_fb_list, (size_t)h->ntot))) return rc;
    }
    if (h->noise_mode == 1) {
        // host-supplied normals: one buffer of 5 per particle of the new layout
        if ((rc = dalloc(&h->d_noise, (size_t)h->ntot * 5))) return rc;
    }
    if (keep_lrw.on) {
        cts_lrwake_config c = keep_lrw.cfg;
        c.sectors = keep_lrw.sectors.data();
        c.res = keep_lrw.res.data();
        if ((rc = cts_set_lrwake(h, &c))) return rc;
    }
    if (keep_bbi.on) {
        cts_bbimp_config c = keep_bbi.cfg;
        const double **dst[16] = {&c.zz_re,    &c.zz_im,   &c.zqx_re,  &c.zqx_im,  &c.zqy_re,  &c.zqy_im,
                                  &c.zdx_re,   &c.zdx_im,  &c.zdy_re,  &c.zdy_im,  &c.wake_pos, &c.wake_z,
                                  &c.wake_dx,  &c.wake_dy, &c.wake_qx, &c.wake_qy};
        for (int i = 0; i < 16; i++) *dst[i] = keep_bbi.arr[i].empty() ? nullptr : keep_bbi.arr[i].data();
        if ((rc = cts_set_bbimp(h, &c))) return rc;
    }
    if (keep_bl.on && (rc = cts_set_beam_loading(h, &keep_bl.cfg))) return rc;
    for (int k = 0; k < (int)ions.size(); k++)
        if (ions[k].nspec > 0 && (rc = cts_set_ions_ip(h, k, &ions[k]))) return rc;
    return CTS_OK;
}

int cts_upload(cts_handle *h, const double *x, const double *px, const double *y, const double *py,
               const double *z, const double *dp, const int32_t *lost)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!x || !px || !y || !py || !z || !dp) return fail(CTS_ERR_ARG, "cts_upload: null array");
    const size_t bytes = (size_t)h->ntot * sizeof(double);
    HIPCHK(hipMemcpy(h->d_x, x, bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_px, px, bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_y, y, bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_py, py, bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_z, z, bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_dp, dp, bytes, hipMemcpyHostToDevice));
    std::vector<uint8_t> fl((size_t)h->ntot, 0);
    if (lost) {
        for (int64_t i = 0; i < h->ntot; i++) {
            if (lost[i] < 0 || lost[i] > 2) return fail(CTS_ERR_ARG, "cts_upload: loss code outside 0..2");
            fl[i] = (uint8_t)lost[i];
        }
    }
    HIPCHK(hipMemcpy(h->d_flag, fl.data(), fl.size(), hipMemcpyHostToDevice));
    if (h->lrw_T > 0) {
        // a new run starts from an empty centroid history and no pending
        // long-range kick
        HIPCHK(hipMemset(h->d_lrw_ring, 0, (size_t)h->lrw_T * h->nb * 3 * sizeof(double)));
        HIPCHK(hipMemset(h->d_lrw_kick, 0, (size_t)h->nb * 3 * sizeof(double)));
    }
    // the moment sums are taken about a per-bunch shift that follows the
    // centroids; a new run starts from zero, as a new handle does, so that
    // its rounding is the same
    HIPCHK(hipMemset(h->d_shift, 0, (size_t)h->nb * 6 * sizeof(double)));
    h->lr_pending = false;
    h->kick_pending = false;
    h->dp_pending = false;
    h->last_turn = -1;
    h->cur_valid = false;
    if ((rc = reset_status(h))) return rc;
    // moments of the uploaded state (the turn-0 MPBeamRMSCal)
    return cts_rms(h, nullptr);
}

int cts_download(cts_handle *h, double *x, double *px, double *y, double *py, double *z, double *dp,
                 int32_t *lost)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if ((rc = flush_kick(h))) return rc;
    const size_t bytes = (size_t)h->ntot * sizeof(double);
    HIPCHK(hipStreamSynchronize(h->stream));
    if (x) HIPCHK(hipMemcpy(x, h->d_x, bytes, hipMemcpyDeviceToHost));
    if (px) HIPCHK(hipMemcpy(px, h->d_px, bytes, hipMemcpyDeviceToHost));
    if (y) HIPCHK(hipMemcpy(y, h->d_y, bytes, hipMemcpyDeviceToHost));
    if (py) HIPCHK(hipMemcpy(py, h->d_py, bytes, hipMemcpyDeviceToHost));
    if (z) HIPCHK(hipMemcpy(z, h->d_z, bytes, hipMemcpyDeviceToHost));
    if (dp) HIPCHK(hipMemcpy(dp, h->d_dp, bytes, hipMemcpyDeviceToHost));
    if (lost) {
        std::vector<uint8_t> fl((size_t)h->ntot);
        HIPCHK(hipMemcpy(fl.data(), h->d_flag, fl.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < h->ntot; i++) lost[i] = fl[i] & 3u;
    }
    return CTS_OK;
}

int cts_set_wake(cts_handle *h, int32_t nbins, int32_t mode, double grid_lo, double grid_hi, int32_t nres,
                 const double *rs, const double *q, const double *fr_hz)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (nbins < 0 || nbins > 1024 || (mode != 0 && mode != 1) || nres < 0 || nres > CTS_MAX_RES)
        return fail(CTS_ERR_ARG, "cts_set_wake: nbins in [0,1024], mode 0/1, nres <= 8");
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    if (mode == 1 && nbins > 0 && !(grid_hi > grid_lo)) return fail(CTS_ERR_ARG, "cts_set_wake: empty grid");
    for (int r = 0; r < nres; r++)
        if (!(q[r] > 0.5)) return fail(CTS_ERR_ARG, "cts_set_wake: BBR needs Q > 1/2");
    if ((rc = flush_kick(h))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    void *old[] = {h->d_bhist, h->d_counts, h->d_kick, h->d_poten, h->d_edges_prev, h->d_edges_now,
                   h->d_table, h->d_scale};
    for (void *p : old)
        if (p) (void)hipFree(p);
    h->d_bhist = nullptr; h->d_counts = nullptr;
    h->d_kick = h->d_poten = h->d_edges_prev = h->d_edges_now = h->d_table = h->d_scale = nullptr;
    h->nbins = (nres > 0) ? nbins : 0;
    h->wake_mode = mode;
    h->grid_lo = grid_lo;
    h->grid_hi = grid_hi;
    h->nres = nres;
    h->rs.assign(rs, rs + nres);
    h->q.assign(q, q + nres);
    h->omega.resize(nres);
    for (int r = 0; r < nres; r++) h->omega[r] = fr_hz[r] * 2 * kPi;  // src/WakeFunction.cpp:275
    h->kick_pending = false;
    if (h->nbins == 0) return CTS_OK;
    const int nb = h->nb, n = h->nbins;
    h->hist_parity = 0;
    if ((rc = dalloc(&h->d_bhist, (size_t)2 * nb * n)) || (rc = dalloc(&h->d_counts, (size_t)nb * n)) ||
        (rc = dalloc(&h->d_kick, (size_t)nb * n)) || (rc = dalloc(&h->d_poten, (size_t)nb * n)) ||
        (rc = dalloc(&h->d_edges_prev, (size_t)nb * 2)) || (rc = dalloc(&h->d_edges_now, (size_t)nb * 2)) ||
        (rc = dalloc(&h->d_table, (size_t)(mode == 0 ? nb : 1) * (2 * n - 1))) || (rc = dalloc(&h->d_scale, nb)))
        return rc;
    // scale of srWakePoten: (-1)*e*q/E (src/MPBunch.cpp:1996-1998)
    std::vector<double> sc(nb);
    for (int b = 0; b < nb; b++) sc[b] = (-1) * kEcharge * h->macro_charge[b] / h->ring.energy;
    HIPCHK(hipMemcpy(h->d_scale, sc.data(), nb * sizeof(double), hipMemcpyHostToDevice));
    if (mode == 1) {
        std::vector<double> ed(2 * nb);
        const double dz = (grid_hi - grid_lo) / n;
        for (int b = 0; b < nb; b++) { ed[2 * b] = grid_lo; ed[2 * b + 1] = dz; }
        HIPCHK(hipMemcpy(h->d_edges_now, ed.data(), ed.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->d_edges_prev, ed.data(), ed.size() * sizeof(double), hipMemcpyHostToDevice));
        WakeTableParams t;
        std::memset(&t, 0, sizeof t);
        t.edges = h->d_edges_now;  // bunch 0's width is the grid's
        t.table = h->d_table;
        t.nbins = n;
        t.nres = nres;
        for (int r = 0; r < nres; r++) { t.rs[r] = h->rs[r]; t.q[r] = h->q[r]; t.omega[r] = h->omega[r]; }
        HIPCHK(launch_wake_table(t, 1, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return CTS_OK;
}

int cts_set_synrad(cts_handle *h, int32_t on, int32_t noise_mode, uint64_t seed)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (noise_mode < 0 || noise_mode > 2) return fail(CTS_ERR_ARG, "cts_set_synrad: noise_mode 0..2");
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    for (int i = 0; i < 3; i++)
        if (on && !(h->ring.synrad_tau[i] > 0)) return fail(CTS_ERR_ARG, "cts_set_synrad: damping times must be > 0");
    h->synrad_on = on ? 1 : 0;
    h->noise_mode = noise_mode;
    h->seed = seed;
    if (noise_mode == 1 && !h->d_noise) {
        if ((rc = dalloc(&h->d_noise, (size_t)h->ntot * 5))) return rc;
    }
    return CTS_OK;
}

int cts_push_noise(cts_handle *h, const double *noise)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!h->d_noise) return fail(CTS_ERR_STATE, "cts_push_noise: noise_mode 1 not set");
    HIPCHK(hipMemcpyAsync(h->d_noise, noise, (size_t)h->ntot * 5 * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CTS_OK;
}

// MPBeam::LRWakeBeamIntaction (src/MPBeam.cpp:3232-3351) after the turn's
// moments: the kick of every bunch, applied by the next pass (or a flush).
static int run_lrwake(cts_handle *h, int32_t turn)
{
    LrwParams L;
    std::memset(&L, 0, sizeof L);
    L.table = h->d_lrw_table;
    L.np = h->lrw.np;
    L.nturns = h->lrw_T;
    L.nb = h->nb;
    L.harmonic = h->ring.harmonic;
    L.dt = h->lrw.dt;
    L.trf = h->lrw.trf;
    L.rbeta = h->ring.rbeta;
    L.scale = kEcharge / h->ring.energy / std::pow(h->ring.rbeta, 2);
    L.harm = h->d_lrw_harm;
    L.ne = h->d_lrw_ne;
    L.cen = h->d_lrw_cen;
    L.ring = h->d_lrw_ring;
    L.kick = h->d_lrw_kick;
    L.cur = h->d_cur;
    L.hist = h->d_hist;
    L.hist_cap = h->hist_cap;
    L.fix_dp = h->nbins > 0 ? 0 : 1;  // with a short-range wake the dp moments close in the next pass
    L.turn = (uint32_t)turn;
    if (h->capturing) {
        L.turn_dev = h->d_turn;
        L.turn = (uint32_t)h->cap_off;
    }
    L.status = h->d_status;
    TimedScope ts(h, "lrwake");
    HIPCHK(launch_lrwake(L, h->stream));
    h->lr_pending = true;
    return CTS_OK;
}

// MPBeam::BBImpBeamInteraction (src/MPBeam.cpp:1321-1345) after the
// turn's pass: bin, induced voltage, kick; the delta moments of the turn's
// record are then completed by the next pass (as with a wake kick)
int run_bbi(cts_handle *h)
{
    const BbiParams p = bbi_params(h);
    {
        TimedScope ts(h, "bbi_bin");
        HIPCHK(launch_bbi_bin(p, h->nwg, h->stream));
    }
    {
        TimedScope ts(h, "bbi_wake");
        HIPCHK(launch_bbi_wake(p, h->stream));
    }
    TimedScope ts(h, "bbi_kick");
    HIPCHK(launch_bbi_kick(p, h->nwg, h->stream));
    h->kick_pending = true;
    return CTS_OK;
}

// MPBeam::Run body for one section (src/MPBeam.cpp:1059-1102), fused.
// The kernels of one turn (everything cts_turn enqueues after its status
// check); also what a graph capture records, twice.
static int turn_body(cts_handle *h, int32_t turn)
{
    int rc;
    uint32_t ops = OP_MAP | OP_RF | OP_DRIFT | OP_U0 | OP_LOSS | OP_MOMA | OP_WRITE;
    if (h->ring.skew_k != 0) ops |= OP_SKEW;
    if (h->synrad_on) ops |= OP_SYNRAD;
    // beam-ion phase at interaction point 0, before the map
    // (src/MPBeam.cpp:1061-1068)
    if (h->ips[0].nspec > 0 && (rc = ion_phase(h, turn, 0))) return rc;
    const bool wake = h->nbins > 0;
    if (h->kick_pending) ops |= OP_KICK | OP_MOMB;
    if (wake) ops |= (h->wake_mode == 1) ? OP_BIN : OP_ZSTAT;
    else if (!h->bbi_on) ops |= OP_MOMD;  // no deferred kick: dp moments close this turn
    if (h->lr_pending) ops |= OP_LRK;
    if (h->lrw_T > 0) ops |= OP_LRC;
    if (h->bl_on) {
        // beam-loaded cavities bin every bunch over its z range after the
        // section map (src/MPBeam.cpp:2637, 2716-2752): a read-only pass
        // runs the stages up to the map (and the skew quad) for that range
        // alone, then the turn's own pass repeats them (same operations on
        // the same inputs, same values) and applies the kick
        const uint32_t opsA = (ops & (OP_KICK | OP_LRK | OP_MAP | OP_SKEW)) | OP_ZSTAT;
        PassParams pa = base_pass(h, opsA, 0);
        if (!h->capturing) pa.turn = (uint32_t)turn;
        if ((rc = run_pass(h, pa, "pass_a"))) return rc;
        if ((rc = run_finalize(h, base_fin(h, OP_ZSTAT, -1, -1)))) return rc;
        if (h->bl.walk) {
            // the cavity state depends on the particles: a second read-only
            // pass counts every bunch's survivors per bin over that range,
            // then one wave per cavity walks the train
            const uint32_t opsB = (ops & (OP_KICK | OP_LRK | OP_MAP | OP_SKEW)) | OP_BLH;
            PassParams pb = base_pass(h, opsB, 0);
            if (!h->capturing) pb.turn = (uint32_t)turn;
            if ((rc = run_pass(h, pb, "pass_b"))) return rc;
            BlWalkParams w{};
            w.state = h->d_bl_state; w.hist = h->d_bl_hist; w.zrange = h->d_zrange; w.gap = h->d_bl_gap;
            w.ne = h->d_bl_ne; w.mq = h->d_bl_mq; w.kick = h->d_bl_kick; w.fb = h->d_bl_fb; w.info = h->d_bl_info;
            w.nb = h->nb; w.nbins = h->bl.nbins; w.nbl = h->bl.nbl; w.nres = h->bl.nres;
            for (int q = 0; q < h->bl.nbl; q++) {
                w.bl_index[q] = h->bl.bl_index[q];
                w.kfac[q] = h->bl.kfac[q];
            }
            w.nturns = h->bl.nturns;
            w.turn = (uint32_t)turn;
            if (h->capturing) {
                w.turn_dev = h->d_turn;
                w.turn = (uint32_t)h->cap_off;
            }
            w.trf = h->ring.t0 / h->ring.harmonic;
            w.rbeta = h->ring.rbeta;
            w.energy = h->ring.energy;
            w.beta2 = std::pow(h->ring.rbeta, 2);
            {
                TimedScope ts(h, "bl_walk");
                HIPCHK(launch_bl_walk(w, h->stream));
            }
            if (!h->capturing) h->bl_info_turn = std::max(h->bl_info_turn, turn);
        }
        ops |= OP_BL;
    }
    PassParams p = base_pass(h, ops, 0);
    if (!h->capturing) p.turn = (uint32_t)turn;
    if ((rc = run_pass(h, p))) return rc;
    // finalize: dp part of the previous record (MOMB), this turn's record
    FinParams f = base_fin(h, ops & (OP_MOMB | OP_MOMA | OP_MOMD | OP_LOSS | OP_ZSTAT | OP_LRC), slot_of(h, turn),
                           slot_of(h, h->last_turn));
    if (wake && h->wake_mode == 1) {
        // fixed grid: moments and wake in one launch
        TimedScope ts(h, "post");
        HIPCHK(launch_post(f, wake_params(h), h->stream));
        h->hist_parity = 1 - h->hist_parity;
        h->kick_pending = true;
    } else {
        if ((rc = run_finalize(h, f))) return rc;
        if (wake) {
            if ((rc = run_bin(h))) return rc;
            if ((rc = run_wake(h))) return rc;
            h->kick_pending = true;
        }
    }
    if (h->bbi_on && (rc = run_bbi(h))) return rc;
    if (h->lrw_T > 0 && (rc = run_lrwake(h, turn))) return rc;
    h->last_turn = turn;
    h->dp_pending = wake || h->bbi_on;
    h->cur_valid = true;
    return CTS_OK;
}

int cts_turn(cts_handle *h, int32_t turn)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (h->secs.size() != 1) return fail(CTS_ERR_STATE, "cts_turn fuses one section; use the seam calls");
    if (h->bl_on && !h->bl.walk && (turn < 0 || turn >= h->bl.nturns))
        return fail(CTS_ERR_ARG, "cts_turn: turn outside the beam-loading table (cts_set_beam_loading nturns)");
    // a bunch below 50% transmission ends the run, as the reference exits
    if ((rc = check_status(h, false))) return rc;
    if ((rc = turn_body(h, turn))) return rc;
    return enqueue_status_copy(h);
}

// Graph of turns (t, t+1) for the handle's present configuration, captured
// from turn_body with the turn read from d_turn.  The host state that the
// capture advances (buffer parities, pending flags) is put back: the graph
// only takes effect when it is launched.
static int capture_pair(cts_handle *h, hipGraphExec_t *out)
{
    const int parity = h->hist_parity, ion_buf = h->ips[0].buf, last_turn = h->last_turn;
    const bool kick_pending = h->kick_pending, dp_pending = h->dp_pending, cur_valid = h->cur_valid;
    const bool lr_pending = h->lr_pending;
    const bool timing = h->timing;
    const int64_t ion_len_ub = h->ips[0].len_ub;
    h->timing = false;
    h->capturing = true;
    hipGraph_t g = nullptr;
    int rc = CTS_OK;
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int off = 0; off < 2 && rc == CTS_OK; off++) {
            h->cap_off = off;
            rc = turn_body(h, last_turn + 1 + off);
        }
        if (rc == CTS_OK) {
            e = launch_advance_turn(h->d_turn, 2u, h->stream);
            if (e != hipSuccess) rc = fail(CTS_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        }
        const hipError_t e2 = hipStreamEndCapture(h->stream, &g);
        if (rc == CTS_OK && e2 != hipSuccess)
            rc = fail(CTS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
    } else {
        rc = fail(CTS_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    }
    h->capturing = false;
    h->cap_off = 0;
    h->timing = timing;
    h->hist_parity = parity;
    h->ips[0].buf = ion_buf;
    h->last_turn = last_turn;
    h->kick_pending = kick_pending;
    h->dp_pending = dp_pending;
    h->cur_valid = cur_valid;
    h->lr_pending = lr_pending;
    h->ips[0].len_ub = ion_len_ub;
    if (rc == CTS_OK) {
        e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = fail(CTS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    if (g) (void)hipGraphDestroy(g);
    return rc;
}

// The surrogate parameters a graph captured (pointer and normalisation are
// kernel arguments).
static void sur_state(cts_handle *h, const float **params, SurNorm *norm)
{
    *params = nullptr;
    *norm = SurNorm{0.f, 0.f, 0.f, 0.f};
    if (h->sur && h->sur_mode) sur_export(h->sur, params, norm);
}

int cts_run_turns(cts_handle *h, int32_t first, int32_t nturns, int32_t use_graph)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (h->secs.size() != 1) return fail(CTS_ERR_STATE, "cts_run_turns fuses one section; use the seam calls");
    if (nturns < 0 || first < 0) return fail(CTS_ERR_ARG, "cts_run_turns: first, nturns >= 0");
    if (h->bl_on && !h->bl.walk && first + nturns > h->bl.nturns)
        return fail(CTS_ERR_ARG, "cts_run_turns: turns beyond the beam-loading table (cts_set_beam_loading nturns)");
    int t = first;
    const int end = first + nturns;
    while (t < end) {
        // a replay starts from the steady state of the turn loop: the last
        // turn was t - 1 and, with a wake, its kick is pending; host noise
        // (pushed per turn) is not replayable
        const bool steady = use_graph && end - t >= 2 && h->cur_valid && h->last_turn == t - 1 &&
                            h->kick_pending == (h->nbins > 0 || h->bbi_on) && h->lr_pending == (h->lrw_T > 0) &&
                            h->noise_mode != 1;
        if (!steady) {
            if ((rc = cts_turn(h, t))) return rc;
            t++;
            continue;
        }
        if ((rc = check_status(h, false))) return rc;
        const float *sp;
        SurNorm sn;
        sur_state(h, &sp, &sn);
        cts_handle::GraphSlot *slot = nullptr;
        for (auto &g : h->graphs)
            if (g.exec && g.parity == h->hist_parity && g.ion_buf == h->ips[0].buf) slot = &g;
        if (slot && (slot->sur_params != sp || std::memcmp(&slot->sur_norm, &sn, sizeof sn) != 0)) {
            graph_reset(h);
            slot = nullptr;
        }
        if (!slot) {
            // a graph captured one turn later in the buffer cycle serves
            // after one eager turn (cheaper than a capture)
            const int p1 = h->nbins > 0 ? 1 - h->hist_parity : h->hist_parity;
            const int b1 = ion_buf_after_turn(h);
            bool later = false;
            for (auto &g : h->graphs)
                if (g.exec && g.parity == p1 && g.ion_buf == b1 && g.sur_params == sp &&
                    std::memcmp(&g.sur_norm, &sn, sizeof sn) == 0)
                    later = true;
            if (later && end - t >= 3) {
                if ((rc = cts_turn(h, t))) return rc;
                t++;
                continue;
            }
            for (auto &g : h->graphs)
                if (!g.exec) { slot = &g; break; }
            if (!slot) {
                graph_reset(h);
                slot = &h->graphs[0];
            }
            if ((rc = capture_pair(h, &slot->exec))) return rc;
            slot->parity = h->hist_parity;
            slot->ion_buf = h->ips[0].buf;
            slot->sur_params = sp;
            slot->sur_norm = sn;
        }
        if (h->turn_dev_next != t) HIPCHK(launch_set_turn(h->d_turn, (uint32_t)t, h->stream));
        HIPCHK(hipGraphLaunch(slot->exec, h->stream));
        h->turn_dev_next = t + 2;
        // the host state after two turns: parities back where they were
        h->last_turn = t + 1;
        h->dp_pending = h->nbins > 0 || h->bbi_on;
        h->lr_pending = h->lrw_T > 0;
        h->cur_valid = true;
        if (h->ips[0].nspec > 0) h->ips[0].len_ub = h->ips[0].cap;
        if (h->bl_on && h->bl.walk) h->bl_info_turn = std::max(h->bl_info_turn, t + 1);
        if ((rc = enqueue_status_copy(h))) return rc;
        t += 2;
    }
    return CTS_OK;
}

int cts_get_layout(cts_handle *h, int32_t *nwg, int32_t *tile)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (nwg) *nwg = h->nwg;
    if (tile) *tile = h->tile;
    return CTS_OK;
}

int cts_set_layout(cts_handle *h, int32_t nwg, int32_t tile)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (nwg < 1 || tile < 0) return fail(CTS_ERR_ARG, "cts_set_layout: nwg >= 1 and tile >= 0");
    if ((rc = flush_kick(h))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    h->wg_target = nwg;
    return layout_segments(h, nwg, tile);
}

// Layout of the streaming pass (workgroup count and tile), picked by
// timing: contiguous chunks for multiples of the CU count, round-robin
// tiles over 2..4 workgroups per CU, and the current setting; each is timed
// on a copy of the particles with the turn's own kernel variant.
int cts_autotune(cts_handle *h, int32_t *nwg_out)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if ((rc = flush_kick(h))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, h->dev));
    const int cus = prop.multiProcessorCount;
    // the turn's pass, as cts_turn builds it (with a pending kick)
    uint32_t ops = OP_MAP | OP_RF | OP_DRIFT | OP_U0 | OP_LOSS | OP_MOMA | OP_WRITE;
    if (h->ring.skew_k != 0) ops |= OP_SKEW;
    if (h->synrad_on) ops |= OP_SYNRAD;
    const bool wake = h->nbins > 0;
    if (wake) ops |= OP_KICK | OP_MOMB | ((h->wake_mode == 1) ? OP_BIN : OP_ZSTAT);
    else ops |= OP_MOMD;
    // scratch copies of the particles and of the histogram
    const int64_t n = h->ntot;
    double *cx[6] = {};
    uint8_t *cf = nullptr;
    uint32_t *chist = nullptr;
    auto release = [&]() {
        for (double *p : cx)
            if (p) (void)hipFree(p);
        if (cf) (void)hipFree(cf);
        if (chist) (void)hipFree(chist);
    };
    double *src[6] = {h->d_x, h->d_px, h->d_y, h->d_py, h->d_z, h->d_dp};
    for (int c = 0; c < 6 && !rc; c++) {
        rc = dalloc(&cx[c], (size_t)n);
        if (!rc && hipMemcpy(cx[c], src[c], (size_t)n * sizeof(double), hipMemcpyDeviceToDevice) != hipSuccess)
            rc = fail(CTS_ERR_HIP, "autotune copy");
    }
    if (!rc) rc = dalloc(&cf, (size_t)n);
    if (!rc && hipMemcpy(cf, h->d_flag, (size_t)n, hipMemcpyDeviceToDevice) != hipSuccess)
        rc = fail(CTS_ERR_HIP, "autotune copy");
    if (!rc && wake) rc = dalloc(&chist, (size_t)h->nb * h->nbins);
    if (rc) {
        release();
        return rc;
    }
    // (workgroups, tile): contiguous chunks for 1..8 workgroups per CU, then
    // round-robin tiles of 2048..16384 particles over 2..4 per CU
    std::vector<std::pair<int, int>> cand;
    for (int m = 1; m <= 8; m++) cand.push_back({cus * m, 0});
    for (int m = 2; m <= 4; m++)
        for (int t = 2048; t <= 16384; t *= 2) cand.push_back({cus * m, t});
    cand.push_back({h->wg_target, h->tile});
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    std::pair<int, int> best = {h->wg_target, h->tile};
    float best_ms = 1e30f;
    for (const auto &c : cand) {
        if ((rc = layout_segments(h, c.first, c.second))) break;
        PassParams p = base_pass(h, ops, 0);
        p.x = cx[0]; p.px = cx[1]; p.y = cx[2]; p.py = cx[3]; p.z = cx[4]; p.dp = cx[5];
        p.flag = cf;
        p.bunch_hist = chist;
        hipError_t e = launch_pass(p, h->nwg, h->stream);
        if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
        for (int r = 0; r < 5 && e == hipSuccess; r++) e = launch_pass(p, h->nwg, h->stream);
        if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            rc = fail(CTS_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
            break;
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    release();
    if (rc) return rc;
    h->wg_target = best.first;
    if ((rc = layout_segments(h, best.first, best.second))) return rc;
    if (nwg_out) *nwg_out = h->nwg;
    return CTS_OK;
}

int cts_flush(cts_handle *h)
{
    int rc = check_ready(h);
    if (rc) return rc;
    return flush_kick(h);
}

static int seam(cts_handle *h, uint32_t ops, int k, uint32_t turn)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (k < 0 || k >= (int)h->secs.size()) return fail(CTS_ERR_ARG, "section index out of range");
    if ((rc = flush_kick(h))) return rc;
    PassParams p = base_pass(h, ops | OP_WRITE, k);
    p.turn = turn;
    h->cur_valid = false;
    return run_pass(h, p);
}

int cts_section_map(cts_handle *h, int32_t k) { return seam(h, OP_MAP, k, 0); }
int cts_skew_quad(cts_handle *h) { return seam(h, OP_SKEW, 0, 0); }
int cts_rf_kick(cts_handle *h) { return seam(h, OP_RF, 0, 0); }
int cts_drift(cts_handle *h) { return seam(h, OP_DRIFT, 0, 0); }
int cts_energy_loss(cts_handle *h) { return seam(h, OP_U0, 0, 0); }
int cts_synrad(cts_handle *h, int32_t turn) { return seam(h, OP_SYNRAD, 0, (uint32_t)turn); }

int cts_sr_wake(cts_handle *h, const double *sigma_z_stale)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (h->nbins == 0) return CTS_OK;
    if ((rc = flush_kick(h))) return rc;
    if (h->wake_mode == 0) {
        // edges from the survivors' z range (GetZMinMax) and the rms bunch
        // length of the last MPBeamRMSCal (src/MPBunch.cpp:1900-1908)
        PassParams p = base_pass(h, OP_ZSTAT, 0);
        if ((rc = run_pass(h, p, "zstat"))) return rc;
        FinParams f = base_fin(h, OP_ZSTAT, -1, -1);
        if ((rc = run_finalize(h, f))) return rc;
        if (sigma_z_stale) {
            std::vector<double> zr(2 * h->nb), ed(2 * h->nb);
            HIPCHK(hipMemcpyAsync(zr.data(), h->d_zrange, zr.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            for (int b = 0; b < h->nb; b++) {
                const double zMin = zr[2 * b] - 1.e-6, zMax = zr[2 * b + 1] + 1.e-6;
                const double lo = zMin - 2 * sigma_z_stale[b], hi = zMax + 2 * sigma_z_stale[b];
                ed[2 * b] = lo;
                ed[2 * b + 1] = (hi - lo) / h->nbins;
            }
            HIPCHK(hipMemcpy(h->d_edges_now, ed.data(), ed.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        if ((rc = run_bin(h))) return rc;
    } else {
        PassParams p = base_pass(h, OP_BIN, 0);
        if ((rc = run_pass(h, p, "bin"))) return rc;
    }
    if ((rc = run_wake(h))) return rc;
    // immediate kick (the seam applies it within the call)
    h->kick_pending = true;
    PassParams p = base_pass(h, OP_KICK | OP_WRITE, 0);
    if ((rc = run_pass(h, p, "kick"))) return rc;
    h->kick_pending = false;
    return CTS_OK;
}

int cts_mark_lost(cts_handle *h, double *transmission)
{
    int rc = seam(h, OP_LOSS, 0, 0);
    if (rc) return rc;
    FinParams f = base_fin(h, OP_LOSS, -1, -1);
    if ((rc = run_finalize(h, f))) return rc;
    if ((rc = enqueue_status_copy(h))) return rc;
    if (transmission) {
        std::vector<cts_moments> m(h->nb);
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(m.data(), h->d_cur, m.size() * sizeof(cts_moments), hipMemcpyDeviceToHost));
        for (int b = 0; b < h->nb; b++) transmission[b] = m[b].transmission;
    }
    return check_status(h, true);
}

int cts_set_lrwake(cts_handle *h, const cts_lrwake_config *cfg)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!cfg || cfg->nturns < 0) return fail(CTS_ERR_ARG, "cts_set_lrwake: nturns >= 0");
    if ((rc = flush_kick(h))) return rc;  // a pending kick belongs to the old settings
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    free_lrw_state(h);
    h->keep_lrw.on = false;
    if (cfg->nturns == 0) return CTS_OK;
    if (h->secs.empty()) return fail(CTS_ERR_STATE, "cts_set_lrwake: no sections");
    if (cfg->n_sectors < 0 || cfg->n_res < 0 || (cfg->n_sectors > 0 && !cfg->sectors) ||
        (cfg->n_res > 0 && !cfg->res) || cfg->n_sectors + cfg->n_res == 0)
        return fail(CTS_ERR_ARG, "cts_set_lrwake: pipe sectors and/or resonators required");
    if (!(cfg->beta_x > 0) || !(cfg->beta_y > 0)) return fail(CTS_ERR_ARG, "cts_set_lrwake: beta functions > 0");
    for (int i = 0; i < cfg->n_sectors; i++) {
        const double *s = cfg->sectors + 7 * i;
        if (!(s[0] > 0) || !(s[1] > 0) || !(s[6] > 0))
            return fail(CTS_ERR_ARG, "cts_set_lrwake: sector radii and conductivity must be > 0");
    }
    for (int r = 0; r < cfg->n_res; r++) {
        const double *q = cfg->res + 9 * r;
        if (!(q[1] > 0.5) || !(q[4] > 0.5) || !(q[7] > 0.5))
            return fail(CTS_ERR_ARG, "cts_set_lrwake: resonator Q must exceed 1/2");
    }
    std::string err;
    LrwTable tab;
    if ((rc = lrwake_table(*cfg, h->ring.t0, h->ring.harmonic, tab, err))) return fail(rc, err);
    const int nb = h->nb;
    if ((rc = dalloc(&h->d_lrw_table, (size_t)3 * tab.np)) ||
        (rc = dalloc(&h->d_lrw_ring, (size_t)cfg->nturns * nb * 3)) || (rc = dalloc(&h->d_lrw_cen, (size_t)nb * 3)) ||
        (rc = dalloc(&h->d_lrw_kick, (size_t)nb * 3)) || (rc = dalloc(&h->d_lrw_ne, (size_t)nb)) ||
        (rc = dalloc(&h->d_lrw_harm, (size_t)nb))) {
        free_lrw_state(h);
        return rc;
    }
    // bunch harmonic numbers from the gaps (only differences enter)
    std::vector<int32_t> harm(nb, 0);
    for (int b = 1; b < nb; b++) harm[b] = harm[b - 1] + h->gap[b - 1];
    HIPCHK(hipMemcpy(h->d_lrw_table, tab.w.data(), tab.w.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_lrw_ne, h->n_e.data(), nb * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_lrw_harm, harm.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice));
    h->lrw = std::move(tab);
    h->lrw_T = cfg->nturns;
    h->lr_pending = false;
    cts_handle::KeepLrw k;
    k.on = true;
    k.cfg = *cfg;
    k.sectors.assign(cfg->sectors, cfg->sectors + (cfg->n_sectors > 0 ? 7 * (size_t)cfg->n_sectors : 0));
    k.res.assign(cfg->res, cfg->res + (cfg->n_res > 0 ? 9 * (size_t)cfg->n_res : 0));
    k.cfg.sectors = k.cfg.res = nullptr;
    h->keep_lrw = std::move(k);
    return CTS_OK;
}

int cts_set_bbimp(cts_handle *h, const cts_bbimp_config *cfg)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!cfg) return fail(CTS_ERR_ARG, "cts_set_bbimp: null config");
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    free_bbi_state(h);
    h->keep_bbi.on = false;
    if (cfg->flags[0] == 0 && cfg->flags[1] == 0 && cfg->flags[2] == 0 && cfg->flags[3] == 0 && cfg->flags[4] == 0)
        return CTS_OK;  // doTracking = 0 (src/MPBunch.cpp:2152-2157)
    if (h->secs.size() != 1) return fail(CTS_ERR_STATE, "cts_set_bbimp: the fused turn (one section) only");
    std::string err;
    if ((rc = bbimp_table(*cfg, h->ring, h->bbi, err))) {
        h->bbi = BbiTable{};
        return fail(rc, err);
    }
    const int n = h->bbi.n, nb = h->nb;
    const bool dip = h->bbi.planes[1] || h->bbi.planes[2];
    const size_t ng = h->bbi.g.size();
    if ((rc = dalloc(&h->d_bbi_hist, (size_t)nb * n)) || (rc = dalloc(&h->d_bbi_prof, 3 * (size_t)nb * n)) ||
        (rc = dalloc(&h->d_bbi_wake, 5 * (size_t)nb * n)) || (rc = dalloc(&h->d_bbi_g, ng)) ||
        (rc = dalloc(&h->d_bbi_scale, (size_t)nb)) || (dip && (rc = dalloc(&h->d_bbi_dsum, 4 * (size_t)nb * n)))) {
        free_bbi_state(h);
        return rc;
    }
    // charge per macroparticle: a current [C/s] per bin in the frequency
    // domain (src/MPBunch.cpp:2166-2171), a charge [C] in the time domain
    // (src/MPBunch.cpp:2053)
    std::vector<double> sc(nb);
    for (int b = 0; b < nb; b++)
        sc[b] = h->bbi.td ? h->macro_charge[b] * kEcharge : h->macro_charge[b] * kEcharge / h->bbi.dz * kClight;
    HIPCHK(hipMemcpy(h->d_bbi_g, h->bbi.g.data(), ng * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_bbi_scale, sc.data(), nb * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(h->d_bbi_hist, 0, (size_t)nb * n * sizeof(uint32_t)));
    if (dip) HIPCHK(hipMemset(h->d_bbi_dsum, 0, 4 * (size_t)nb * n * sizeof(unsigned long long)));
    HIPCHK(hipMemset(h->d_bbi_wake, 0, 5 * (size_t)nb * n * sizeof(double)));
    HIPCHK(hipMemset(h->d_bbi_prof, 0, 3 * (size_t)nb * n * sizeof(double)));
    h->bbi_on = true;
    cts_handle::KeepBbi k;
    k.on = true;
    k.cfg = *cfg;
    const double *src[16] = {cfg->zz_re,   cfg->zz_im,   cfg->zqx_re,  cfg->zqx_im,   cfg->zqy_re, cfg->zqy_im,
                             cfg->zdx_re,  cfg->zdx_im,  cfg->zdy_re,  cfg->zdy_im,   cfg->wake_pos, cfg->wake_z,
                             cfg->wake_dx, cfg->wake_dy, cfg->wake_qx, cfg->wake_qy};
    // only the tables of the chosen method are read (and kept)
    for (int i = 0; i < 16; i++)
        if (src[i] && (i < 10) == (h->bbi.td == 0)) k.arr[i].assign(src[i], src[i] + (i < 10 ? cfg->nimp : cfg->nwake));
    k.cfg.zz_re = k.cfg.zz_im = k.cfg.zqx_re = k.cfg.zqx_im = k.cfg.zqy_re = k.cfg.zqy_im = nullptr;
    k.cfg.zdx_re = k.cfg.zdx_im = k.cfg.zdy_re = k.cfg.zdy_im = nullptr;
    k.cfg.wake_pos = k.cfg.wake_z = k.cfg.wake_dx = k.cfg.wake_dy = k.cfg.wake_qx = k.cfg.wake_qy = nullptr;
    h->keep_bbi = std::move(k);
    return CTS_OK;
}

int cts_get_bbimp_wake(cts_handle *h, int32_t *n, double *wake, double *prof)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!h->bbi_on) return fail(CTS_ERR_STATE, "cts_get_bbimp_wake: no broadband impedance");
    if (n) *n = h->bbi.n;
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t sz = (size_t)h->nb * h->bbi.n * sizeof(double);
    if (wake) HIPCHK(hipMemcpy(wake, h->d_bbi_wake, 5 * sz, hipMemcpyDeviceToHost));
    if (prof) HIPCHK(hipMemcpy(prof, h->d_bbi_prof, 3 * sz, hipMemcpyDeviceToHost));
    return CTS_OK;
}

int cts_set_beam_loading(cts_handle *h, const cts_beam_loading_config *cfg)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!cfg || cfg->nres < 0 || cfg->nres > CTS_MAX_RF) return fail(CTS_ERR_ARG, "cts_set_beam_loading: nres 0..4");
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    free_bl_state(h);
    h->keep_bl.on = false;
    if (cfg->nres == 0) return CTS_OK;
    if (h->secs.size() != 1) return fail(CTS_ERR_STATE, "cts_set_beam_loading: the fused turn (one section) only");
    if (cfg->nbins < 1 || cfg->nturns < 1) return fail(CTS_ERR_ARG, "cts_set_beam_loading: nbins, nturns >= 1");
    {
        // the per-bin kicks of a bunch sit in LDS next to the wake's bins
        int nbl = 0;
        for (int i = 0; i < cfg->nres && i < CTS_MAX_RF; i++) nbl += cfg->rf_mode[i] == 1;
        if ((int64_t)nbl * (cfg->nbins + 1) > kBlMaxLds)
            return fail(CTS_ERR_ARG, "cts_set_beam_loading: beam-loaded cavities x (nbins + 1) above " +
                                         std::to_string(kBlMaxLds));
    }
    for (int i = 0; i < cfg->nres; i++) {
        if (cfg->rf_mode[i] != 0 && cfg->rf_mode[i] != 1) return fail(CTS_ERR_ARG, "cts_set_beam_loading: rf_mode 0 or 1");
        if (!(cfg->q0[i] > 0) || cfg->res_harm[i] < 1 || !(cfg->coupling[i] >= 0))
            return fail(CTS_ERR_ARG, "cts_set_beam_loading: Q0 > 0, harmonic >= 1, coupling >= 0");
        if (cfg->rf_mode[i] == 1 && !(cfg->rs[i] > 0)) return fail(CTS_ERR_ARG, "cts_set_beam_loading: Rs > 0");
    }
    // the bin-by-bin variant or the direct feedback: the state is walked on
    // the device every turn
    bool walk = false;
    int nbl_all = 0;
    for (int i = 0; i < cfg->nres; i++)
        if (cfg->rf_mode[i] == 1) {
            walk = walk || cfg->excite_instability[i] != 0 || cfg->dir_fb[i] != 0;
            nbl_all++;
        }
    if (walk && (int64_t)nbl_all * (2 * cfg->nbins + 1) > kBlMaxLds)
        return fail(CTS_ERR_ARG, "cts_set_beam_loading: beam-loaded cavities x (2 nbins + 1) above " +
                                     std::to_string(kBlMaxLds) + " with the bin-by-bin variant or the feedback");
    if (cfg->nres == 1 && !(std::fabs(h->ring.u0 / cfg->vol[0]) <= 1))
        return fail(CTS_ERR_ARG, "cts_set_beam_loading: one cavity needs |U0 / V| <= 1");
    std::string err;
    BlTable t;
    if (walk) {
        const double bytes = (double)cfg->nturns * h->nb * cfg->nres * 8 * 8;
        if (bytes > 4e9) return fail(CTS_ERR_ARG, "cts_set_beam_loading: cavity record above 4 GB (fewer turns)");
        if ((rc = beam_loading_walk_state(*cfg, h->ring, h->n_e, h->gap, t, err))) return fail(rc, err);
        const int nb = h->nb, nbl = t.nbl, nbins = t.nbins;
        std::vector<double> mq(nb);
        for (int b = 0; b < nb; b++) mq[b] = h->macro_charge[b];
        if ((rc = dalloc(&h->d_bl_state, t.state.size())) || (rc = dalloc(&h->d_bl_hist, (size_t)nb * nbins)) ||
            (rc = dalloc(&h->d_bl_kick, (size_t)nb * nbl * nbins * 2)) || (rc = dalloc(&h->d_bl_fb, (size_t)nb * nbl)) ||
            (rc = dalloc(&h->d_bl_info, (size_t)t.nturns * nb * t.nres * 8)) || (rc = dalloc(&h->d_bl_ne, (size_t)nb)) ||
            (rc = dalloc(&h->d_bl_mq, (size_t)nb)) || (rc = dalloc(&h->d_bl_gap, (size_t)nb))) {
            free_bl_state(h);
            return rc;
        }
        HIPCHK(hipMemcpy(h->d_bl_state, t.state.data(), t.state.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->d_bl_ne, h->n_e.data(), nb * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->d_bl_mq, mq.data(), nb * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->d_bl_gap, h->gap.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(h->d_bl_hist, 0, (size_t)nb * nbins * sizeof(uint32_t)));
        HIPCHK(hipMemset(h->d_bl_fb, 0, (size_t)nb * nbl * sizeof(double)));
        HIPCHK(hipMemset(h->d_bl_info, 0, (size_t)t.nturns * nb * t.nres * 8 * sizeof(double)));
        h->bl_fbmask = 0;
        for (int q = 0; q < nbl; q++)
            if (t.state[(size_t)q * kBlS + BS_DFB] != 0.0) h->bl_fbmask |= 1u << q;
        h->bl_info_turn = -1;
    } else {
        const double bytes = (double)cfg->nturns * h->nb * CTS_MAX_RF * 3 * 8;
        if (bytes > 4e9) return fail(CTS_ERR_ARG, "cts_set_beam_loading: table above 4 GB (fewer turns per call)");
        if ((rc = beam_loading_tables(*cfg, h->ring, h->n_e, h->gap, t, err))) return fail(rc, err);
        if ((rc = dalloc(&h->d_bl_tab, t.dev.size()))) return rc;
        HIPCHK(hipMemcpy(h->d_bl_tab, t.dev.data(), t.dev.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    h->bl = std::move(t);
    h->bl_on = true;
    h->keep_bl.on = true;
    h->keep_bl.cfg = *cfg;
    return CTS_OK;
}

int cts_get_cavity_info(cts_handle *h, int32_t turn, double *info)
{
    if (!h || !info) return fail(CTS_ERR_ARG, "null argument");
    if (!h->bl_on) return fail(CTS_ERR_STATE, "cts_get_cavity_info: no beam loading");
    const size_t n = (size_t)h->nb * h->bl.nres * 8;
    if (h->bl.walk) {
        // the walk records turns 0 .. nturns-1 as it runs them
        if (turn < 0 || turn >= h->bl.nturns || turn > h->bl_info_turn)
            return fail(CTS_ERR_ARG, "cts_get_cavity_info: turn not walked yet or past the record (nturns)");
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(info, h->d_bl_info + (size_t)turn * n, n * sizeof(double), hipMemcpyDeviceToHost));
        return CTS_OK;
    }
    if (turn < 0 || turn >= h->bl.nturns) return fail(CTS_ERR_ARG, "cts_get_cavity_info: turn outside the table");
    std::memcpy(info, h->bl.info.data() + (size_t)turn * n, n * sizeof(double));
    return CTS_OK;
}

int cts_get_lrwake_table(cts_handle *h, int32_t *np, double *time, double *w, double *rw, double *bbr)
{
    if (!h || !np) return fail(CTS_ERR_ARG, "null argument");
    if (h->lrw_T == 0) return fail(CTS_ERR_STATE, "cts_get_lrwake_table: no long-range wake");
    const int n = h->lrw.np;
    if (*np < n) {
        *np = n;
        return fail(CTS_ERR_ARG, "cts_get_lrwake_table: capacity below np (np returned)");
    }
    *np = n;
    if (time) std::memcpy(time, h->lrw.time.data(), (size_t)n * sizeof(double));
    if (w) std::memcpy(w, h->lrw.w.data(), (size_t)3 * n * sizeof(double));
    if (rw) std::memcpy(rw, h->lrw.rw.data(), (size_t)3 * n * sizeof(double));
    if (bbr) std::memcpy(bbr, h->lrw.bbr.data(), (size_t)3 * n * sizeof(double));
    return CTS_OK;
}

int cts_get_lrwake_kick(cts_handle *h, double *kick)
{
    if (!h || !kick) return fail(CTS_ERR_ARG, "null argument");
    if (h->lrw_T == 0) return fail(CTS_ERR_STATE, "cts_get_lrwake_kick: no long-range wake");
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(kick, h->d_lrw_kick, (size_t)h->nb * 3 * sizeof(double), hipMemcpyDeviceToHost));
    return CTS_OK;
}

int cts_set_ions(cts_handle *h, const cts_ion_config *cfg) { return cts_set_ions_ip(h, 0, cfg); }

int cts_set_ions_ip(cts_handle *h, int32_t k, const cts_ion_config *cfg)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!cfg || cfg->nspec < 0 || cfg->nspec > CTS_MAX_SPECIES) return fail(CTS_ERR_ARG, "cts_set_ions: nspec 0..4");
    if (k < 0 || k >= (int)h->ips.size()) return fail(CTS_ERR_ARG, "cts_set_ions_ip: no such section");
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    IonIp &Ip = h->ips[k];
    free_ion_ip(Ip);
    if (cfg->nspec == 0) return CTS_OK;
    if (cfg->n_new < 1 || cfg->max_ions < 1 || !(cfg->temperature > 0) || !(cfg->length >= 0) ||
        !(cfg->ap_x > 0) || !(cfg->ap_y > 0) || !(cfg->loss_boundary > 0))
        return fail(CTS_ERR_ARG, "cts_set_ions: n_new, max_ions, temperature, apertures and loss boundary must be > 0");
    for (int p = 0; p < cfg->nspec; p++)
        if (!(cfg->mass_number[p] > 0)) return fail(CTS_ERR_ARG, "cts_set_ions: mass number must be > 0");
    // one passage appends n_new per species; a phase appends nb * n_new
    // before its last compaction
    const int64_t cap = cfg->max_ions + (int64_t)h->nb * cfg->n_new;
    if (cap > (int64_t)1 << 31) return fail(CTS_ERR_ARG, "cts_set_ions: cloud capacity above 2^31 slots");
    // waves of a segment: the compacted cloud, then kIonSegment passages' new ions
    const int64_t wstride = (cap + 63) / 64 + (int64_t)kIonSegment * ((cfg->n_new + 63) / 64) + 1;
    Ip.cfg = *cfg;
    Ip.cap = cap;
    Ip.wstride = wstride;
    auto fail_free = [&](int code) {
        free_ion_ip(Ip);
        return code;
    };
    for (int b = 0; b < 2; b++)
        for (int p = 0; p < cfg->nspec; p++) {
            for (int c = 0; c < 4; c++)
                if ((rc = dalloc(&Ip.d_ion[b][p][c], (size_t)cap))) return fail_free(rc);
            if ((rc = dalloc(&Ip.d_alive[b][p], (size_t)cap))) return fail_free(rc);
        }
    if ((rc = dalloc(&Ip.d_len, 3 * CTS_MAX_SPECIES)) || (rc = dalloc(&Ip.d_nalive, CTS_MAX_SPECIES)) ||
        (rc = dalloc(&Ip.d_mean, 2 * CTS_MAX_SPECIES)) || (rc = dalloc(&Ip.d_smax, 2)) ||
        (rc = dalloc(&Ip.d_q, (size_t)h->nb * cfg->nspec)) ||
        (rc = dalloc(&Ip.d_cloud, (size_t)h->nb * CTS_MAX_SPECIES * 6)) ||
        (rc = dalloc(&Ip.d_coeff, (size_t)h->nb * CTS_MAX_SPECIES)) || (rc = dalloc(&Ip.d_dt, (size_t)h->nb)) ||
        (rc = dalloc(&Ip.d_acc, (size_t)(kIonReplicas + 1) * cfg->nspec * h->nb * kIonAcc * 3)) ||
        (rc = dalloc(&Ip.d_wcnt, (size_t)(cfg->nspec * wstride))) || (rc = dalloc(&Ip.d_wofs, (size_t)wstride)))
        return fail_free(rc);
    // per bunch: ion kick coefficient 2 N_e r_e (m_e / m_u) / A c
    // (src/MPBunch.cpp:1117) and drift time C / h * gap / c
    // (src/LatticeInterActionPoint.cpp:1169-1170)
    std::vector<double> coeff((size_t)h->nb * CTS_MAX_SPECIES, 0.0), dtv((size_t)h->nb);
    for (int b = 0; b < h->nb; b++) {
        for (int p = 0; p < cfg->nspec; p++)
            coeff[(size_t)b * CTS_MAX_SPECIES + p] =
                2.0 * h->n_e[b] * kRe * kElectronMassEV / kIonMassEV / cfg->mass_number[p] * kClight;
        dtv[b] = h->ring.circ / h->ring.harmonic * h->gap[b] / kClight;
    }
    HIPCHK(hipMemcpy(Ip.d_coeff, coeff.data(), coeff.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(Ip.d_dt, dtv.data(), dtv.size() * sizeof(double), hipMemcpyHostToDevice));
    // macro-ion charge per bunch passage: line density sigma P / (k_B T) N_e
    // over the interaction length, shared by n_new macro-ions
    // (src/LatticeInterActionPoint.cpp:974-981, 998)
    std::vector<double> q((size_t)h->nb * cfg->nspec);
    for (int b = 0; b < h->nb; b++)
        for (int p = 0; p < cfg->nspec; p++) {
            const double line = cfg->cross_section[p] * cfg->pressure[p] / cfg->temperature / kBoltzmann * h->n_e[b];
            q[(size_t)b * cfg->nspec + p] = line * cfg->length / cfg->n_new;
        }
    HIPCHK(hipMemcpy(Ip.d_q, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
    Ip.q = std::move(q);
    Ip.nspec = cfg->nspec;
    return CTS_OK;
}

int cts_set_field_model(cts_handle *h, const cts_surrogate *s, int32_t mode)
{
    if (!h) return fail(CTS_ERR_ARG, "null handle");
    if (mode < 0 || mode > 3) return fail(CTS_ERR_ARG, "cts_set_field_model: mode 0..3");
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    h->sur = mode ? s : nullptr;
    h->sur_mode = s ? mode : 0;
    if ((h->sur_mode & 2) && !h->d_fb_list) {
        int rc;
        if ((rc = dalloc(&h->d_fb_count, 1)) || (rc = dalloc(&h->d_fb_list, (size_t)h->ntot))) return rc;
    }
    return CTS_OK;
}

int cts_ion_phase(cts_handle *h, int32_t turn) { return cts_ion_phase_ip(h, turn, 0); }

int cts_ion_phase_ip(cts_handle *h, int32_t turn, int32_t k)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (k < 0 || k >= (int)h->ips.size()) return fail(CTS_ERR_ARG, "cts_ion_phase_ip: no such section");
    if (h->ips[k].nspec == 0) return fail(CTS_ERR_STATE, "cts_ion_phase: no ions at this interaction point");
    if ((rc = flush_kick(h))) return rc;
    if ((rc = ion_phase(h, turn, k))) return rc;
    return enqueue_status_copy(h);
}

int cts_get_ion_cloud(cts_handle *h, int32_t species, int64_t cap, int64_t *n, double *x, double *y, double *vx,
                      double *vy)
{
    return cts_get_ion_cloud_ip(h, 0, species, cap, n, x, y, vx, vy);
}

int cts_get_ion_cloud_ip(cts_handle *h, int32_t k, int32_t species, int64_t cap, int64_t *n, double *x, double *y,
                         double *vx, double *vy)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (k < 0 || k >= (int)h->ips.size()) return fail(CTS_ERR_ARG, "cts_get_ion_cloud: no such section");
    const IonIp &Ip = h->ips[k];
    if (species < 0 || species >= Ip.nspec || !n) return fail(CTS_ERR_ARG, "cts_get_ion_cloud: bad species");
    HIPCHK(hipStreamSynchronize(h->stream));
    int64_t len = 0;
    HIPCHK(hipMemcpy(&len, Ip.d_len + Ip.buf * CTS_MAX_SPECIES + species, sizeof len, hipMemcpyDeviceToHost));
    *n = len;
    const int64_t m = std::min(len, cap);
    double *dst[4] = {x, y, vx, vy};
    for (int c = 0; c < 4; c++)
        if (dst[c] && m > 0)
            HIPCHK(hipMemcpy(dst[c], Ip.d_ion[Ip.buf][species][c], (size_t)m * sizeof(double),
                             hipMemcpyDeviceToHost));
    return CTS_OK;
}

int cts_get_ion_passages(cts_handle *h, double *out) { return cts_get_ion_passages_ip(h, 0, out); }

int cts_get_ion_passages_ip(cts_handle *h, int32_t k, double *out)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (k < 0 || k >= (int)h->ips.size() || h->ips[k].nspec == 0 || !out)
        return fail(CTS_ERR_STATE, "cts_get_ion_passages: no ions at this interaction point");
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(out, h->ips[k].d_cloud, (size_t)h->nb * CTS_MAX_SPECIES * 6 * sizeof(double),
                     hipMemcpyDeviceToHost));
    return CTS_OK;
}

int cts_get_ion_charge(cts_handle *h, double *total)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!total) return fail(CTS_ERR_ARG, "cts_get_ion_charge: null output");
    HIPCHK(hipStreamSynchronize(h->stream));
    double sum = 0.0;
    for (const IonIp &Ip : h->ips) {
        if (Ip.nspec == 0) continue;
        int64_t n[CTS_MAX_SPECIES] = {};
        HIPCHK(hipMemcpy(n, Ip.d_nalive, Ip.nspec * sizeof(int64_t), hipMemcpyDeviceToHost));
        // survivors of the last passage times the last bunch's macro-ion charge
        for (int p = 0; p < Ip.nspec; p++) sum += (double)n[p] * Ip.q[(size_t)(h->nb - 1) * Ip.nspec + p];
    }
    *total = sum;
    return CTS_OK;
}

int cts_rms(cts_handle *h, cts_moments *out)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if ((rc = enqueue_moments(h))) return rc;
    if (out) {
        HIPCHK(hipMemcpyAsync(out, h->d_cur, h->nb * sizeof(cts_moments), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return CTS_OK;
}

int cts_get_moments(cts_handle *h, cts_moments *out)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if ((rc = flush_kick(h))) return rc;
    HIPCHK(hipMemcpyAsync(out, h->d_cur, h->nb * sizeof(cts_moments), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CTS_OK;
}

int cts_set_history(cts_handle *h, int32_t capacity)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (capacity < 0) return fail(CTS_ERR_ARG, "cts_set_history: capacity < 0");
    HIPCHK(hipStreamSynchronize(h->stream));
    graph_reset(h);
    if (h->d_hist) (void)hipFree(h->d_hist);
    h->d_hist = nullptr;
    h->hist_cap = capacity;
    if (capacity > 0) {
        if ((rc = dalloc(&h->d_hist, (size_t)capacity * h->nb))) return rc;
    }
    return CTS_OK;
}

int cts_get_history(cts_handle *h, int32_t first, int32_t nturns, cts_moments *out)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (h->hist_cap == 0) return fail(CTS_ERR_STATE, "cts_get_history: no history buffer");
    if (first < 0 || nturns < 0 || first + nturns - 1 > h->last_turn || h->last_turn - first >= h->hist_cap)
        return fail(CTS_ERR_ARG, "cts_get_history: turns not held");
    if ((rc = flush_kick(h))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int t = first; t < first + nturns; t++) {
        const int s = t % h->hist_cap;
        HIPCHK(hipMemcpy(out + (size_t)(t - first) * h->nb, h->d_hist + (size_t)s * h->nb,
                         h->nb * sizeof(cts_moments), hipMemcpyDeviceToHost));
    }
    return CTS_OK;
}

int cts_growth_rates(cts_handle *h, int32_t first, int32_t interval, int32_t nsamples, int32_t fit_start,
                     int32_t fit_end, double nuz, double beta_z, double *rates)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!rates || interval < 1 || nsamples < 2 || first < 0) return fail(CTS_ERR_ARG, "cts_growth_rates");
    const int last = first + (nsamples - 1) * interval;
    if (h->hist_cap == 0 || last > h->last_turn || h->last_turn - first >= h->hist_cap)
        return fail(CTS_ERR_ARG, "cts_growth_rates: turns not in the history");
    if ((rc = flush_kick(h))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<cts_moments> m((size_t)nsamples * h->nb);
    for (int n = 0; n < nsamples; n++) {
        const int t = first + n * interval;
        HIPCHK(hipMemcpy(m.data() + (size_t)n * h->nb, h->d_hist + (size_t)(t % h->hist_cap) * h->nb,
                         h->nb * sizeof(cts_moments), hipMemcpyDeviceToHost));
    }
    std::vector<int32_t> harm(h->nb, 0);
    for (int b = 1; b < h->nb; b++) harm[b] = harm[b - 1] + h->gap[b - 1];
    const cts_section &S = h->secs[0];
    const double tunes[3] = {h->ring.nux, h->ring.nuy, nuz};
    const double twiss[5] = {S.alpha_x, S.beta_x, S.alpha_y, S.beta_y, beta_z};
    rc = cts_cbm_growth_from_moments(m.data(), nsamples, h->nb, harm.data(), h->ring.harmonic, tunes, twiss, interval,
                                     fit_start, fit_end, h->ring.t0, rates, nullptr);
    if (rc) return fail(rc, "cts_growth_rates: fit window outside the samples");
    return CTS_OK;
}

int cts_get_wake_state(cts_handle *h, int32_t *counts, double *edges, double *kick)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (h->nbins == 0) return fail(CTS_ERR_STATE, "wake is off");
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t n = (size_t)h->nb * h->nbins;
    if (counts) HIPCHK(hipMemcpy(counts, h->d_counts, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (edges) HIPCHK(hipMemcpy(edges, h->d_edges_prev, 2 * h->nb * sizeof(double), hipMemcpyDeviceToHost));
    if (kick) HIPCHK(hipMemcpy(kick, h->d_kick, n * sizeof(double), hipMemcpyDeviceToHost));
    return CTS_OK;
}

int cts_derive_moments(cts_handle *h, const cts_moments *m, int32_t k, cts_derived *o)
{
    if (!h || !m || !o || k < 0 || k >= (int)h->secs.size()) return fail(CTS_ERR_ARG, "cts_derive_moments");
    const double *c = m->c2, *mu = m->mean;
    // src/MPBunch.cpp:431-485
    o->emit[0] = std::sqrt(c[0] * c[2] - c[1] * c[1]);
    o->emit[1] = std::sqrt(c[3] * c[5] - c[4] * c[4]);
    o->emit[2] = std::sqrt(c[6] * c[8] - c[7] * c[7]);
    const double ux = c[0] + mu[0] * mu[0], upx = c[2] + mu[1] * mu[1], uxpx = c[1] + mu[0] * mu[1];
    const double uy = c[3] + mu[2] * mu[2], upy = c[5] + mu[3] * mu[3], uypy = c[4] + mu[2] * mu[3];
    const double uz = c[6] + mu[4] * mu[4], ud = c[8] + mu[5] * mu[5], uzd = c[7] + mu[4] * mu[5];
    o->emit_eff[0] = std::sqrt(ux * upx - uxpx * uxpx);
    o->emit_eff[1] = std::sqrt(uy * upy - uypy * uypy);
    o->emit_eff[2] = std::sqrt(uz * ud - uzd * uzd);
    o->rms_x = std::sqrt(o->emit[0] * h->secs[k].beta_x);
    o->rms_y = std::sqrt(o->emit[1] * h->secs[k].beta_y);
    o->rms_eff_x = std::sqrt(o->emit_eff[0] * h->secs[k].beta_x);
    o->rms_eff_y = std::sqrt(o->emit_eff[1] * h->secs[k].beta_y);
    o->sigma_z = std::sqrt(c[6]);
    o->sigma_dp = std::sqrt(c[8]);
    return CTS_OK;
}

// Determinant by LU decomposition with partial pivoting (the largest
// magnitude in the column), in the operation order of the GSL LU
// factorisation behind GetEigenEmit's get_det (src/MPBunch.cpp:606):
// multipliers by the pivot's reciprocal, the determinant as the sign times
// the product of the diagonal.
static double det_lu(int n, double *a)
{
    int sign = 1;
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int i = k + 1; i < n; i++)
            if (std::fabs(a[i * n + k]) > std::fabs(a[piv * n + k])) piv = i;
        if (piv != k) {
            for (int j = 0; j < n; j++) std::swap(a[k * n + j], a[piv * n + j]);
            sign = -sign;
        }
        const double akk = a[k * n + k];
        if (akk == 0.0) return 0.0;
        const double inv = 1.0 / akk;
        for (int i = k + 1; i < n; i++) {
            const double l = a[i * n + k] * inv;
            for (int j = k + 1; j < n; j++) a[i * n + j] -= l * a[k * n + j];
        }
    }
    double det = sign;
    for (int k = 0; k < n; k++) det *= a[k * n + k];
    return det;
}

// std::accumulate(first, last, 0): the int running sum truncates every
// partial sum (src/MPBunch.cpp:522-525).
static double accumulate_int(const std::vector<double> &v)
{
    int s = 0;
    for (double e : v) s = (int)(s + e);
    return (double)s;
}

int cts_eigen_emit(cts_handle *h, double *out)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!out) return fail(CTS_ERR_ARG, "cts_eigen_emit: null output");
    if ((rc = flush_kick(h))) return rc;
    const int nb = h->nb;
    int64_t nmax = 0;
    for (int64_t c : h->counts) nmax = std::max(nmax, c);
    const int nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(256, (nmax + 8191) / 8192));
    const size_t nsum = (size_t)nb * nchunk * 15;
    if (h->eig_cap < nsum) {
        if (h->d_eig) (void)hipFree(h->d_eig);
        h->d_eig = nullptr;
        h->eig_cap = 0;
        if ((rc = dalloc(&h->d_eig, nsum))) return rc;
        h->eig_cap = nsum;
    }
    const cts_section &S0 = h->secs[0];
    const double eta[4] = {S0.eta_x, S0.etap_x, S0.eta_y, S0.etap_y};
    HIPCHK(launch_eigen_sums(h->d_x, h->d_px, h->d_y, h->d_py, h->d_dp, h->d_bunch_off, nb, nchunk, eta, h->d_eig,
                             h->stream));
    std::vector<double> part(nsum);
    HIPCHK(hipMemcpyAsync(part.data(), h->d_eig, nsum * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int b = 0; b < nb; b++) {
        const double n = (double)h->counts[b];
        double s[10] = {0}, mx = 0.0;
        for (int c = 0; c < nchunk; c++) {
            const double *q = part.data() + ((size_t)b * nchunk + c) * 15;
            for (int k = 0; k < 10; k++) s[k] += q[k];
            mx = std::max(mx, q[14]);
        }
        if (mx >= 1.0) {
            // a coordinate of 1 or more: the integer accumulate can leave a
            // non-zero "mean"; the bunch's moments are taken on the host as
            // the reference takes them
            const size_t m = (size_t)h->counts[b];
            std::vector<double> X[5];
            double *src[5] = {h->d_x, h->d_px, h->d_y, h->d_py, h->d_dp};
            for (int k = 0; k < 5; k++) {
                X[k].resize(m);
                HIPCHK(hipMemcpy(X[k].data(), src[k] + h->offsets[b], m * sizeof(double), hipMemcpyDeviceToHost));
            }
            for (size_t i = 0; i < m; i++) {
                X[0][i] -= eta[0] * X[4][i];
                X[1][i] -= eta[1] * X[4][i];
                X[2][i] -= eta[2] * X[4][i];
                X[3][i] -= eta[3] * X[4][i];
            }
            // the reference's order: x, y, xp, yp
            const double ax = accumulate_int(X[0]), ay = accumulate_int(X[2]);
            const double axp = accumulate_int(X[1]), ayp = accumulate_int(X[3]);
            for (double &v : s) v = 0.0;
            for (size_t i = 0; i < m; i++) {
                const double u0 = X[0][i] - ax, u1 = X[1][i] - axp, u2 = X[2][i] - ay, u3 = X[3][i] - ayp;
                s[0] += u0 * u0; s[1] += u0 * u1; s[2] += u0 * u2; s[3] += u0 * u3; s[4] += u1 * u1;
                s[5] += u1 * u2; s[6] += u1 * u3; s[7] += u2 * u2; s[8] += u2 * u3; s[9] += u3 * u3;
            }
        }
        for (double &v : s) v /= n;
        // Sigma (x, px, y, py), J, tr((Sigma J)^2), det Sigma
        const double Sg[16] = {s[0], s[1], s[2], s[3], s[1], s[4], s[5], s[6],
                               s[2], s[5], s[7], s[8], s[3], s[6], s[8], s[9]};
        double CJ[16];
        for (int i = 0; i < 4; i++) {
            // (Sigma J)[i][j]: J has +1 at (0,1), (2,3) and -1 at (1,0), (3,2)
            CJ[i * 4 + 0] = -Sg[i * 4 + 1];
            CJ[i * 4 + 1] = Sg[i * 4 + 0];
            CJ[i * 4 + 2] = -Sg[i * 4 + 3];
            CJ[i * 4 + 3] = Sg[i * 4 + 2];
        }
        double tr = 0.0;
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 4; k++) tr += CJ[i * 4 + k] * CJ[k * 4 + i];
        double A[16];
        std::memcpy(A, Sg, sizeof A);
        const double det = det_lu(4, A);
        double *o = out + 4 * b;
        o[0] = std::sqrt(-tr + std::sqrt(tr * tr - 16 * det)) / 2;
        o[1] = std::sqrt(-tr - std::sqrt(tr * tr - 16 * det)) / 2;
        o[2] = std::sqrt(std::sqrt(det));
        // x-y tilt of the beam ellipse (src/MPBunch.cpp:642-656)
        const double a = s[0], bb = s[2], c = s[7];
        const double lam = (a + c) / 2 + std::sqrt((a - c) * (a - c) / 4 + bb * bb);
        o[3] = bb == 0 ? (a >= c ? 0.0 : kPi / 2) : std::atan2(lam - a, bb);
    }
    return CTS_OK;
}

int cts_sync(cts_handle *h)
{
    if (!h) return fail(CTS_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return check_status(h, true);
}

int cts_set_timing(cts_handle *h, int32_t on)
{
    if (!h) return fail(CTS_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    resolve_timers(h);
    h->timers.clear();
    h->timing = on != 0;
    return CTS_OK;
}

int cts_kernel_time(cts_handle *h, const char *name, double *avg_ms, int64_t *launches)
{
    if (!h || !name) return fail(CTS_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(h->stream));
    resolve_timers(h);
    auto it = h->timers.find(name);
    if (it == h->timers.end() || it->second.n == 0) {
        if (avg_ms) *avg_ms = 0;
        if (launches) *launches = 0;
        return CTS_OK;
    }
    if (avg_ms) *avg_ms = it->second.total_ms / (double)it->second.n;
    if (launches) *launches = it->second.n;
    return CTS_OK;
}

int cts_debug_faddeeva(cts_handle *h, int64_t n, const double *re, const double *im, double *wr, double *wi)
{
    if (!h || n <= 0 || !re || !im || !wr || !wi) return fail(CTS_ERR_ARG, "cts_debug_faddeeva");
    double *d = nullptr;
    int rc = dalloc(&d, (size_t)n * 4);
    if (rc) return rc;
    hipError_t e = hipMemcpy(d, re, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, im, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_w_debug(d, d + n, n, d + 2 * n, d + 3 * n, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(wr, d + 2 * n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(wi, d + 3 * n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CTS_ERR_HIP, std::string("faddeeva debug: ") + hipGetErrorString(e));
    return CTS_OK;
}

int cts_debug_kick_field(cts_handle *h, int64_t n, const double *in, double *out)
{
    if (!h || n <= 0 || !in || !out) return fail(CTS_ERR_ARG, "cts_debug_kick_field");
    double *d = nullptr;
    int rc = dalloc(&d, (size_t)n * 6);
    if (rc) return rc;
    hipError_t e = hipMemcpy(d, in, (size_t)n * 4 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_kick_field_debug(d, n, d + 4 * n, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d + 4 * n, (size_t)n * 2 * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CTS_ERR_HIP, std::string("kick field debug: ") + hipGetErrorString(e));
    return CTS_OK;
}

int cts_debug_qe_normals(cts_handle *h, uint64_t seed, int32_t turn, uint64_t base, int64_t n, double *noise_out)
{
    if (!h || n <= 0 || !noise_out) return fail(CTS_ERR_ARG, "cts_debug_qe_normals");
    double *d = nullptr;
    int rc = dalloc(&d, (size_t)n * 5);
    if (rc) return rc;
    hipError_t e = launch_qe_debug(h->d_ntab, seed, (uint32_t)turn, base, n, d, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(noise_out, d, (size_t)n * 5 * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CTS_ERR_HIP, std::string("qe debug: ") + hipGetErrorString(e));
    return CTS_OK;
}

}  // extern "C"
