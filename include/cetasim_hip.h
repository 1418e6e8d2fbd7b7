/*
 * cetasim_hip.h -- C ABI of the MI355X multi-particle tracking turn.
 *
 * The reference has no plugin or FFI layer for this path: its seams are the
 * MPBeam member functions called from MPBeam::Run (src/MPBeam.cpp:1010-1173).
 * Each entry point below names the seam it replaces.  A host integration
 * keeps MPBeam/MPBunch/LatticeInterActionPoint/WakeFunction and calls these
 * functions instead of looping over beamVec.
 *
 * Conventions
 *   - every function returns CTS_OK (0) or a negative/positive status code;
 *     cts_last_error() returns the message of the last failure;
 *   - all pointers are host pointers; the library owns device memory;
 *   - particle arrays are structure-of-arrays, all bunches concatenated in
 *     bunch order (bunch b occupies [offset_b, offset_b + count_b));
 *   - loss codes are the reference's eSurive values (0 alive, 1 transverse,
 *     2 longitudinal), src/Bunch.cpp:172-200;
 *   - one HIP stream per handle; a handle is not thread-safe.
 */
#ifndef CETASIM_HIP_H
#define CETASIM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTS_OK 0
#define CTS_ERR_ARG (-1)
#define CTS_ERR_HIP (-2)
#define CTS_ERR_STATE (-3)
/* transmission < 0.5: the reference prints a message and exits
 * (src/Bunch.cpp:193-198) */
#define CTS_LOW_TRANSMISSION 1
/* ion cloud above ionCalIonMaxNumber: the reference exits
 * (src/LatticeInterActionPoint.cpp:1041-1049) */
#define CTS_ION_OVERFLOW 2
/* long-range wake: a source bunch lies ahead of the receiving one in time
 * (tau > 0): the reference exits (src/WakeFunction.cpp:206-210) */
#define CTS_LRWAKE_TAU 3

#define CTS_MAX_RF 4
#define CTS_MAX_RES 8
#define CTS_MAX_SPECIES 4

typedef struct cts_handle cts_handle;

/* Ring constants: ReadInputSettings derived quantities
 * (src/ReadInputSettings.cpp:986-1077) and the cavity phasors of
 * CavityResonator::Initial (src/CavityResonator.cpp:121-132). */
typedef struct {
    double circ;            /* circRing [m] */
    double energy;          /* electronBeamEnergy [eV] */
    double rbeta, rgamma;   /* rBeta, rGamma */
    double t0, f0;          /* revolution period [s], frequency [Hz] */
    int32_t harmonic;       /* ringHarm */
    int32_t n_rf;           /* resNum (rfMode 0 cavities) */
    double u0;              /* energy loss per turn [eV] */
    double alphac[3];       /* momentum compaction a1 a2 a3 */
    double nux, nuy;        /* workQx, workQy */
    double chrom[2];        /* chromaticity */
    double adtx[5], adty[5];/* amplitude-dependent tune shift coefficients */
    double skew_k;          /* thin skew quadrupole K (0: off) */
    double synrad_tau[3];   /* damping times [turns] */
    double emit_nat[3];     /* natural emittances (emitNat) */
    double rf_vr[CTS_MAX_RF], rf_vi[CTS_MAX_RF]; /* resCavVolReq [V] */
    int32_t rf_harm[CTS_MAX_RF];                 /* resHarm */
    double ap_x, ap_y;      /* pipeAperatureX/Y[0] used by the loss test */
} cts_ring;

/* One section / interaction point k (LatticeInterActionPoint,
 * src/LatticeInterActionPoint.cpp:181-431). */
typedef struct {
    double b1h1[36];        /* symplecticMapB1H1[k], row-major */
    double h1b1[36];        /* symplecticMapInvH1InvB1[k] */
    double h2b2[36];        /* symplecticMapInvH2InvB2[k] */
    double eta_x, etap_x, eta_y, etap_y;
    double alpha_x, beta_x, alpha_y, beta_y;
    double phase_adv_x, phase_adv_y; /* phaseAdvX12[k], phaseAdvY12[k] */
} cts_section;

/* Per-bunch moments, the quantities of MPBunch::GetMPBunchRMS
 * (src/MPBunch.cpp:342-498) in raw form: centred second moments c2 =
 * <xx> <xpx> <pxpx> <yy> <ypy> <pypy> <zz> <zdp> <dpdp>.  Emittances and
 * sizes follow from them as in the reference (cts_derive_moments). */
typedef struct {
    double n_alive;
    double mean[6];         /* x px y py z dp */
    double c2[9];
    double zmin, zmax;      /* survivors' z range, without the +-1e-6 pad */
    double loss_count;      /* MarkLostParticle's count (double counting kept) */
    double transmission;
} cts_moments;

/* Derived quantities of GetMPBunchRMS. */
typedef struct {
    double emit[3], emit_eff[3];
    double rms_x, rms_y, rms_eff_x, rms_eff_y;
    double sigma_z, sigma_dp;
} cts_derived;

/* ---------------------------------------------------------------- setup */
int cts_create(const cts_ring *ring, const cts_section *secs, int32_t nsec,
               cts_handle **out);
int cts_destroy(cts_handle *h);
const char *cts_last_error(void);
/* Device and build description (for logs). */
int cts_device_info(char *buf, int32_t len);

/* MPBeam::Initial (src/MPBeam.cpp:64-161): bunch layout. counts[b] =
 * macroEleNumPerBunch of bunch b (uneven counts allowed), macro_charge[b] =
 * macroEleCharge (electrons per macroparticle), n_e[b] = electronNumPerBunch,
 * gap[b] = bunchGap. */
int cts_set_bunches(cts_handle *h, int32_t nb, const int64_t *counts,
                    const double *macro_charge, const double *n_e, const int32_t *gap);

/* Copy particle state in / out (the host vectors of Bunch,
 * include/Bunch.h:43-55).  lost may be NULL on upload (all alive).
 * Download applies any wake kick still pending from cts_turn first. */
int cts_upload(cts_handle *h, const double *x, const double *px, const double *y,
               const double *py, const double *z, const double *dp, const int32_t *lost);
int cts_download(cts_handle *h, double *x, double *px, double *y, double *py,
                 double *z, double *dp, int32_t *lost);

/* WakeFunction::InitialSRWake / BBRWakeParaReadIn (src/WakeFunction.cpp:
 * 232-290): broadband resonators (Rs [Ohm], Q, fr [Hz]) and the binning of
 * MPBunch::BunchTransferDueToSRWake.  mode 0: the reference's edges
 * [zmin-2*sigz, zmax+2*sigz]; mode 1: fixed grid [grid_lo, grid_hi].
 * nbins <= 1024.  nres = 0 switches the wake off. */
int cts_set_wake(cts_handle *h, int32_t nbins, int32_t mode, double grid_lo,
                 double grid_hi, int32_t nres, const double *rs, const double *q,
                 const double *fr_hz);

/* Long-range (coupled-bunch) wake: WakeFunction::InitialLRWake
 * (src/WakeFunction.cpp:49-205) with RWWakeParaReadIn / BBRWakeParaReadIn
 * (src/WakeFunction.cpp:251-350).  The wake table W(tau) of the resistive
 * wall (GetRWLRWakeFun, src/WakeFunction.cpp:499-555) plus the broadband
 * resonators (GetBBRWakeFun, src/WakeFunction.cpp:655-720) is tabulated at
 * tau = -i * T_rf / 2, i < np = int(nturns * h * T_rf / (T_rf / 2)) + 2, and
 * every turn MPBeam::LRWakeBeamIntaction (src/MPBeam.cpp:3232-3351) sums it
 * over the bunches' post-map centroids of the last nturns turns into one
 * px, py, delta kick per bunch (Bunch::BunchTransferDueToWake,
 * src/Bunch.cpp:1091-1101), applied with the turn's deferred kicks.
 * nturns = lrwNTurnsWakeTrunction + 1 (the value the reference stores);
 * 0 switches the wake off.  Fused turn (one section) only. */
typedef struct {
    int32_t nturns;
    int32_t n_sectors;       /* resistive-wall pipe sectors (0: none) */
    const double *sectors;   /* [n_sectors][7]: radiusX, radiusY [m], length [m], betaX,
                                betaY [m], number, conductivity [1/(Ohm m)] (the pipe file's columns) */
    int32_t n_res;           /* broadband resonators (0: none) */
    const double *res;       /* [n_res][9]: Rs [Ohm], Q, f [Hz], x Rs [Ohm/m^2], Q, f, y Rs, Q, f
                                (the BBR file's columns) */
    double beta_x, beta_y;   /* beta functions at the first twiss point [m] */
} cts_lrwake_config;
int cts_set_lrwake(cts_handle *h, const cts_lrwake_config *cfg);
/* The wake table (time[np] >= 0 and w[3][np]: W_x, W_y [V/C/m], W_z [V/C]
 * at tau = -time; rw[3][np] and bbr[3][np] its two parts, may be null);
 * *np in: capacity, out: np. */
int cts_get_lrwake_table(cts_handle *h, int32_t *np, double *time, double *w, double *rw, double *bbr);
/* The last turn's long-range kick of every bunch: kick[nb][3] (px, py, delta). */
int cts_get_lrwake_kick(cts_handle *h, double *kick);

/* Beam loading of rfMode-1 cavities (transient, bunch by bunch):
 * CavityResonator::Initial (src/CavityResonator.cpp:36-176), the initial
 * beam-induced and generator voltages of MPBeam::InitialcavityResonator
 * (src/MPBeam.cpp:236-465), the cavity's evolution over the train
 * (Resonator::ResonatorDynamics / GetBeamInducedVol,
 * src/Resonator.cpp:205-261, driven by MPBeam::BeamMomtumUpdateDueToRFTest,
 * src/MPBeam.cpp:2624-2706) and the kick of
 * MPBunch::BunchMomentumUpdateDueToRFModeStable (src/MPBunch.cpp:1323-1417):
 * survivors binned over the bunch's z range after the section map get
 * delta += Re(V exp(-i z_bin 2 pi h f_rf / (c beta))) / (E beta^2) +
 * Re(V_b0) / 2 / (E beta^2).  The cavity voltages do not depend on the
 * particles in this model, so they are tabulated on the host for turns
 * 0 .. nturns-1.  The ring's rf_* fields keep the rfMode-0 cavities; this
 * call lists every cavity in input order (rfMode-0 entries only enter the
 * cavity set-up).  nres = 0 switches it off.
 * With excite_instability = 1 on a cavity (rfResExciteIntability,
 * MPBunch::BunchMomentumUpdateDueToRFMode, src/MPBunch.cpp:1419-1514) each
 * bin's own charge induces V_b0 = -2 pi f_r R_s / Q_0 q n_bin e, the later
 * bins of the bunch see it decayed and turned bin by bin, and the voltage
 * decays to the next bunch over the time from this bunch's tail to the
 * next one's head (MPBeam::GetTimeDisToNextBunch, src/MPBeam.cpp:2716-2752);
 * with dir_fb = 1 (rfResDirFB, Bunch::GetLongiKickDueToCavFB,
 * src/Bunch.cpp:431-449) every particle of a bunch, lost ones included,
 * gets Re(g |dV| exp(i(arg dV + phase + 2 pi f_r delay))) / (E beta^2),
 * dV = V_req - V_cav sampled at the bunch's bucket
 * (Resonator::GetResonatorInfoAtNTrf, src/Resonator.cpp:263-275).  Either
 * makes the cavity state depend on the particles: it then lives on the
 * device and is walked bunch by bunch every turn (no turn limit; a second
 * read-only pass bins the bunches over their z range first), and the
 * cavity values of turns 0 .. nturns-1 are recorded as they run. */
typedef struct {
    int32_t nres;
    int32_t rf_mode[CTS_MAX_RF];      /* 0 ideal, 1 beam loading */
    int32_t res_type[CTS_MAX_RF];     /* rfResType: 0 passive, 1 active */
    int32_t res_harm[CTS_MAX_RF];     /* rfResHarm */
    double rs[CTS_MAX_RF];            /* rfResShuntImpRs [Ohm] */
    double q0[CTS_MAX_RF];            /* rfResQualityQ0 */
    double coupling[CTS_MAX_RF];      /* rfResCouplingBeta */
    double detune_hz[CTS_MAX_RF];     /* rfResDetuneFre [Hz] */
    double vol[CTS_MAX_RF];           /* rfResVol [V] */
    double phase[CTS_MAX_RF];         /* rfResPhase [rad] (single cavity: set from U0) */
    int32_t cold[CTS_MAX_RF];         /* rfResCold */
    int32_t excite_instability[CTS_MAX_RF];  /* rfResExciteIntability */
    int32_t dir_fb[CTS_MAX_RF];              /* rfResDirFB */
    int32_t nbins;                    /* rfBunchBinNum; (rfMode-1 cavities) x (nbins + 1) <= 4096, x (2 nbins + 1)
                                         when the state is walked on the device */
    int32_t nturns;                   /* turns tabulated (stable model without feedback) */
    double dir_fb_gain[CTS_MAX_RF];   /* rfResDirFBGain */
    double dir_fb_phase[CTS_MAX_RF];  /* rfResDirFBPhase [rad] */
    double dir_fb_delay[CTS_MAX_RF];  /* rfResDirFBDelay [s] */
} cts_beam_loading_config;
int cts_set_beam_loading(cts_handle *h, const cts_beam_loading_config *cfg);
/* What bunch b saw from every cavity at turn t (the reference's
 * bunchRFModeInfo): info[nb][nres][8] = cavity, beam-induced, generator
 * and self-loss voltages (re, im); ideal cavities give zeros here. */
int cts_get_cavity_info(cts_handle *h, int32_t turn, double *info);

/* Broadband impedance in the frequency domain, longitudinal plane
 * (MPBunch::BBImpBunchInteraction, src/MPBunch.cpp:2143-2216, with the
 * impedance table of BoardBandImp::ReadInImp, src/BoardBandImp.cpp:42-90,
 * and the profile smoothing of GetSmoothedBunchProfileGassionFilter,
 * src/MPBunch.cpp:2348-2410).  zz[k] = Z_||(f_k) for the file's nimp
 * frequencies (the reference's zZImp: column 2 - i column 7); f_max is the
 * last frequency.  Every turn, after the radiation stage, each bunch's
 * survivors are binned on the centred grid of 2 nimp - 1 bins
 * (dz = c beta / (2 f_max)), the profile over one RF bucket is smoothed by
 * a 51-point Gaussian filter (alpha 10, end values padded, sum kept), the
 * induced voltage -IDFT(DFT(profile) Z) is formed as a circular convolution
 * with the real kernel of Z, and delta += V(bin) / (E beta^2).  The
 * quadrupolar planes (flags[3], flags[4], src/MPBunch.cpp:2218-2250) use
 * the same profile with i Z_Qx, i Z_Qy and kick px += V_Qx(bin) x /
 * (E beta^2 beta_x), py likewise.  The dipolar planes (flags[1], flags[2],
 * src/MPBunch.cpp:2252-2312) use the x- and y-weighted profiles (charge
 * times x summed per bin, smoothed the same way) with i Z_Dx, i Z_Dy and
 * kick px += V_Dx(bin) / (E beta^2 beta_x), py likewise, after the
 * quadrupolar kicks; the per-bin sums are two-limb fixed-point integers, so
 * they do not depend on the order of the atomics.  The grid holds at most
 * 16383 bins (nimp <= 8192).
 * time_domain 1 (the reference's default; MPBunch::BBImpBunchInteractionTD,
 * src/MPBunch.cpp:2029-2142, dispatched at src/MPBeam.cpp:1337): the
 * quasi-green wake potential of MPBeam::GetQuasiWakePoten
 * (src/MPBeam.cpp:748-812: the wake file's columns times -1, positions
 * from 0, the longitudinal wake folded about z = 0) is passed as nwake rows
 * wake_pos, wake_z, wake_dx, wake_dy, wake_qx, wake_qy (nimp, f_max and the
 * impedance tables are not read).  Every turn each bunch's survivors are
 * binned over one RF bucket, (int)((z + L/2)/dz) for the bucket length
 * L = c T0/h and dz = |wake_pos[1] - wake_pos[0]|, on (int)(L/dz) <= 65535
 * bins; the charge profile (charge per bin [C]) is smoothed by the same
 * Gaussian filter; per bin the mean x and y of its particles; then
 * V_i = -sum_{j >= i} w[j - i] rho_j (dipolar: rho_j <x>_j) over the
 * wake's rows, and the kicks delta += V_z/(E beta^2), px += V_Dx/(E beta^2
 * beta_x), px += V_Qx/(E beta^2 beta_x) x in that order (py likewise).
 * All flags 0 switches the impedance off. */
typedef struct {
    int32_t nimp;                     /* rows of the impedance file, >= 2 */
    double f_max;                     /* its last frequency [Hz] */
    const double *zz_re, *zz_im;      /* [nimp] Z_|| [Ohm] (flags[0]) */
    int32_t flags[5];                 /* BBIimpedSimLandTFlag: Z, Dx, Dy, Qx, Qy */
    int32_t time_domain;              /* BBIimpedSimTimeOrFreFlag */
    const double *zqx_re, *zqx_im;    /* [nimp] Z_Qx [Ohm/m] (flags[3]) */
    const double *zqy_re, *zqy_im;    /* [nimp] Z_Qy (flags[4]) */
    double beta_x, beta_y;            /* beta functions at the interaction point [m] */
    const double *zdx_re, *zdx_im;    /* [nimp] Z_Dx [Ohm/m] (flags[1]; the reference's zDxImp:
                                         -column 8 - i column 3) */
    const double *zdy_re, *zdy_im;    /* [nimp] Z_Dy (flags[2]; -column 9 - i column 4) */
    int32_t nwake;                    /* time_domain 1: rows of the wake table, >= 2 */
    const double *wake_pos;           /* [nwake] binPosZ [m] */
    const double *wake_z;             /* [nwake] wz [V/C] (flags[0]) */
    const double *wake_dx, *wake_dy;  /* [nwake] wDx, wDy [V/C/m] (flags[1], flags[2]) */
    const double *wake_qx, *wake_qy;  /* [nwake] wQx, wQy [V/C/m] (flags[3], flags[4]) */
} cts_bbimp_config;
int cts_set_bbimp(cts_handle *h, const cts_bbimp_config *cfg);
/* The induced voltages of every bunch from the last turn:
 * wake[5][nb][2 nimp - 1] in the flags' order (longitudinal, dipolar x, y,
 * quadrupolar x, y; bin k at z = -z_max + k dz; zeros for planes off), and
 * the smoothed profiles prof[3][nb][2 nimp - 1] if prof is not NULL (charge
 * [C/s], x- and y-weighted [C m/s]; zeros for dipolar planes off); n
 * receives 2 nimp - 1.  Time domain: n = (int)(L/dz) bins from -L/2, the
 * sums V of every plane in wake, and in prof the smoothed charge profile
 * [C] and the mean x, y of each bin [m] (zeros for dipolar planes off). */
int cts_get_bbimp_wake(cts_handle *h, int32_t *n, double *wake, double *prof);

/* Radiation damping switch and quantum-excitation noise source
 * (BunchSynRadDamping, src/Bunch.cpp:940-1044).  noise_mode 0: damping
 * only; 1: host-supplied normals (cts_push_noise) for parity runs;
 * 2: on-device Philox4x32-10 stream keyed by seed. */
int cts_set_synrad(cts_handle *h, int32_t on, int32_t noise_mode, uint64_t seed);
/* noise[5][ntot]: r0 r1 r2 r3 r5 of every particle for the next synrad. */
int cts_push_noise(cts_handle *h, const double *noise);

/* Residual-gas ions at interaction point 0: the ion part of
 * LatticeInterActionPoint::Initial / InitialLatticeIonInfo
 * (src/LatticeInterActionPoint.cpp:39-75, 293-324) and the ionCal* settings
 * of input.dat (src/ReadInputSettings.cpp:480-550).  pressure[p] is
 * gasPercent[p] times the IP's vacuum pressure.
 * New macro-ions come from the counter-based Philox stream keyed by seed
 * (the reference draws them from a random_device-seeded mt19937). */
typedef struct {
    int32_t nspec;                 /* ionCalGasSpec, <= CTS_MAX_SPECIES; 0 = ions off */
    int32_t n_new;                 /* ionCalMacroIonNumberGeneratedPerIP */
    int64_t max_ions;              /* ionCalIonMaxNumber, per species */
    double mass_number[CTS_MAX_SPECIES];   /* ionCalIonMassNumber */
    double cross_section[CTS_MAX_SPECIES]; /* corssSectionEI [m^2] */
    double pressure[CTS_MAX_SPECIES];      /* partial pressure at the IP [Pa] */
    double temperature;            /* [K] */
    double length;                 /* interactionLength [m] */
    double loss_boundary;          /* ionCalIonlossboundary, in largest bunch rms sizes */
    double ap_x, ap_y;             /* pipe aperture at the IP [m] */
    uint64_t seed;
} cts_ion_config;
/* Set up (or switch off, nspec = 0) the ion cloud at interaction point 0;
 * the cloud starts empty. */
int cts_set_ions(cts_handle *h, const cts_ion_config *cfg);
/* The same at interaction point k (section k, 0 <= k < nsec): a cloud per
 * interaction point, as LatticeInterActionPoint keeps one per point
 * (src/LatticeInterActionPoint.cpp:974-1228, ionAccumu*[k][p]); pressure,
 * temperature, length and apertures are those of point k. */
int cts_set_ions_ip(cts_handle *h, int32_t k, const cts_ion_config *cfg);

/* ------------------------------------------------------------ the turn
 * One turn of MPBeam::Run (src/MPBeam.cpp:1059-1102) for every bunch:
 * [beam-ion phase at IP 0] -> section map -> skew quad -> RF -> drift ->
 * U0 -> synrad -> SR wake -> loss marking -> moments.  One fused streaming
 * pass; the wake kick is applied at the start of the next pass (or by
 * cts_flush).  Returns CTS_LOW_TRANSMISSION when a bunch drops below 0.5
 * (state stays consistent). Only one section is fused (nsec == 1). */
int cts_turn(cts_handle *h, int32_t turn);
/* Apply a pending wake kick and complete the last turn's moments. */
int cts_flush(cts_handle *h);
/* Turns first .. first+nturns-1 of MPBeam::Run's loop (src/MPBeam.cpp:
 * 1010-1173), each as cts_turn.  use_graph = 1 replays them two at a time
 * from a HIP graph captured from the same kernels (the turn number is read
 * from device memory), once the loop is in its steady state (the previous
 * turn was first-1 and, with a wake, its kick is pending); other turns and
 * host-supplied noise run eagerly.  Results are bit-identical to cts_turn.
 * The graphs are rebuilt after any cts_set_* call or cts_autotune; the
 * transmission/ion status is polled after every replay. */
int cts_run_turns(cts_handle *h, int32_t first, int32_t nturns, int32_t use_graph);
/* Autotuner (no counterpart in the reference): times the turn's streaming
 * pass on a copy of the particles for two layouts of the particle range
 * (one contiguous chunk per workgroup, for workgroup counts that are
 * multiples of the device's compute units; tiles of 2048-16384 particles
 * dealt round-robin to 2-4 workgroups per compute unit) and keeps the
 * fastest; *nwg_out gets the chosen workgroup count.  Moments and bin
 * counts do not depend on the layout beyond the order of floating-point
 * partial sums. */
int cts_autotune(cts_handle *h, int32_t *nwg_out);
/* Sets the layout the autotuner chooses between (no counterpart in the
 * reference): tile = 0 cuts the particles into nwg contiguous chunks,
 * tile > 0 into tiles of that many particles (rounded up to 256) dealt
 * round-robin to min(nwg, tiles) workgroups. */
int cts_set_layout(cts_handle *h, int32_t nwg, int32_t tile);
/* The layout in use: workgroups of the streaming pass and tile (0: chunks). */
int cts_get_layout(cts_handle *h, int32_t *nwg, int32_t *tile);

/* ------------------------------------------- seams, one kernel each
 * MPBeam::BeamTransferPerInteractionPointDueToLatticeT (src/MPBeam.cpp:2613) */
int cts_section_map(cts_handle *h, int32_t k);
/* MPBeam::BeamTransferDueToSkewQuad (src/MPBeam.cpp:1210) */
int cts_skew_quad(cts_handle *h);
/* MPBeam::BeamMomtumUpdateDueToRFTest, rfMode 0 (src/MPBeam.cpp:2624) */
int cts_rf_kick(cts_handle *h);
/* MPBeam::BeamLongiPosTransferOneTurn (src/MPBeam.cpp:1204) */
int cts_drift(cts_handle *h);
/* MPBeam::BeamEnergyLossOneTurn (src/MPBeam.cpp:2708) */
int cts_energy_loss(cts_handle *h);
/* MPBeam::BeamSynRadDamping (src/MPBeam.cpp:1219); turn keys the device
 * stream in noise_mode 2. */
int cts_synrad(cts_handle *h, int32_t turn);
/* MPBeam::SRWakeBeamIntaction (src/MPBeam.cpp:2117).  sigma_z_stale[b] is
 * each bunch's rmsBunchLength as of the last MPBeamRMSCal (mode 0 only;
 * NULL uses the library's last post-map value). */
int cts_sr_wake(cts_handle *h, const double *sigma_z_stale);
/* MPBeam::MarkParticleLostInBunch (src/MPBeam.cpp:1358); transmission[nb]
 * may be NULL. */
int cts_mark_lost(cts_handle *h, double *transmission);
/* MPBeam::SSBeamIonEffectOneInteractionPoint at IP 0 (src/MPBeam.cpp:
 * 2585-2611) preceded by its MPBeamRMSCal/MPGetBeamInfo (src/MPBeam.cpp:
 * 1061-1063): every bunch passes the cloud in train order (generation,
 * cloud moments, Bassetti-Erskine kicks both ways, ion drift and loss);
 * the electrons' kicks are applied.  turn keys the ion generation stream.
 * cts_turn calls this itself when ions are on. */
int cts_ion_phase(cts_handle *h, int32_t turn);
/* The ion phase at interaction point k, before section map k
 * (MPBeam::SSBeamIonEffectOneInteractionPoint(k), src/MPBeam.cpp:1059-1066):
 * bunch views from the present moments with section k's beta functions. */
int cts_ion_phase_ip(cts_handle *h, int32_t turn, int32_t k);
/* The cloud of one species after the last ion phase (survivors in order):
 * *n gets the count; arrays of `cap` entries may be NULL. */
int cts_get_ion_cloud(cts_handle *h, int32_t species, int64_t cap, int64_t *n, double *x,
                      double *y, double *vx, double *vy);
int cts_get_ion_cloud_ip(cts_handle *h, int32_t k, int32_t species, int64_t cap, int64_t *n,
                         double *x, double *y, double *vx, double *vy);
/* Per passage of the last ion phase: out[(b * CTS_MAX_SPECIES + p) * 6 ..]
 * = macro-ion count, mean x, mean y, rms x, rms y, ion number seen by bunch
 * b (IonRMSCal(k, p) and GetTotIonCharge inputs). out has nb*CTS_MAX_SPECIES*6. */
int cts_get_ion_passages(cts_handle *h, double *out);
int cts_get_ion_passages_ip(cts_handle *h, int32_t k, double *out);
/* LatticeInterActionPoint::GetTotIonCharge after the last ion phase
 * (src/LatticeInterActionPoint.cpp:1213-1226): the ions of every
 * interaction point and species that survived its last passage, times the
 * last passage's macro-ion charge (0 with ions off). */
int cts_get_ion_charge(cts_handle *h, double *total);
/* MPBeam::MPBeamRMSCal (src/MPBeam.cpp:2126): moments of the current state
 * (applies a pending wake kick first). out[nb]. */
int cts_rms(cts_handle *h, cts_moments *out);

/* ------------------------------------------------------------ results */
/* Moments at the end of the last completed turn (out[nb]). */
int cts_get_moments(cts_handle *h, cts_moments *out);
/* Turn-by-turn moments: out[(t - first) * nb + b] for turns
 * first .. first+nturns-1 (kept for the last `capacity` turns). */
int cts_set_history(cts_handle *h, int32_t capacity);
int cts_get_history(cts_handle *h, int32_t first, int32_t nturns, cts_moments *out);
/* Last binning: counts[nb*nbins], edges[nb*2] = {poszMin, dzBin},
 * kick[nb*nbins] = srWakePoten[2] of each bunch. */
int cts_get_wake_state(cts_handle *h, int32_t *counts, double *edges, double *kick);
/* GetMPBunchRMS derived quantities from raw moments at section k. */
int cts_derive_moments(cts_handle *h, const cts_moments *m, int32_t k, cts_derived *out);
/* MPBunch::GetEigenEmit (src/MPBunch.cpp:500-658) of every bunch in its
 * present state: out[nb][4] = eigenEmitX, eigenEmitY, emitXY4D,
 * xyCouplingAlpha.  As the reference: every particle, lost ones included,
 * less the dispersion of section 0 times delta; the "means" are the
 * integer-valued std::accumulate of the coordinates (zero while every
 * coordinate is below 1 in magnitude) and are not divided by N. */
int cts_eigen_emit(cts_handle *h, double *out);

/* Wait for all work on the handle's stream. */
int cts_sync(cts_handle *h);

/* Kernel timing with HIP events on the handle's stream.  name is one of
 * "pass", "finalize", "wake_table", "bin", "ion"; avg_ms is the mean
 * duration of the launches recorded since timing was switched on. */
int cts_set_timing(cts_handle *h, int32_t on);
int cts_kernel_time(cts_handle *h, const char *name, double *avg_ms, int64_t *launches);

/* Test hook: the device's w(z) (the Faddeeva function behind every
 * Bassetti-Erskine kick) at n points. */
int cts_debug_faddeeva(cts_handle *h, int64_t n, const double *re, const double *im, double *wr, double *wi);
/* Test hook: the field the beam-ion kicks apply (MPBunch::SSIonBunchInteraction's
 * branches, src/MPBunch.cpp:1128-1147, per unit line charge) at n points:
 * in = {x[n], y[n], sx[n], sy[n]}, out = {fx[n], fy[n]}. */
int cts_debug_kick_field(cts_handle *h, int64_t n, const double *in, double *out);
/* Test hook: the device quantum-excitation stream for particles
 * [base, base+n) at `turn` (noise_out[5][n]). */
int cts_debug_qe_normals(cts_handle *h, uint64_t seed, int32_t turn, uint64_t base,
                         int64_t n, double *noise_out);

/* ------------------------------------------------------- growth rates
 * MPBeam::GetCBMGR, "ideal" coupled-bunch modes (src/MPBeam.cpp:1496-1891):
 * per sample n and bunch b the normalized phasor
 *   (x/sqrt(bx) - i (sqrt(bx) px + ax/sqrt(bx) x)) e^{-2 pi i nu h_b / h}
 * (and y, and (z/sqrt(bz) + i sqrt(bz) dp) for z), its forward DFT over the
 * bunches, and for every mode the least-squares slope of log|amplitude|
 * against turn over samples [fit_start/interval, fit_end/interval) with unit
 * weights (FittingGSL::FitALinear, src/FittingGSL.cpp:231-243), divided by
 * T0.  m[nsamples][nb] are moment records taken every `interval` turns;
 * harm_num[b] = bunchHarmNum; tunes = {nux, nuy, nuz}; twiss = {alpha_x,
 * beta_x, alpha_y, beta_y, beta_z}.  rates[3][nb] (1/s); amps[nsamples][3][nb]
 * (mode amplitudes) may be NULL. */
int cts_cbm_growth_from_moments(const cts_moments *m, int32_t nsamples, int32_t nb, const int32_t *harm_num,
                                int32_t harmonics, const double *tunes, const double *twiss, int32_t interval,
                                int32_t fit_start, int32_t fit_end, double t0, double *rates, double *amps);
/* The whole of GetCBMGR's analysis (src/MPBeam.cpp:1496-1891) of the same
 * samples; every array may be NULL.  Per sample n, plane p (x, y, z) and
 * index k ([nsamples][3][nb]):
 *   ideal_amp/arg  |mode k| of the phasors above and atan2(Re, Im) (the
 *                  reference's argument order);
 *   iq_amp/arg     |<centroid_i e^{-2 pi i f_k t_i}>_i| and its arg, f_k =
 *                  (k + nu) f0, t_i = h_i T0 / h (IQ decomposition);
 *   ana_re/im      analytic signal of bunch k's centroid history (one-sided
 *                  spectrum over all samples, GetHilbertAnalytical,
 *                  src/MPBeam.cpp:1936-2011);
 *   hil_amp/arg    modes of the phasors built from (Re, -Im) of it.
 * Rates ([3][nb], 1/s) are the slopes of log amplitude over the fit window
 * as above: ideal modes, IQ components, Hilbert modes and, per bunch,
 * |analytic signal| (the reference's BGR). */
typedef struct {
    double *ideal_amp, *ideal_arg;
    double *iq_amp, *iq_arg;
    double *ana_re, *ana_im;
    double *hil_amp, *hil_arg;
    double *rate_ideal, *rate_iq, *rate_hil, *rate_bunch;
} cts_cbm_result;
int cts_cbm_analysis(const cts_moments *m, int32_t nsamples, int32_t nb, const int32_t *harm_num, int32_t harmonics,
                     const double *tunes, const double *twiss, int32_t interval, int32_t fit_start, int32_t fit_end,
                     double t0, const cts_cbm_result *out);
/* gsl_fit_wlinear's line fit y = c0 + c1 x (the fit above). */
int cts_fit_wlinear(const double *x, const double *w, const double *y, int32_t n, double *c0, double *c1);
/* The same from the handle's moment history (cts_set_history): samples at
 * turns first, first + interval, ...; bunch harmonic numbers from the gaps
 * (bunch 0 at 0); nuz and beta_z as the caller's ReadInputSettings has
 * them (workQz, rmsBunchLength / rmsEnergySpread). */
int cts_growth_rates(cts_handle *h, int32_t first, int32_t interval, int32_t nsamples, int32_t fit_start,
                     int32_t fit_end, double nuz, double beta_z, double *rates);

/* ---------------------------------------------------- field surrogate
 * No counterpart in the reference, which evaluates every Bassetti-Erskine
 * kick through the Faddeeva function (src/Bunch.cpp:119-211).  A 3-64-64-2
 * MLP (GELU, tanh form) approximates the normalized field G = F sigma_x of
 * a Gaussian beam on |x/sigma_x|, |y/sigma_y| <= 6, sigma_y/sigma_x in
 * [0.02, 1] (axes ordered so that sigma_x >= sigma_y); it is trained on the
 * GPU from samples of the FP64 field by the library's own bf16-MFMA forward
 * and backward kernels.  Parameters: 4546 FP32 values (W1[64][3], b1[64],
 * W2[64][64], b2[64], W3[2][64], b3[2]) and norm[4] = target mean / scale. */
typedef struct cts_surrogate cts_surrogate;
int cts_sur_create(cts_surrogate **out);
int cts_sur_destroy(cts_surrogate *s);
const char *cts_sur_last_error(void);
/* deterministic uniform(-1/sqrt(fan_in), +) weights, zero biases */
int cts_sur_init_params(cts_surrogate *s, uint64_t seed);
int cts_sur_set_params(cts_surrogate *s, const float *params, const float *norm);
int cts_sur_get_params(cts_surrogate *s, float *params, float *norm);
/* n_train training records (a0 a1 a2 t0 t1, FP32) and n_val validation
 * points (|u| |v| r Gx Gy, FP64) from the device FP64 field, Philox
 * stream keyed by seed; sets norm from the training targets. */
int cts_sur_generate(cts_surrogate *s, int64_t n_train, int64_t n_val, uint64_t seed);
int cts_sur_get_samples(cts_surrogate *s, int64_t first, int64_t n, float *train5, double *val5);
/* Adam steps on random minibatches of nwg*chunks*64 samples, learning rate
 * cosine from lr_max to lr_min; *loss_out = last step's mean loss. */
int cts_sur_train(cts_surrogate *s, int32_t steps, int32_t nwg, int32_t chunks, float lr_max, float lr_min,
                  uint64_t seed, float *loss_out);
/* Gradient (4546) and mean loss of the normalized-target MSE over the given
 * records (samples5[n][5] in the cts_sur_generate format). */
int cts_sur_gradient(cts_surrogate *s, const float *samples5, int64_t n, float *grad, float *loss);
/* Raw network outputs out2[n][2] for features3[n][3]. */
int cts_sur_eval(cts_surrogate *s, const float *features3, int64_t n, float *out2);
/* cts_sur_eval without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc); no counterpart in the reference.  Inputs and outputs
 * are staged through one static 5 MiB device buffer per GPU (held until the
 * process ends): n <= 262144 points, larger n is refused with CTS_ERR_ARG.
 * Every call on a GPU shares that buffer, so calls that overlap on one GPU
 * (other threads, other surrogates) race; calls in sequence do not. */
int cts_sur_eval_noalloc(cts_surrogate *s, const float *features3, int64_t n, float *out2);
/* Field error on the validation set: per point |G_sur - G| / |G|;
 * err_out[n_val] may be NULL. */
int cts_sur_validate(cts_surrogate *s, double *max_err, double *rms_err, double *err_out);
/* Field model of the beam-ion kicks of a tracking handle: the surrogate
 * inside its domain, the exact FP64 field outside it.  mode bit 0: ions
 * kicked by the bunches (MPBunch::SSIonBunchInteraction, src/MPBunch.cpp:
 * 1113-1160), bit 1: electrons kicked by the ion cloud (1165-1206); mode 0
 * or s == NULL: exact field everywhere.  s must outlive its use. */
int cts_set_field_model(cts_handle *h, const cts_surrogate *s, int32_t mode);

#ifdef __cplusplus
}
#endif
#endif
