/*
 * cts_oracle.h -- CPU oracle for the CETASim multi-particle tracking turn.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the
 * reference's per-turn multi-particle algorithms so that the HIP path in
 * cetasim_amd/ can be checked against it.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The
 * product path never calls into this library.
 *
 * Every function cites the reference source it restates (paths relative to
 * the reference repository root).  Floating-point operation order follows
 * the reference; the library is built with -ffp-contract=off so that no
 * multiply-add is fused, matching the reference's x86-64 build.
 *
 * Pinning (what checks each function, besides the HIP path itself):
 *   orc_faddeeva_w      57 known-answer points of src/Faddeeva.cpp:1922-2090,
 *                       the reference Faddeeva.cpp compiled into oracle/_ref,
 *                       w(-z)=2exp(-z^2)-w(z), w(iy)=erfcx(y), asymptote.
 *   orc_bassetti_erskine  symmetries, far field -(x,y)/r^2, round-beam limit,
 *                       brute-force quadrature of the Gaussian field integral.
 *   orc_section_map     pure rotation with tune nu, symplecticity M^T J M = J,
 *                       period-1/frac(nu) return for rational tunes.
 *   orc_rf_kick / drift / energy_loss   explicit closed forms.
 *   orc_synrad          exact geometric decay by lambda per turn (no noise).
 *   orc_sr_wake         brute-force binning, sum of counts = N, lag-table
 *                       identity tau=(i-j)*dt.
 *   orc_rms             direct two-pass sums on small arrays.
 *   orc_philox4x32_10   Random123 known-answer vectors.
 *   orc_fit_linear      closed form OLS on an exact line.
 */
#ifndef CTS_ORACLE_H
#define CTS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- constants
 * Physical constants as the reference defines them (include/Global.h:26-40).
 */
#define ORC_CLIGHT 299792458.0
#define ORC_ECHARGE 1.602017733e-19
#define ORC_EMASS_EV 0.51099906e6
#define ORC_IONMASS_EV 931.49432e+6
#define ORC_RE 2.81794092E-15
#define ORC_KBOLTZ 1.3806505E-23

/* ------------------------------------------------------------------ lattice
 * Twiss values of one interaction point / section start
 * (src/LatticeInterActionPoint.cpp:225-243).
 */
typedef struct {
    double beta_x, alpha_x, eta_x, etap_x;
    double beta_y, alpha_y, eta_y, etap_y;
    double beta_z;
} orc_twiss;

/* Section matrices B1H1, InvH1InvB1 and InvH2InvB2 (row-major 6x6) built as
 * in LatticeInterActionPoint::InitialLatticeSympMat
 * (src/LatticeInterActionPoint.cpp:326-431). t2 is the twiss of the next
 * section start (section 0 for the last section). */
void orc_section_matrices(const orc_twiss *t1, const orc_twiss *t2,
                          double b1h1[36], double h1b1[36], double h2b2[36]);

/* C = A*B with A (M x K), B (K x N), row-major; accumulation order and the
 * skip of zero A entries follow gslcblas dgemm (GSL 2.8, source_gemm_r.h,
 * NoTrans/NoTrans branch), which the reference links (Makefile:2-3). */
void orc_dgemm(int M, int N, int K, const double *A, const double *B, double *C);

/* --------------------------------------------------------------- particles */
typedef struct {
    int64_t n;                 /* macroEleNumPerBunch */
    double *x, *px, *y, *py, *z, *dp;
    int32_t *lost;             /* eSurive: 0 alive, 1 transverse, 2 longitudinal */
    double *fx_ion, *fy_ion;   /* eFxDueToIon / eFyDueToIon (may be NULL) */
} orc_bunch;

/* Scalars of the section map (Bunch::BunchTransferDueToLatticeTSymplectic,
 * src/Bunch.cpp:625-755). */
typedef struct {
    const double *b1h1, *h2b2;      /* 36 each */
    double eta_x, etap_x, eta_y, etap_y;
    double alpha_x, beta_x, alpha_y, beta_y;
    double phase_adv_x, phase_adv_y; /* phaseAdvX12[k], phaseAdvY12[k] */
    double nux, nuy;                 /* workQx, workQy */
    double chrom_x, chrom_y;
    double adtx[5], adty[5];
} orc_map;

void orc_section_map(orc_bunch *b, const orc_map *m);

/* Thin skew quadrupole (src/Bunch.cpp:757-795). */
void orc_skew_quad(orc_bunch *b, double k);

/* Ideal RF cavity kick, rfMode = 0 (src/Bunch.cpp:453-481). One call per
 * resonator: vr + i*vi = resCavVolReq, res_harm = resHarm, f_rf = f0*h. */
void orc_rf_kick(orc_bunch *b, double vr, double vi, int res_harm, double f_rf,
                 double rbeta, double energy);

/* z -= C*(a1*d + a2*d^2 + a3*d^3) (src/Bunch.cpp:494-501). */
void orc_drift(orc_bunch *b, double circ, const double alphac[3]);

/* d -= U0/(E*beta^2) (src/Bunch.cpp:483-491). */
void orc_energy_loss(orc_bunch *b, double u0, double energy, double rbeta);

/* Radiation damping and quantum excitation (src/Bunch.cpp:940-1044).
 * tau[3] in turns, emit_nat[3]; noise is SoA [5][n] holding r0,r1,r2,r3,r5
 * for particle i at noise[j*n+i] (NULL: no excitation, as for one particle). */
void orc_synrad(orc_bunch *b, const double b1h1_0[36], const double h1b1_0[36],
                const double tau[3], const double emit_nat[3], const double *noise);

/* Loss marking (src/Bunch.cpp:172-200). Returns transmission; *count gets the
 * reference's (double-counting) loss count. */
double orc_mark_lost(orc_bunch *b, double t0, int harm, double ap_x, double ap_y,
                     int64_t *count);

/* GetZMinMax (src/MPBunch.cpp:1860-1885): min/max z over survivors -/+ 1e-6. */
void orc_zminmax(const orc_bunch *b, double *zmin, double *zmax);

/* ----------------------------------------------------------------- moments
 * GetMPBunchRMS + GetEigenEmit (src/MPBunch.cpp:342-658). */
typedef struct {
    int64_t n_alive;
    double mean[6];              /* x px y py z dp */
    double emit[3];              /* emittanceX/Y/Z (centred) */
    double emit_eff[3];          /* rmsEffectiveRingEmitX/Y/Z (uncentred) */
    double rms_x, rms_y;         /* rmsRx = sqrt(emitX*betaX[k]) */
    double rms_eff_x, rms_eff_y; /* rmsEffectiveRx/Ry */
    double sigma_z, sigma_dp;    /* rmsBunchLength, rmsEnergySpread */
    double zmin, zmax;           /* zMinCurrentTurn / zMaxCurrentTurn */
    double eigen_x, eigen_y, emit4d, xy_alpha;
    double c2[9];                /* centred <xx> <xpx> <pxpx> <yy> <ypy> <pypy> <zz> <zdp> <dpdp> */
} orc_moments;

void orc_rms(const orc_bunch *b, double beta_x_k, double beta_y_k,
             double eta_x0, double etap_x0, double eta_y0, double etap_y0,
             orc_moments *out);

/* Determinant through LU with partial pivoting (Global.cpp:180-196, which
 * calls gsl_linalg_LU_decomp / gsl_linalg_LU_det). */
double orc_det(int n, const double *a);

/* -------------------------------------------------------------- SR wake
 * BBR pseudo-Green function of a 1 mm Gaussian bunch, summed over
 * resonators (src/WakeFunction.cpp:558-652). out[0..2] = x, y, z. */
void orc_bbr_wake(double tau, int nres, const double *lrs, const double *lq,
                  const double *lomega, const double *txrs, const double *txq,
                  const double *txomega, const double *tyrs, const double *tyq,
                  const double *tyomega, double out[3]);

/* Binning + wake sum + dp kick (src/MPBunch.cpp:1887-2026).
 * mode 0: reference edges [zmin-2*sigz, zmax+2*sigz] with zmin/zmax from
 *         GetZMinMax and sigz = sigma_z_stale;
 * mode 1: fixed grid [grid_lo, grid_hi].
 * poten[3*nbins] carries srWakePoten between calls (quirk: never zeroed).
 * counts[nbins] (out), kick[nbins] (out, = poten[2*nbins+i] after scaling),
 * edges[2] (out) = {poszMin, dzBin}.  The bin is int((z-poszMin)/dzBin),
 * truncated toward zero; particles whose bin is outside [0,nbins) are
 * skipped (the reference indexes out of range there).
 * Returns the number of skipped particles. */
int64_t orc_sr_wake(orc_bunch *b, int nbins, int mode, double sigma_z_stale,
                    double grid_lo, double grid_hi, double macro_charge,
                    double energy, int nres, const double *lrs, const double *lq,
                    const double *lomega, double *poten, int32_t *counts,
                    double *edges, const double *wake_table);

/* The wake sum as the reference evaluates it, one pseudo-Green call per
 * bin pair (src/MPBunch.cpp:1983-1993): pz[i] += W((i-j) dt) counts[j] for
 * i, j < nbins, j inner.  Same values as the lag table (the same arguments),
 * at the reference's cost; the CPU baseline times it. */
void orc_wake_pairs(int nbins, double dt, int nres, const double *lrs, const double *lq,
                    const double *lomega, const int32_t *counts, double *pz);

/* Longitudinal wake table W[l] = orc_bbr_wake((l-(nbins-1))*dt)[2],
 * l = 0..2*nbins-2 (lag-indexed). */
void orc_wake_table(int nbins, double dt, int nres, const double *lrs,
                    const double *lq, const double *lomega, double *table);

/* -------------------------------------------------------------- Faddeeva */
void orc_faddeeva_w(double re, double im, double *wr, double *wi);
double orc_erfcx(double x);
double orc_w_im(double x);

/* Bunch::BassettiErskine1 (src/Bunch.cpp:119-170) and
 * Bunch::GaussianField (src/Bunch.cpp:202-211). Results in 1/m. */
void orc_bassetti_erskine(double x, double y, double sx, double sy,
                          double *fx, double *fy);
void orc_gaussian_field(double x, double y, double sx, double sy,
                        double *fx, double *fy);
/* Branch selection used by both kick directions of
 * MPBunch::SSIonBunchInteraction (src/MPBunch.cpp:1128-1147, 1181-1200). */
void orc_ion_field(double x, double y, double sx, double sy, double *fx, double *fy);

/* ------------------------------------------------------------------- RNG
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  The reference draws
 * its noise from std::random_device-seeded mt19937 (src/Bunch.cpp:989-991),
 * which cannot be reproduced; this project's counter-based stream replaces
 * it and is shared by the oracle and the HIP path. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Five standard normals (r0 r1 r2 r3 r5 of BunchSynRadDamping) for particle
 * `index` at `turn`, float Box-Muller with project-written log/sin/cos. */
void orc_qe_normals(uint64_t seed, uint32_t turn, uint64_t index, double out[5]);
/* One standard-normal Box-Muller pair from two 32-bit words. */
void orc_box_muller(uint32_t a, uint32_t b, float *n0, float *n1);
/* the quantum-excitation pair transform on 21-bit uniforms, and its
 * tables (out[1024]: sc[256][2] then lg[256][2]) */
void orc_box_muller_u21(uint32_t a, uint32_t b, float *n0, float *n1);
void orc_normal_tables(float *out);
/* Fill SoA noise[5][n] for one bunch, global particle offset `base`. */
void orc_qe_fill(uint64_t seed, uint32_t turn, uint64_t base, int64_t n, double *noise);

/* ------------------------------------------------------------- ion cloud
 * LatticeInterActionPoint ion cloud of one interaction point, one or more
 * gas species (src/LatticeInterActionPoint.cpp:974-1228). */
typedef struct {
    int nspec;
    int64_t cap;                 /* ionMaxNumberOneInterPoint */
    int64_t n[4];                /* ionAccumuNumber[k][p] */
    double *x[4], *y[4], *vx[4], *vy[4], *fx[4], *fy[4];
    double macro_charge[4];      /* macroIonCharge[k][p] */
    double mean_x[4], mean_y[4], rms_x[4], rms_y[4];
    double mass_number[4], cross_section[4], pressure[4]; /* per species */
    double temperature, length;  /* temperature[k], interactionLength[k] */
    int n_new;                   /* macroIonNumberGeneratedPerIP */
    double loss_boundary;        /* ionLossBoundary */
    double ap_x, ap_y;           /* pipeAperatureX/Y[k] */
    double circ; int harm;       /* circRing, harmonics */
} orc_ion_cloud;

/* Bunch moments the ion phase reads. */
typedef struct {
    double xmean, ymean, rms_x, rms_y;
    double n_e;                  /* electronNumPerBunch */
    int gap;                     /* bunchGap */
    double gamma;                /* rGamma */
} orc_bunch_view;

/* One bunch passage at one interaction point, in the order of
 * MPBeam::SSBeamIonEffectOneInteractionPoint (src/MPBeam.cpp:2585-2611):
 * ion count, generation (project stream, see orc_ion_generate), append,
 * moments, BE both ways, electron kick, ion kick-drift and loss.
 * size_x_max / size_y_max = strongStrongBunchInfo->bunchSizeX/YMax.
 * Returns 0, or 1 if the cloud exceeds its capacity (the reference exits).
 * Electron kicks are written to b->fx_ion/fy_ion and applied to px/py of
 * every particle (BunchTransferDueToIon, src/Bunch.cpp:213-220). */
int orc_ion_passage(orc_ion_cloud *c, orc_bunch *b, const orc_bunch_view *v,
                    double size_x_max, double size_y_max, uint64_t seed,
                    uint32_t turn, uint32_t passage, double *cloud_out /* 5*nspec or NULL */);

/* Generation of new ions of species p for one passage: Box-Muller pairs of
 * the project stream, rejection outside the 2-sigma ellipse
 * (src/LatticeInterActionPoint.cpp:985-1019). */
void orc_ion_generate(uint64_t seed, uint32_t turn, uint32_t passage, int species,
                      int n_new, double xmean, double ymean, double rms_x, double rms_y,
                      double *gx, double *gy);

/* Per-species moments (IonRMSCal(k,p), src/LatticeInterActionPoint.cpp:1113-1155). */
void orc_ion_rms(orc_ion_cloud *c, int p);

/* ------------------------------------------------------------ growth rate
 * gsl_fit_wlinear (FittingGSL::FitALinear, src/FittingGSL.cpp:231-243). */
/* GetCBMGR's analysis (ideal modes, IQ, analytic signals, Hilbert modes,
 * rates) of centroids mean[ns][nb][6] with direct DFTs; layout as
 * cts_cbm_analysis.  Returns -1 when the fit window lies outside. */
/* ------------------------------------------------------ long-range wake
 * (lrwake.c) WakeFunction::InitialLRWake's table, the Yokoya factors of a
 * pipe sector and MPBeam::LRWakeBeamIntaction's per-turn sum. */
void orc_lrw_yokoya(double rx, double ry, double out[5]);
int orc_lrw_table(int nturns, int harmonic, double t0, int nsec, const double *sectors, int nres,
                  const double *res, double beta_x, double beta_y, int cap, double *time, double *w,
                  double *rw, double *bbr);
int orc_lrw_kick(int nturns, int nb, const int *harm, int harmonic, double t0, double rbeta, double energy,
                 const double *ne, const double *pos, const double *w, int np, double *kick);

/* ------------------------------------------------------ beam loading
 * (beamload.c) rfMode-1 cavities, the reference's stable variant. */
int orc_bl_tables(int nres, const int *ip, const double *dp, int harmonic, double f0, double t0, double u0,
                  double rbeta, int nb, const double *ne, const int *gap, int nturns, double *dev, double *info);
void orc_bl_kick(orc_bunch *b, double vr, double vi, double vb0, double zmin, double zmax, int nbins, int res_harm,
                 double f_rf, double rbeta, double energy);
int orc_bl_state_size(void);
void orc_bl_state(int nres, const int *ip, const double *dp, int harmonic, double f0, double u0, int nb,
                  const double *ne, const int *gap, double *state);
void orc_bl_walk(double *state, int nres, orc_bunch *bs, int nb, const int *gap, const double *ne,
                 const double *macro_charge, int nbins, int harmonic, double f0, double t0, double rbeta,
                 double energy, double *info, double *fb);

/* Broadband impedance, longitudinal, dipolar and quadrupolar planes (bbimp.c);
 * orc_bbi_kick: planes[5], zre/zim [5][nimp] and wake [5][n] in the
 * reference's flag order (Z, Dx, Dy, Qx, Qy), prof [3][n]. */
void orc_bbi_smooth(double *p, int n);
void orc_bbi_potential(const double *p, int nimp, const double *zre, const double *zim, double *wake);
void orc_bbi_kick(orc_bunch *b, int nimp, double f_max, const int *planes, const double *zre, const double *zim,
                  double bx, double by, double macro_charge, double rbeta, double t0, int harmonic, double energy,
                  double *wake, double *prof);
/* The time-domain method: the wake table from the file's rows [n][6]
 * (out [6][n]: binPosZ, wz, wDx, wDy, wQx, wQy), the grid size and one
 * bunch's interaction; w [5][nw] in flag order, V [5][n], prof [3][n]. */
int orc_bbi_td_wake(int n, const double *rows, double *out);
int orc_bbi_td_bins(double t0, int harmonic, double dz);
void orc_bbi_td_kick(orc_bunch *b, int nw, const double *pos, const double *w, const int *planes, double bx,
                     double by, double macro_charge, double t0, int harmonic, double rbeta, double energy, double *V,
                     double *prof);

int orc_cbm_analysis(const double *mean, int ns, int nb, const int *harm, int harmonics, const double *tunes,
                     const double *twiss, int interval, int fit_start, int fit_end, double t0,
                     double *ideal_amp, double *ideal_arg, double *iq_amp, double *iq_arg, double *ana_re,
                     double *ana_im, double *hil_amp, double *hil_arg, double *rate_ideal, double *rate_iq,
                     double *rate_hil, double *rate_bunch);
void orc_fit_wlinear(const double *x, const double *w, const double *y, int n,
                     double *c0, double *c1);

#ifdef __cplusplus
}
#endif
#endif
