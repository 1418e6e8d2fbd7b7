// cts_lattice.cpp -- host-side lattice and ring set-up of the tracking turn:
// the section matrices of LatticeInterActionPoint::InitialLatticeSympMat
// (src/LatticeInterActionPoint.cpp:326-431), the derived ring quantities of
// ReadInputSettings::ParamRead (src/ReadInputSettings.cpp:986-1077) and the
// single-cavity phasor of CavityResonator::Initial
// (src/CavityResonator.cpp:121-132).
#include <cmath>
#include <cstring>

#include "../../include/cetasim_hip.h"
#include "../../include/cetasim_host.h"

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kC = 299792458.0;
constexpr double kMeEV = 0.51099906e6;

// C = A*B, row-major, gslcblas dgemm order (zero A entries skipped).
void matmul(int M, int N, int K, const double *A, const double *B, double *C)
{
    for (int i = 0; i < M * N; i++) C[i] = 0.0;
    for (int i = 0; i < M; i++)
        for (int k = 0; k < K; k++) {
            const double a = 1.0 * A[i * K + k];
            if (a != 0.0)
                for (int j = 0; j < N; j++) C[i * N + j] += a * B[k * N + j];
        }
}

void dispersion(const cts_twiss &t, double sign, double H[36])
{
    std::memset(H, 0, 36 * sizeof(double));
    for (int i = 0; i < 6; i++) H[i * 7] = 1.0;
    H[0 * 6 + 5] = sign * -t.eta_x;
    H[1 * 6 + 5] = sign * -t.etap_x;
    H[2 * 6 + 5] = sign * -t.eta_y;
    H[3 * 6 + 5] = sign * -t.etap_y;
    H[4 * 6 + 0] = sign * t.etap_x;
    H[4 * 6 + 1] = sign * -t.eta_x;
    H[4 * 6 + 2] = sign * t.etap_y;
    H[4 * 6 + 3] = sign * -t.eta_y;
}

}  // namespace

extern "C" int cts_build_section(const cts_twiss *t1, const cts_twiss *t2, double phase_adv_x,
                                 double phase_adv_y, cts_section *out)
{
    if (!t1 || !t2 || !out) return CTS_ERR_ARG;
    if (!(t1->beta_x > 0 && t1->beta_y > 0 && t1->beta_z > 0 && t2->beta_x > 0 && t2->beta_y > 0 &&
          t2->beta_z > 0))
        return CTS_ERR_ARG;
    double H1[36], iH1[36], iH2[36], B1[36], iB1[36], iB2[36];
    dispersion(*t1, 1.0, H1);
    // inverse: the same entries with opposite sign (389-395); 1*-x keeps -0
    // where the reference writes +x, so build it explicitly
    std::memset(iH1, 0, sizeof iH1);
    std::memset(iH2, 0, sizeof iH2);
    for (int i = 0; i < 6; i++) { iH1[i * 7] = 1.0; iH2[i * 7] = 1.0; }
    const cts_twiss *ts[2] = {t1, t2};
    double *iHs[2] = {iH1, iH2};
    for (int q = 0; q < 2; q++) {
        const cts_twiss &t = *ts[q];
        double *m = iHs[q];
        m[0 * 6 + 5] = t.eta_x;  m[1 * 6 + 5] = t.etap_x;
        m[2 * 6 + 5] = t.eta_y;  m[3 * 6 + 5] = t.etap_y;
        m[4 * 6 + 0] = -t.etap_x; m[4 * 6 + 1] = t.eta_x;
        m[4 * 6 + 2] = -t.etap_y; m[4 * 6 + 3] = t.eta_y;
    }
    std::memset(B1, 0, sizeof B1);
    std::memset(iB1, 0, sizeof iB1);
    std::memset(iB2, 0, sizeof iB2);
    B1[0] = 1 / std::sqrt(t1->beta_x); B1[6] = t1->alpha_x / std::sqrt(t1->beta_x); B1[7] = std::sqrt(t1->beta_x);
    B1[14] = 1 / std::sqrt(t1->beta_y); B1[20] = t1->alpha_y / std::sqrt(t1->beta_y); B1[21] = std::sqrt(t1->beta_y);
    B1[28] = 1 / std::sqrt(t1->beta_z); B1[35] = std::sqrt(t1->beta_z);
    const cts_twiss *tb[2] = {t1, t2};
    double *iBs[2] = {iB1, iB2};
    for (int q = 0; q < 2; q++) {
        const cts_twiss &t = *tb[q];
        double *m = iBs[q];
        m[0] = std::sqrt(t.beta_x); m[6] = -t.alpha_x / std::sqrt(t.beta_x); m[7] = 1 / std::sqrt(t.beta_x);
        m[14] = std::sqrt(t.beta_y); m[20] = -t.alpha_y / std::sqrt(t.beta_y); m[21] = 1 / std::sqrt(t.beta_y);
        m[28] = std::sqrt(t.beta_z); m[35] = 1 / std::sqrt(t.beta_z);
    }
    matmul(6, 6, 6, B1, H1, out->b1h1);
    matmul(6, 6, 6, iH1, iB1, out->h1b1);
    matmul(6, 6, 6, iH2, iB2, out->h2b2);
    out->eta_x = t1->eta_x; out->etap_x = t1->etap_x;
    out->eta_y = t1->eta_y; out->etap_y = t1->etap_y;
    out->alpha_x = t1->alpha_x; out->beta_x = t1->beta_x;
    out->alpha_y = t1->alpha_y; out->beta_y = t1->beta_y;
    out->phase_adv_x = phase_adv_x;
    out->phase_adv_y = phase_adv_y;
    return CTS_OK;
}

extern "C" int cts_derive_ring(const cts_ring_input *in, cts_ring *r, cts_ring_derived *d)
{
    if (!in || !r) return CTS_ERR_ARG;
    std::memset(r, 0, sizeof *r);
    const double E = in->energy;
    const double rGamma = E / kMeEV;
    const double eta = in->alphac[0] - 1. / std::pow(rGamma, 2);
    const double rBeta = std::sqrt(1.E0 - std::pow(rGamma, -2));
    const double t0 = in->circ / kC / rBeta;
    const double f0 = 1 / t0;
    const double vRF = in->rf_volt;
    const double synPhase = kPi - std::asin(in->u0 / vRF);
    const double workQz = std::sqrt(-1 * eta * in->harmonic * vRF * std::cos(synPhase) / (2 * kPi * rBeta * E));
    const double naturalBunchLength = eta * kC * in->sdelta0 / (2 * kPi * workQz * f0);
    r->circ = in->circ;
    r->energy = E;
    r->rbeta = rBeta;
    r->rgamma = rGamma;
    r->t0 = t0;
    r->f0 = f0;
    r->harmonic = in->harmonic;
    r->u0 = in->u0;
    for (int i = 0; i < 3; i++) r->alphac[i] = in->alphac[i];
    r->nux = in->nux;
    r->nuy = in->nuy;
    r->chrom[0] = in->chrom[0];
    r->chrom[1] = in->chrom[1];
    for (int i = 0; i < 5; i++) { r->adtx[i] = in->adtx[i]; r->adty[i] = in->adty[i]; }
    r->skew_k = in->skew_k;
    for (int i = 0; i < 3; i++) r->synrad_tau[i] = in->synrad_tau[i];
    const double coupling = in->coupling;
    r->emit_nat[0] = in->natural_emit / (1 + coupling);
    r->emit_nat[1] = in->natural_emit / (1 + coupling) * coupling;
    r->emit_nat[2] = naturalBunchLength * in->sdelta0;
    // one ideal cavity: phase set so that Re(Vc) = U0
    r->n_rf = 1;
    const double phase = kPi - std::asin(in->u0 / vRF) - kPi / 2.;
    r->rf_vr[0] = vRF * std::cos(phase);
    r->rf_vi[0] = vRF * std::sin(phase);
    r->rf_harm[0] = 1;
    r->ap_x = in->ap_x;
    r->ap_y = in->ap_y;
    if (d) {
        d->eta = eta;
        d->workqz = workQz;
        d->natural_bunch_length = naturalBunchLength;
        d->beta_z = naturalBunchLength / in->sdelta0;
    }
    return CTS_OK;
}
