// cts_bl.cpp -- host model of the beam-loaded (rfMode 1) cavities:
// CavityResonator::Initial (src/CavityResonator.cpp:36-176), the initial
// beam-induced and generator voltages of MPBeam::InitialcavityResonator
// (src/MPBeam.cpp:236-465), the generator current of
// Resonator::GetInitialResonatorGenIg (src/Resonator.cpp:38-57), and the
// cavity's evolution along the train in MPBeam::BeamMomtumUpdateDueToRFTest
// (src/MPBeam.cpp:2624-2706): at each bunch the voltage it sees, its rigid
// self-induced voltage V_b0, then the generator's dynamics bucket by bucket
// (Resonator::ResonatorDynamics, src/Resonator.cpp:235-261) and the decay and
// rotation of the beam-induced voltage to the next bunch
// (Resonator::GetBeamInducedVol, src/Resonator.cpp:205-232).  In the
// reference's stable variant none of this depends on the particles, so the
// voltages of every bunch are tabulated here for the run's turns and the
// device only applies the kick.  With the bin-by-bin variant or the direct
// feedback the state depends on the particles: beam_loading_walk_state
// sets it up for the device walk (k_bl_walk).
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "cts_internal.hpp"

namespace cts {
namespace {

using cplx = std::complex<double>;
constexpr double kPiB = 3.14159265358979323846;
const cplx kI(0.0, 1.0);

struct Cavity {
    int mode, type, harm, cold;
    double rs, q0, beta, detune;
    double ql, fre, psi, tf;
    cplx req, vb, gen, ig;
};

// CavityResonator::Initial and MPBeam::InitialcavityResonator: the cavities'
// constants, the induced voltage after 100 filling times, the generator
// voltage and current
std::vector<Cavity> setup_cavities(const cts_beam_loading_config &c, const cts_ring &ring, const std::vector<double> &ne,
                                   const std::vector<int32_t> &gap)
{
    const int nres = c.nres, nb = (int)ne.size(), H = ring.harmonic;
    const double f0 = ring.f0;
    std::vector<Cavity> cv(nres);
    // CavityResonator::Initial
    for (int i = 0; i < nres; i++) {
        Cavity &r = cv[i];
        r.mode = c.rf_mode[i];
        r.type = c.res_type[i];
        r.harm = c.res_harm[i];
        r.cold = c.cold[i];
        r.rs = c.rs[i];
        r.q0 = c.q0[i];
        r.beta = c.coupling[i];
        r.detune = c.detune_hz[i];
        r.ql = r.q0 / (1 + r.beta);
        r.fre = r.harm * H * f0 + r.detune;
        const double tan_psi = r.ql * (r.fre / (r.harm * H * f0) - r.harm * H * f0 / r.fre);
        r.psi = std::atan(tan_psi);
        r.tf = 2 * r.ql / (2 * kPiB * r.fre);
    }
    for (int i = 0; i < nres; i++) {
        double ph = c.phase[i];
        // one cavity: the synchronous phase that restores U0
        if (nres == 1) ph = kPiB - std::asin(ring.u0 / c.vol[0]) - kPiB / 2.;
        cv[i].req = cplx(c.vol[i] * std::cos(ph), c.vol[i] * std::sin(ph));
    }
    // bunch harmonic numbers from the gaps (a common offset leaves the
    // steady state below unchanged)
    std::vector<int> harm(nb, 0);
    for (int b = 1; b < nb; b++) harm[b] = harm[b - 1] + gap[b - 1];
    // MPBeam::InitialcavityResonator: beam-induced voltage after 100 filling
    // times, averaged over the last turn's buckets, and the mean self-loss
    std::vector<cplx> kick_aver(nres);
    for (int i = 0; i < nres; i++) {
        Cavity &r = cv[i];
        const int nturns = (int)std::ceil(100 * r.tf * f0);
        cplx vb(0, 0), aver2(0, 0), kick_accum(0, 0);
        const double tB = 1.0 / f0 / H;
        const double dL = tB / r.tf, cpsi = dL * std::tan(r.psi);
        for (int n = 0; n < nturns; n++) {
            int k = 0;
            for (int j = 0; j < H; j++) {
                cplx vb0(0, 0);
                if (k < nb && harm[k] == j) {
                    vb0 = cplx(-1 * r.fre * 2 * kPiB * r.rs / r.q0, 0.E0) * ne[k] * kEcharge;
                    k++;
                    if (n == nturns - 1) kick_accum += vb0 / 2.0;
                }
                vb += vb0;
                vb = vb * std::exp(-dL) * std::exp(kI * cpsi);
                if (n == nturns - 1) aver2 += vb;
            }
        }
        r.vb = aver2 / double(H);
        kick_aver[i] = kick_accum / double(nb);
    }
    for (int i = 0; i < nres; i++) {
        Cavity &r = cv[i];
        r.gen = r.type == 0 ? cplx(0, 0) : r.req - r.vb - kick_aver[i];
        if (r.cold || r.mode == 0) r.vb = cplx(0, 0);
        if (r.type == 0) {
            r.ig = cplx(0, 0);
        } else {
            const double arg_vgr = std::arg(r.gen) - r.psi;
            const double abs_vgr = std::abs(r.gen) / std::abs(std::cos(r.psi));
            r.ig = abs_vgr * std::exp(kI * arg_vgr) * (1.0 + r.beta) / r.rs;
        }
    }
    return cv;
}

}  // namespace

int beam_loading_tables(const cts_beam_loading_config &c, const cts_ring &ring, const std::vector<double> &ne,
                        const std::vector<int32_t> &gap, BlTable &t, std::string &err)
{
    const int nres = c.nres, nb = (int)ne.size(), H = ring.harmonic;
    const double f0 = ring.f0;
    std::vector<Cavity> cv = setup_cavities(c, ring, ne, gap);
    // the beam-loaded cavities, in input order
    t = BlTable{};
    t.nres = nres;
    t.nturns = c.nturns;
    t.nbins = c.nbins;
    for (int i = 0; i < nres; i++)
        if (cv[i].mode == 1) {
            t.bl_index.push_back(i);
            t.kfac.push_back(2. * kPiB * (double)cv[i].harm * (f0 * H) / kClight / ring.rbeta);
        }
    t.nbl = (int)t.bl_index.size();
    if (t.nbl == 0) {
        err = "cts_set_beam_loading: no rfMode-1 cavity";
        return CTS_ERR_ARG;
    }
    const size_t per_turn = (size_t)nb * t.nbl * 3;
    t.dev.assign((size_t)c.nturns * per_turn, 0.0);
    t.info.assign((size_t)c.nturns * nb * nres * 8, 0.0);
    const double trf = ring.t0 / H;
    for (int n = 0; n < c.nturns; n++) {
        for (int q = 0; q < t.nbl; q++) {
            Cavity &r = cv[t.bl_index[q]];
            const double sigma = 2.0 * kPiB * r.fre / (2.0 * r.ql);
            const double dom = 2.0 * kPiB * r.detune;
            const double es = std::exp(-sigma * trf), sn = std::sin(dom * trf), cs = std::cos(dom * trf);
            const double alpha = dom * es * sn - sigma * es * cs + sigma;
            const double beta = sigma * es * sn + dom * es * cs - dom;
            const double coefB = 2 * kPiB * r.fre / 2.0 * r.rs / r.q0 / (sigma * sigma + dom * dom);
            const cplx gen_step = std::exp(-sigma * trf) * std::exp(kI * dom * trf);
            const cplx gen_add(coefB * (alpha * r.ig.real() + beta * r.ig.imag()),
                               coefB * (-beta * r.ig.real() + alpha * r.ig.imag()));
            for (int b = 0; b < nb; b++) {
                // BunchMomentumUpdateDueToRFModeStable
                const cplx vb0 = cplx(-1 * 2 * kPiB * r.fre * r.rs / r.q0, 0.E0) * ne[b] * kEcharge;
                const cplx gen = r.gen, cav = r.vb + gen;
                double *d = t.dev.data() + (size_t)n * per_turn + ((size_t)b * t.nbl + q) * 3;
                d[0] = cav.real();
                d[1] = cav.imag();
                d[2] = vb0.real();
                double *f = t.info.data() + (((size_t)n * nb + b) * nres + t.bl_index[q]) * 8;
                const cplx ind = r.vb + vb0, self = vb0 / 2.0;
                f[0] = cav.real(); f[1] = cav.imag();
                f[2] = ind.real(); f[3] = ind.imag();
                f[4] = gen.real(); f[5] = gen.imag();
                f[6] = self.real(); f[7] = self.imag();
                r.vb += vb0;
                // the generator over the bunch's gap, bucket by bucket
                for (int k = 0; k < gap[b]; k++) {
                    r.gen *= gen_step;
                    r.gen += gen_add;
                }
                // beam-induced voltage decayed and rotated to the next bunch
                const double T = gap[b] * trf;
                r.vb *= std::exp(-T / r.tf) * std::exp(kI * (2.0 * kPiB * r.fre * T));
            }
        }
    }
    return CTS_OK;
}

int beam_loading_walk_state(const cts_beam_loading_config &c, const cts_ring &ring, const std::vector<double> &ne,
                            const std::vector<int32_t> &gap, BlTable &t, std::string &err)
{
    const int nres = c.nres, H = ring.harmonic;
    const double f0 = ring.f0;
    std::vector<Cavity> cv = setup_cavities(c, ring, ne, gap);
    t = BlTable{};
    t.walk = true;
    t.nres = nres;
    t.nturns = c.nturns;
    t.nbins = c.nbins;
    for (int i = 0; i < nres; i++) {
        if (cv[i].mode != 1) continue;
        const Cavity &r = cv[i];
        t.bl_index.push_back(i);
        t.kfac.push_back(2. * kPiB * (double)r.harm * (f0 * H) / kClight / ring.rbeta);
        double s[kBlS] = {};
        s[BS_EXC] = c.excite_instability[i] != 0;
        s[BS_DFB] = c.dir_fb[i] != 0;
        s[BS_HARM] = r.harm;
        s[BS_RS] = r.rs;
        s[BS_Q0] = r.q0;
        s[BS_QL] = r.ql;
        s[BS_DET] = r.detune;
        s[BS_FRE] = r.fre;
        s[BS_TF] = r.tf;
        s[BS_REQ] = r.req.real(); s[BS_REQ + 1] = r.req.imag();
        s[BS_VB] = r.vb.real(); s[BS_VB + 1] = r.vb.imag();
        s[BS_GEN] = r.gen.real(); s[BS_GEN + 1] = r.gen.imag();
        s[BS_IG] = r.ig.real(); s[BS_IG + 1] = r.ig.imag();
        s[BS_FBG] = c.dir_fb_gain[i];
        s[BS_FBP] = c.dir_fb_phase[i];
        s[BS_FBD] = c.dir_fb_delay[i];
        t.state.insert(t.state.end(), s, s + kBlS);
    }
    t.nbl = (int)t.bl_index.size();
    if (t.nbl == 0) {
        err = "cts_set_beam_loading: no rfMode-1 cavity";
        return CTS_ERR_ARG;
    }
    return CTS_OK;
}

}  // namespace cts
