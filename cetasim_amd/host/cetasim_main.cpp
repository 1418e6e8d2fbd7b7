// cetasim_main.cpp -- the tracking program: main() of the reference
// (src/main.cpp) for calSetting 2, multi-particle tracking, on the MI355X
// library.
//
//   cetasim_run input.dat            track and write the output files
//   cetasim_run --check input.dat    read the input, build the lattice, the
//                                    filling pattern and the distributions
//                                    on the host only, print them as JSON
//   cetasim_run --dump DIR input.dat write what the run would hand the
//                                    library (ring, sections, bunches,
//                                    particles, settings) into DIR, host only
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "cetasim_model.hpp"

using namespace cetasim;

namespace {

void print_array(const char *name, const std::vector<double> &v, bool last = false)
{
    std::printf("\"%s\": [", name);
    for (size_t i = 0; i < v.size(); i++) std::printf("%s%.17g", i ? ", " : "", v[i]);
    std::printf("]%s\n", last ? "" : ",");
}

void print_iarray(const char *name, const std::vector<int> &v, bool last = false)
{
    std::printf("\"%s\": [", name);
    for (size_t i = 0; i < v.size(); i++) std::printf("%s%d", i ? ", " : "", v[i]);
    std::printf("]%s\n", last ? "" : ",");
}

// Everything the host set-up computed, for tests of the reader.
void print_check(ReadInputSettings &in, Train &train, LatticeInterActionPoint &lat, MPBeam &beam)
{
    const auto *B = in.ringParBasic;
    std::printf("{\n");
    std::printf("\"rGamma\": %.17g, \"rBeta\": %.17g, \"t0\": %.17g, \"f0\": %.17g, \"eta\": %.17g,\n", B->rGamma,
                B->rBeta, B->t0, B->f0, B->eta);
    std::printf("\"workQz\": %.17g, \"naturalBunchLength\": %.17g, \"harmonics\": %d,\n", B->workQz,
                B->naturalBunchLength, B->harmonics);
    std::printf("\"emittanceX\": %.17g, \"emittanceY\": %.17g, \"twissBetaZ\": %.17g,\n",
                in.ringBunchPara->emittanceX, in.ringBunchPara->emittanceY, lat.twissBetaZ);
    std::printf("\"nTurns\": %d, \"beamIonFlag\": %d, \"sRWakeFlag\": %d, \"synRadDampingFlag\": %d, \"seed\": %llu,\n",
                in.ringRun->nTurns, in.ringRun->beamIonFlag, in.ringRun->sRWakeFlag, in.ringRun->synRadDampingFlag,
                (unsigned long long)in.ringRun->seed);
    print_iarray("trainStart", train.trainStart);
    print_array("phaseAdvX12", lat.phaseAdvX12);
    print_array("phaseAdvY12", lat.phaseAdvY12);
    print_array("interactionLength", lat.interactionLength);
    print_array("vacuumPressure", lat.vacuumPressure);
    print_array("pipeAperatureX", lat.pipeAperatureX);
    print_array("pipeAperatureY", lat.pipeAperatureY);
    print_array("corssSectionEI", lat.corssSectionEI);
    std::vector<double> b1h1;
    for (auto &s : lat.sections) b1h1.insert(b1h1.end(), s.b1h1, s.b1h1 + 36);
    print_array("b1h1", b1h1);
    WakeFunction sr;
    if (sr.InitialSRWake(in) == 0) {
        print_array("lRs", sr.lRs);
        print_array("lQ", sr.lQ);
        print_array("lOmega", sr.lOmega);
    }
    std::vector<int> harm, gap, ncur;
    std::vector<double> cur, ne, q, moments;
    for (auto &b : beam.beamVec) {
        harm.push_back(b.bunchHarmNum);
        gap.push_back(b.bunchGap);
        ncur.push_back(b.normCurrent);
        cur.push_back(b.current);
        ne.push_back(b.electronNumPerBunch);
        q.push_back(b.macroEleCharge);
        // mean and rms of the six coordinates
        const std::vector<double> *c[6] = {&b.ePositionX, &b.eMomentumX, &b.ePositionY,
                                           &b.eMomentumY, &b.ePositionZ, &b.eMomentumZ};
        for (int k = 0; k < 6; k++) {
            double s = 0, s2 = 0;
            for (double v : *c[k]) s += v;
            const double m = s / (double)c[k]->size();
            for (double v : *c[k]) s2 += (v - m) * (v - m);
            moments.push_back(m);
            moments.push_back(std::sqrt(s2 / (double)c[k]->size()));
        }
    }
    print_iarray("bunchHarmNum", harm);
    print_iarray("bunchGap", gap);
    print_iarray("normCurrent", ncur);
    print_array("current", cur);
    print_array("electronNumPerBunch", ne);
    print_array("macroEleCharge", q);
    print_array("moments", moments, true);
    std::printf("}\n");
}

template <class T>
bool write_bin(const std::string &path, const T *p, size_t n)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, sizeof(T), n, f) == n;
    return std::fclose(f) == 0 && ok;
}

// The inputs of the library calls of MPBeam::Initial and MPBeam::Run as
// files: ring.bin (cts_ring), sections.bin (cts_section[nsec]),
// bunches.bin (int64 counts, double macro charge, double electrons,
// int32 gap, int32 harmonic number, each [nb]), particles.bin (double
// x px y py z dp, each [ntot]) and settings.json.
int dump(const std::string &dir, ReadInputSettings &in, LatticeInterActionPoint &lat, MPBeam &beam)
{
    cts_ring ring{};
    RunCalls c;
    if (beam.DeviceRing(lat, in, ring) || beam.RunSettings(lat, in, c)) return 1;
    const size_t nb = beam.beamVec.size();
    std::vector<int64_t> counts;
    std::vector<double> q, ne, X[6];
    std::vector<int32_t> gap, harm;
    for (auto &b : beam.beamVec) {
        counts.push_back(b.macroEleNumPerBunch);
        q.push_back(b.macroEleCharge);
        ne.push_back(b.electronNumPerBunch);
        gap.push_back(b.bunchGap);
        harm.push_back(b.bunchHarmNum);
        const std::vector<double> *src[6] = {&b.ePositionX, &b.eMomentumX, &b.ePositionY,
                                             &b.eMomentumY, &b.ePositionZ, &b.eMomentumZ};
        for (int k = 0; k < 6; k++) X[k].insert(X[k].end(), src[k]->begin(), src[k]->end());
    }
    bool ok = write_bin(dir + "/ring.bin", &ring, 1) &&
              write_bin(dir + "/sections.bin", lat.sections.data(), lat.sections.size());
    FILE *f = std::fopen((dir + "/bunches.bin").c_str(), "wb");
    ok = ok && f && std::fwrite(counts.data(), 8, nb, f) == nb && std::fwrite(q.data(), 8, nb, f) == nb &&
         std::fwrite(ne.data(), 8, nb, f) == nb && std::fwrite(gap.data(), 4, nb, f) == nb &&
         std::fwrite(harm.data(), 4, nb, f) == nb;
    if (f) ok = (std::fclose(f) == 0) && ok;
    f = std::fopen((dir + "/particles.bin").c_str(), "wb");
    for (int k = 0; k < 6 && f; k++) ok = ok && std::fwrite(X[k].data(), 8, X[k].size(), f) == X[k].size();
    if (f) ok = (std::fclose(f) == 0) && ok;
    f = std::fopen((dir + "/settings.json").c_str(), "w");
    if (f) {
        auto arr = [&](const char *name, const double *v, size_t n) {
            std::fprintf(f, "\"%s\": [", name);
            for (size_t i = 0; i < n; i++) std::fprintf(f, "%s%.17g", i ? ", " : "", v[i]);
            std::fprintf(f, "],\n");
        };
        const auto *R = in.ringRun;
        std::fprintf(f, "{\n\"nb\": %zu, \"nsec\": %zu, \"ntot\": %zu,\n", nb, lat.sections.size(), X[0].size());
        std::fprintf(f, "\"nTurns\": %d, \"interval\": %d, \"fit_start\": %d, \"fit_end\": %d,\n", R->nTurns,
                     R->bunchInfoPrintInterval, R->growthRateFittingStart, R->growthRateFittingEnd);
        std::fprintf(f, "\"skew\": %d, \"wake\": %d, \"nbins\": %d,\n", in.ringParBasic->skewQuadK != 0, c.wake,
                     c.nbins);
        arr("rs", c.rs.data(), c.rs.size());
        arr("q", c.q.data(), c.q.size());
        arr("fr", c.fr.data(), c.fr.size());
        std::fprintf(f, "\"synrad\": %d, \"seed\": %llu, \"ions\": %d,\n", c.synrad, (unsigned long long)c.seed,
                     c.ions);
        // interaction point 0 under the plain keys, every point in ion_ips
        const cts_ion_config ic = c.ion.empty() ? cts_ion_config{} : c.ion[0];
        std::fprintf(f, "\"ion_ips\": [");
        for (size_t k = 0; k < c.ion.size(); k++) {
            const cts_ion_config &e = c.ion[k];
            std::fprintf(f, "%s{\"pressure\": [%.17g, %.17g, %.17g, %.17g], \"temperature\": %.17g, "
                            "\"length\": %.17g, \"ap\": [%.17g, %.17g], \"seed\": %llu}",
                         k ? ", " : "", e.pressure[0], e.pressure[1], e.pressure[2], e.pressure[3], e.temperature,
                         e.length, e.ap_x, e.ap_y, (unsigned long long)e.seed);
        }
        std::fprintf(f, "],\n");
        std::fprintf(f, "\"ion_nspec\": %d, \"ion_n_new\": %d, \"ion_max\": %lld,\n", ic.nspec, ic.n_new,
                     (long long)ic.max_ions);
        arr("ion_mass", ic.mass_number, CTS_MAX_SPECIES);
        arr("ion_cross", ic.cross_section, CTS_MAX_SPECIES);
        arr("ion_pressure", ic.pressure, CTS_MAX_SPECIES);
        std::fprintf(f, "\"ion_temperature\": %.17g, \"ion_length\": %.17g, \"ion_loss\": %.17g,\n",
                     ic.temperature, ic.length, ic.loss_boundary);
        std::fprintf(f, "\"ion_ap\": [%.17g, %.17g], \"ion_seed\": %llu,\n", ic.ap_x, ic.ap_y,
                     (unsigned long long)ic.seed);
        std::fprintf(f, "\"lrw\": %d, \"lrw_nturns\": %d, \"lrw_beta\": [%.17g, %.17g],\n", c.lrw, c.lrw_nturns,
                     c.lrw_beta[0], c.lrw_beta[1]);
        arr("lrw_sectors", c.lrw_sectors.data(), c.lrw_sectors.size());
        arr("lrw_res", c.lrw_res.data(), c.lrw_res.size());
        std::fprintf(f, "\"bbi\": %d, \"bbi_fmax\": %.17g, \"bbi_td\": %d, \"bbi_flags\": [%d, %d, %d, %d, %d],\n",
                     c.bbi, c.bbi_fmax, c.bbi_time_domain, c.bbi_flags[0], c.bbi_flags[1], c.bbi_flags[2],
                     c.bbi_flags[3], c.bbi_flags[4]);
        arr("bbi_re", c.bbi_re.data(), c.bbi_re.size());
        arr("bbi_im", c.bbi_im.data(), c.bbi_im.size());
        arr("bbi_qx_re", c.bbi_qx_re.data(), c.bbi_qx_re.size());
        arr("bbi_qx_im", c.bbi_qx_im.data(), c.bbi_qx_im.size());
        arr("bbi_qy_re", c.bbi_qy_re.data(), c.bbi_qy_re.size());
        arr("bbi_qy_im", c.bbi_qy_im.data(), c.bbi_qy_im.size());
        arr("bbi_dx_re", c.bbi_dx_re.data(), c.bbi_dx_re.size());
        arr("bbi_dx_im", c.bbi_dx_im.data(), c.bbi_dx_im.size());
        arr("bbi_dy_re", c.bbi_dy_re.data(), c.bbi_dy_re.size());
        arr("bbi_dy_im", c.bbi_dy_im.data(), c.bbi_dy_im.size());
        arr("bbi_beta", c.bbi_beta, 2);
        const char *wk[6] = {"bbi_wpos", "bbi_wz", "bbi_wdx", "bbi_wdy", "bbi_wqx", "bbi_wqy"};
        for (int i = 0; i < 6; i++) arr(wk[i], c.bbi_wake[i].data(), c.bbi_wake[i].size());
        const cts_beam_loading_config &b = c.blc;
        std::fprintf(f, "\"bl\": %d, \"bl_nres\": %d, \"bl_nbins\": %d, \"bl_nturns\": %d,\n", c.bl, b.nres,
                     b.nbins, b.nturns);
        std::fprintf(f, "\"bl_int\": [");
        for (int i = 0; i < b.nres; i++)
            std::fprintf(f, "%s[%d, %d, %d, %d, %d, %d]", i ? ", " : "", b.rf_mode[i], b.res_type[i], b.res_harm[i],
                         b.cold[i], b.excite_instability[i], b.dir_fb[i]);
        std::fprintf(f, "],\n\"bl_real\": [");
        for (int i = 0; i < b.nres; i++)
            std::fprintf(f, "%s[%.17g, %.17g, %.17g, %.17g, %.17g, %.17g, %.17g, %.17g, %.17g]", i ? ", " : "",
                         b.rs[i], b.q0[i], b.coupling[i], b.detune_hz[i], b.vol[i], b.phase[i], b.dir_fb_gain[i],
                         b.dir_fb_phase[i], b.dir_fb_delay[i]);
        std::fprintf(f, "],\n");
        std::fprintf(f, "\"tbt_print\": %d, \"ring_current\": %.17g,\n", R->TBTBunchPrintNum,
                     in.ringParBasic->ringCurrent);
        std::fprintf(f, "\"workQz\": %.17g, \"betaZ\": %.17g, \"tbt\": \"%s\", \"cbmgr\": \"%s\"\n}\n",
                     in.ringParBasic->workQz, in.ringBunchPara->rmsBunchLength / in.ringBunchPara->rmsEnergySpread,
                     R->TBTBunchAverData.c_str(), R->runCBMGR.c_str());
        ok = (std::fclose(f) == 0) && ok;
    } else {
        ok = false;
    }
    if (!ok) {
        std::cerr << "cannot write the dump into " << dir << std::endl;
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char *argv[])
{
    bool check_only = false, no_result = false;
    std::string dump_dir;
    int a = 1;
    if (argc > 1 && std::strcmp(argv[1], "--no-result") == 0) {
        // skip result.sdds: the turns then run as HIP-graph replays
        no_result = true;
        a = 2;
    } else if (argc > 1 && std::strcmp(argv[1], "--check") == 0) {
        check_only = true;
        a = 2;
    } else if (argc > 2 && std::strcmp(argv[1], "--dump") == 0) {
        dump_dir = argv[2];
        a = 3;
    }
    char *args[2] = {argv[0], a < argc ? argv[a] : nullptr};
    ReadInputSettings inputParameter;
    if (inputParameter.ParamRead(args[1] ? 2 : 1, args)) return 1;
    if (inputParameter.ringRun->calSetting != 2 || inputParameter.ringBunchPara->macroEleNumPerBunch == 1) {
        std::cerr << "wrong settings about SP and MP tracking: this program runs calSetting 2 "
                     "(multi-particle) with bunchMacroEleNumPerBunch > 1" << std::endl;
        return 1;
    }
    LatticeInterActionPoint latticeInterActionPoint;
    if (latticeInterActionPoint.Initial(inputParameter)) return 1;
    Train train;
    if (train.Initial(inputParameter)) return 1;
    MPBeam mpbeam;
    if (check_only || !dump_dir.empty()) {
        if (mpbeam.InitialBunches(train, latticeInterActionPoint, inputParameter)) return 1;
        if (!dump_dir.empty()) return dump(dump_dir, inputParameter, latticeInterActionPoint, mpbeam);
        print_check(inputParameter, train, latticeInterActionPoint, mpbeam);
        return 0;
    }
    mpbeam.writeResult = !no_result;
    if (mpbeam.Initial(train, latticeInterActionPoint, inputParameter)) return 1;
    if (mpbeam.Run(train, latticeInterActionPoint, inputParameter)) return 1;
    std::cout << "End of Tracking " << inputParameter.ringRun->nTurns << " Turns" << std::endl;
    return 0;
}
