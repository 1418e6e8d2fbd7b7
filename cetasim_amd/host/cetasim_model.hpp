// cetasim_model.hpp -- host classes of the tracking path with the names,
// members and call sequence of the reference's ReadInputSettings, Train,
// LatticeInterActionPoint, WakeFunction, MPBunch and MPBeam
// (include/ReadInputSettings.h, include/Train.h,
// include/LatticeInterActionPoint.h, include/WakeFunction.h,
// include/MPBunch.h, include/MPBeam.h), reduced to what the multi-particle
// tracking turn reads.  Every seam calls the MI355X library through its C
// ABI (include/cetasim_hip.h); the particles live on the GPU and the
// MPBunch vectors are filled only when asked (DownloadParticles).
#pragma once

#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/cetasim_hip.h"
#include "../../include/cetasim_host.h"

namespace cetasim {

// input.dat reader (src/ReadInputSettings.cpp:46-1082): keys are case
// insensitive, blanks inside a key are ignored, the last occurrence wins,
// values are the whitespace-separated tokens after '='.
class ReadInputSettings {
public:
    struct RingParBasic {
        double circRing = 0, workQx = 0, workQy = 0, workQz = 0, chrom[2] = {0, 0};
        double synchRadDampTime[3] = {0, 0, 0}, skewQuadK = 0, pipeAperature[3] = {1, 1, 1};
        double electronBeamEnergy = 0, alphac[3] = {0, 0, 0}, aDTX[5] = {}, aDTY[5] = {};
        double u0 = 0, sdelta0 = 0, ringCurrent = 0, naturalEmit = 0, couplingfactorXY = 0;
        std::string twissInput;
        int ringSectNum = 1;
        // derived (src/ReadInputSettings.cpp:986-1077)
        double rGamma = 0, rBeta = 0, t0 = 0, f0 = 0, eta = 0, naturalBunchLength = 0, emitNat[3] = {0, 0, 0};
        int harmonics = 0;
    };
    struct RingParRf {
        int ringHarm = 0, resNum = 0;
        std::vector<int> resHarm, rfMode;
        std::vector<double> resVol, resPhase;
        // beam-loaded (rfMode 1) cavities (src/ReadInputSettings.cpp:194-368)
        std::vector<int> resType, resCold, resExciteIntability, resDirFB;
        std::vector<double> resDirFBGain, resDirFBPhase, resDirFBDelay;
        std::vector<double> resShuntImpRs, resQualityQ0, resCouplingBeta, resDetuneFre;
        int rfBunchBinNum = 0;
        std::string transResonParWriteTo;
    };
    struct RingFillPatt {
        int totBunchNumber = 1, trainNumber = 1;
        std::vector<int> bunchNumberPerTrain{1}, trainGaps{0}, bunchGaps{0};
        int bunchChargeNum = 0;
        std::vector<int> bunchChargeIndex, bunchCharge;
    };
    struct RingBunchPara {
        double current = 0, emittanceX = 0, emittanceY = 0, kappa = 0, rmsBunchLength = 0, rmsEnergySpread = 0;
        int macroEleNumPerBunch = 1, distributionType = 3;
        double initialStaticOffSet[6] = {0, 0, 0, 0, 0, 0};   // x y z px py dp
        double initialDynamicOffSet[6] = {0, 0, 0, 0, 0, 0};  // rms of a per-bunch random offset
    };
    struct RingIonEffPara {
        int gasSpec = 0, macroIonNumberGeneratedPerIP = 0, numberofIonBeamInterPoint = 1;
        std::vector<double> ionMassNumber, gasPercent, corssSectionEI;
        double ionLossBoundary = 0, ionMaxNumber = 0;
        std::string twissInput;
    };
    struct RingSRWake {
        std::string SRWBBRInput;
        int SRWBunchBinNum = 100;
    };
    struct RingLRWake {
        std::string pipeGeoInput, bbrInput, lrwOutput;
        int nTurnswakeTrunction = 0;  // lrwNTurnsWakeTrunction + 1, as the reference stores it
    };
    struct RingBBImp {  // &BoardBandImpedance (src/ReadInputSettings.cpp:726-752)
        std::string impedInput, wakeInput;
        // timeDomain 1 (tracking in the time domain) unless the input says otherwise
        // (include/ReadInputSettings.h:213)
        int bunchBinNumberZ = 0, impedSimFlag[5] = {0, 0, 0, 0, 0}, timeDomain = 1;
        double quasiWakeBunchLen = 0;
    };
    struct RingRun {
        int bBImpFlag = 0;
        int calSetting = 2, nTurns = 0, bunchInfoPrintInterval = 1, synRadDampingFlag = 0, beamIonFlag = 0;
        int sRWakeFlag = 0, lRWakeFlag = 0, growthRateFittingStart = 0, growthRateFittingEnd = 0;
        std::string TBTBunchAverData = "TBTBunchAverData", runCBMGR = "CBMGR";
        int TBTBunchPrintNum = 0;  // bunches with their own columns in result.sdds
        uint64_t seed = 1;  // seed of the injected distribution, ions and noise (Q13)
    };

    RingParBasic *ringParBasic = &basic_;
    RingParRf *ringParRf = &rf_;
    RingFillPatt *ringFillPatt = &fill_;
    RingBunchPara *ringBunchPara = &bunch_;
    RingIonEffPara *ringIonEffPara = &ion_;
    RingSRWake *ringSRWake = &srw_;
    RingLRWake *ringLRWake = &lrw_;
    RingBBImp *ringBBImp = &bbi_;
    RingRun *ringRun = &run_;

    // Returns 0; prints the reason and returns a non-zero code where the
    // reference exits.
    int ParamRead(int argc, char *argv[]);
    int ParamReadFile(const std::string &path);

private:
    RingParBasic basic_;
    RingParRf rf_;
    RingFillPatt fill_;
    RingBunchPara bunch_;
    RingIonEffPara ion_;
    RingSRWake srw_;
    RingLRWake lrw_;
    RingBBImp bbi_;
    RingRun run_;
    void Derive();
};

// Train::Initial (src/Train.cpp:38-100), including the reference's use of
// bunchGaps[i] in the train-start sum (Q10).
class Train {
public:
    int trainNumber = 0, totBunchNum = 0;
    std::vector<int> bunchNumberPerTrain, trainGaps, bunchGaps, trainStart;
    int Initial(ReadInputSettings &inputParameter);
};

// Twiss of the interaction points and the section maps
// (LatticeInterActionPoint::Initial / InitialLattice /
// InitialLatticeSympMat, src/LatticeInterActionPoint.cpp:39-431).
class LatticeInterActionPoint {
public:
    int numberOfInteraction = 1, gasSpec = 0;
    std::vector<double> elemPos, twissBetaX, twissAlphaX, twissBetaY, twissAlphaY, twissDispX, twissDispPX;
    std::vector<double> twissDispY, twissDispPY, xPhaseAdv, yPhaseAdv, phaseAdvX12, phaseAdvY12;
    std::vector<double> pipeAperatureX, pipeAperatureY, interactionLength, vacuumPressure, temperature;
    std::vector<double> ionMassNumber, corssSectionEI, gasPercent;
    double ionLossBoundary = 0, ionMaxNumberOneInterPoint = 0, twissBetaZ = 0;
    std::vector<cts_section> sections;  // the three maps and twiss of each section
    int Initial(const ReadInputSettings &inputParameter);
};

// Broadband resonators of the short-range wake
// (WakeFunction::BBRWakeParaReadIn, src/WakeFunction.cpp:251-290): 12
// header lines, 9 columns (longitudinal Rs Q f, then transverse x and y).
class WakeFunction {
public:
    std::vector<double> lRs, lQ, lOmega, txRs, txQ, txOmega, tyRs, tyQ, tyOmega;
    // resistive-wall pipe sectors (RWWakeParaReadIn, src/WakeFunction.cpp:
    // 292-320): one header line, 7 columns
    std::vector<double> sectorRadiusX, sectorRadiusY, sectorLength, sectorBetaX, sectorBetaY, sectorNum,
        sectormatSigma;
    double betaFunIntPoint[2] = {0, 0};
    int RWWakeParaReadIn(const std::string &filename);
    // WakeFunction::InitialLRWake (src/WakeFunction.cpp:49-90): the pipe
    // and resonator files of the long-range wake, beta at the first twiss point
    int InitialLRWake(const ReadInputSettings &inputParameter, const LatticeInterActionPoint &lattice);
    // WakeFunction::InitialSRWake (src/WakeFunction.cpp:232-249): reads the
    // resonator file named by srwBBRInput when there is one.
    int InitialSRWake(const ReadInputSettings &inputParameter);
    int BBRWakeParaReadIn(const std::string &inputfilename);
};

// One bunch (include/MPBunch.h, include/Bunch.h): the host mirror of the
// device state and the moments of the last MPBeamRMSCal.
class MPBunch {
public:
    int macroEleNumPerBunch = 0, bunchHarmNum = 0, bunchGap = 0, normCurrent = 1;
    double current = 0, electronNumPerBunch = 0, macroEleCharge = 0;
    std::vector<double> ePositionX, ePositionY, ePositionZ, eMomentumX, eMomentumY, eMomentumZ;
    std::vector<int> eSurive;
    double xAver = 0, pxAver = 0, yAver = 0, pyAver = 0, zAver = 0, pzAver = 0;
    double emittance[3] = {0, 0, 0}, effectiveEmittance[3] = {0, 0, 0};
    double rmsRx = 0, rmsRy = 0, rmsEffectiveRx = 0, rmsEffectiveRy = 0, rmsBunchLength = 0, rmsEnergySpread = 0;
    double transmission = 1;
};

// The library set-up calls MPBeam::Run makes before its turn loop.
struct RunCalls {
    bool wake = false;
    int nbins = 0;
    std::vector<double> rs, q, fr;  // resonators: Rs [Ohm], Q, f [Hz]
    bool synrad = false;
    uint64_t seed = 0;
    bool ions = false;
    std::vector<cts_ion_config> ion;  // one per interaction point (section)
    bool lrw = false;
    int lrw_nturns = 0;
    std::vector<double> lrw_sectors, lrw_res;  // [n][7], [n][9] (cts_lrwake_config's layout)
    double lrw_beta[2] = {0, 0};
    std::string lrw_output;
    bool bbi = false;  // broadband impedance (BoardBandImp::ReadInImp's table)
    int bbi_flags[5] = {0, 0, 0, 0, 0}, bbi_time_domain = 0;
    double bbi_fmax = 0;
    std::vector<double> bbi_re, bbi_im, bbi_qx_re, bbi_qx_im, bbi_qy_re, bbi_qy_im, bbi_dx_re, bbi_dx_im, bbi_dy_re,
        bbi_dy_im;
    double bbi_beta[2] = {1, 1};
    std::vector<double> bbi_wake[6];  // time domain: binPosZ, wz, wDx, wDy, wQx, wQy (GetQuasiWakePoten)
    bool bl = false;  // rfMode-1 cavities present
    cts_beam_loading_config blc{};
};

// The beam of bunches and the per-turn seams of MPBeam::Run
// (src/MPBeam.cpp:1010-1173), each a call into the library.
class MPBeam {
public:
    std::vector<MPBunch> beamVec;
    cts_handle *gpu = nullptr;
    ~MPBeam();
    // MPBeam::Initial (src/MPBeam.cpp:64-161): bunch harmonic numbers,
    // currents, charges, distributions and gaps, then the device set-up
    int Initial(Train &train, LatticeInterActionPoint &lattice, ReadInputSettings &inputParameter);
    // the host half of Initial alone (no device)
    int InitialBunches(Train &train, LatticeInterActionPoint &lattice, ReadInputSettings &inputParameter);
    // the library's ring constants from the input (ReadInputSettings'
    // derived block and CavityResonator::Initial)
    int DeviceRing(const LatticeInterActionPoint &lattice, const ReadInputSettings &inputParameter, cts_ring &ring);
    // wake, radiation and ion settings of Run
    int RunSettings(const LatticeInterActionPoint &lattice, const ReadInputSettings &inputParameter, RunCalls &calls);
    // the wake function is built inside, as the reference's Run does
    // (src/MPBeam.cpp:881-890)
    int Run(Train &train, LatticeInterActionPoint &lattice, ReadInputSettings &inputParameter);
    // seams (src/MPBeam.cpp:2613, 1210, 2624, 1204, 2708, 1219, 2117, 1358,
    // 2126, 2585)
    int BeamTransferPerInteractionPointDueToLatticeT(int k);
    int BeamTransferDueToSkewQuad();
    int BeamMomtumUpdateDueToRFTest();
    int BeamLongiPosTransferOneTurn();
    int BeamEnergyLossOneTurn();
    int BeamSynRadDamping(int turn);
    int SRWakeBeamIntaction();
    int MarkParticleLostInBunch();
    int MPBeamRMSCal(int k);
    int SSBeamIonEffectOneInteractionPoint(int turn, int k);
    int DownloadParticles();
    // growth rates of the coupled-bunch modes from the run's samples
    // (MPBeam::GetCBMGR, src/MPBeam.cpp:1496-1891): ideal modes, IQ,
    // Hilbert modes and per-bunch rates, written to runCBMGR.sdds;
    // rates[3][nb] gets the ideal-mode rates
    int GetCBMGR(ReadInputSettings &inputParameter, std::vector<double> &rates);
    void PrintTBTBunchAverData(std::ofstream &fout, int turn, const cts_moments *m, const double *ionCharge,
                               const double *cavInfo,
                               const ReadInputSettings &inputParameter, bool header);
    // result.sdds (src/MPBeam.cpp:901-1159): the header, then one row per
    // turn from the end-of-turn moments, eigen-emittances and ion charge
    void PrintResultHeader(std::ofstream &fout, const ReadInputSettings &inputParameter);
    void PrintResultRow(std::ofstream &fout, int turn, const cts_moments *m, const double *eigen, double ionCharge,
                        const ReadInputSettings &inputParameter);
    bool writeResult = true;  // result.sdds every turn (one eigen-emittance pass per turn)
    // the long-range wake table file of InitialLRWake (src/WakeFunction.cpp:
    // 100-155) from the library's table
    int PrintLRWakeTable(const std::string &path);

private:
    int check(int rc, const char *what);
    std::vector<cts_moments> samples_;  // [sample][bunch], turns 0, interval, ...
    double alphaX0_ = 0, betaX0_ = 1, alphaY0_ = 0, betaY0_ = 1;
    std::vector<double> cavity_;  // resCavVolReq of each cavity (re, im)
};

// Initial distribution of one bunch (MPBunch::DistriGenerator,
// src/MPBunch.cpp:98-298): longitudinal Gaussian truncated at 3 sigma,
// transverse action-angle sampling at interaction point 0 for
// distributionType 1 (shell), 2 (uniform) and 3 (Gaussian, cut at 3 sigma),
// then the per-bunch offsets and the dispersion.  The draws come from a
// splitmix64 stream keyed by runSeed and the bunch index instead of the
// time-seeded rand() and random_device-seeded mt19937, so a run repeats.
int DistriGenerator(MPBunch &bunch, const LatticeInterActionPoint &lattice, const ReadInputSettings &inputParameter,
                    int bunchIndex);

}  // namespace cetasim
