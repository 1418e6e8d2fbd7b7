// cetasim_model.cpp -- host classes of the tracking path over the MI355X
// library's C ABI (see cetasim_model.hpp).
#include "cetasim_model.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace cetasim {

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kClight = 299792458.0;
constexpr double kEcharge = 1.602017733e-19;

// key (blanks removed, lower case) and the tokens after '='
// (StringVecSplit, src/Global.cpp:210-260)
bool split_line(const std::string &line, std::string &key, std::vector<std::string> &vals)
{
    const size_t eq = line.find('=');
    if (line.empty() || eq == std::string::npos) return false;
    key.clear();
    for (size_t i = 0; i < eq; i++)
        if (!std::isblank((unsigned char)line[i])) key.push_back((char)std::tolower((unsigned char)line[i]));
    vals.clear();
    std::istringstream rest(line.substr(eq + 1));
    std::string tok;
    while (rest >> tok) vals.push_back(tok);
    return true;
}

double num(const std::vector<std::string> &v, size_t i)
{
    if (i >= v.size()) throw std::runtime_error("missing value");
    return std::stod(v[i]);  // parses the leading number, as the reference's stod
}

int inum(const std::vector<std::string> &v, size_t i)
{
    if (i >= v.size()) throw std::runtime_error("missing value");
    return std::stoi(v[i]);
}

uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace

// ------------------------------------------------------------- input.dat
int ReadInputSettings::ParamRead(int argc, char *argv[])
{
    return ParamReadFile(argc >= 2 ? argv[1] : "input.dat");
}

int ReadInputSettings::ParamReadFile(const std::string &path)
{
    std::ifstream fin(path);
    if (!fin.is_open()) {
        std::cerr << "Error opening file " << path << std::endl;
        return 1;
    }
    std::string line, k;
    std::vector<std::string> v;
    int lineno = 0;
    while (std::getline(fin, line)) {
        lineno++;
        if (!split_line(line, k, v)) continue;
        try {
            RingParBasic &B = basic_;
            if (k == "ringcircring") B.circRing = num(v, 0);
            else if (k == "ringworkq") { B.workQx = num(v, 0); B.workQy = num(v, 1); B.workQz = num(v, 2); }
            else if (k == "ringchrom") { B.chrom[0] = num(v, 0); B.chrom[1] = num(v, 1); }
            else if (k == "ringsynchraddamptime") for (int i = 0; i < 3; i++) B.synchRadDampTime[i] = num(v, i);
            else if (k == "ringskewquadk") B.skewQuadK = num(v, 0);
            else if (k == "ringpipeaperaturex") B.pipeAperature[1] = num(v, 0);
            else if (k == "ringpipeaperaturey") B.pipeAperature[2] = num(v, 0);
            else if (k == "ringpipeaperaturer") B.pipeAperature[0] = num(v, 0);
            else if (k == "ringelectronbeamenergy") B.electronBeamEnergy = num(v, 0);
            else if (k == "ringalphac") for (int i = 0; i < 3; i++) B.alphac[i] = num(v, i);
            else if (k == "ringadtx") for (int i = 0; i < 5; i++) B.aDTX[i] = num(v, i);  // Q8: five values
            else if (k == "ringadty") for (int i = 0; i < 5; i++) B.aDTY[i] = num(v, i);
            else if (k == "ringu0") B.u0 = num(v, 0);
            else if (k == "ringsdelta0") B.sdelta0 = num(v, 0);
            else if (k == "ringcurrent") B.ringCurrent = num(v, 0);
            else if (k == "ringnaturalemit") B.naturalEmit = num(v, 0);
            else if (k == "ringcoupling") B.couplingfactorXY = num(v, 0);
            else if (k == "ringtwiss") B.twissInput = v.at(0);
            else if (k == "ringsectnum") B.ringSectNum = inum(v, 0);
            // rf (sized by rfresnum, which must come first)
            else if (k == "rfringharm") rf_.ringHarm = inum(v, 0);
            else if (k == "rfresnum") {
                rf_.resNum = inum(v, 0);
                rf_.resHarm.assign(rf_.resNum, 1);
                rf_.rfMode.assign(rf_.resNum, 0);
                rf_.resVol.assign(rf_.resNum, 0.0);
                rf_.resPhase.assign(rf_.resNum, 0.0);
                rf_.resType.assign(rf_.resNum, 0);
                rf_.resCold.assign(rf_.resNum, 0);
                rf_.resExciteIntability.assign(rf_.resNum, 0);
                rf_.resDirFB.assign(rf_.resNum, 0);
                rf_.resDirFBGain.assign(rf_.resNum, 0.0);
                rf_.resDirFBPhase.assign(rf_.resNum, 0.0);
                rf_.resDirFBDelay.assign(rf_.resNum, 0.0);
                rf_.resShuntImpRs.assign(rf_.resNum, 0.0);
                rf_.resQualityQ0.assign(rf_.resNum, 0.0);
                rf_.resCouplingBeta.assign(rf_.resNum, 0.0);
                rf_.resDetuneFre.assign(rf_.resNum, 0.0);
            } else if (k == "rfresharm") for (int i = 0; i < rf_.resNum; i++) rf_.resHarm[i] = inum(v, i);
            else if (k == "rfmode") for (int i = 0; i < rf_.resNum; i++) rf_.rfMode[i] = inum(v, i);
            else if (k == "rfresvol") for (int i = 0; i < rf_.resNum; i++) rf_.resVol[i] = num(v, i);
            else if (k == "rfresphase") for (int i = 0; i < rf_.resNum; i++) rf_.resPhase[i] = num(v, i);
            else if (k == "rfrestype") for (int i = 0; i < rf_.resNum; i++) rf_.resType[i] = inum(v, i);
            else if (k == "rfresshuntimprs") for (int i = 0; i < rf_.resNum; i++) rf_.resShuntImpRs[i] = num(v, i);
            else if (k == "rfresqualityq0") for (int i = 0; i < rf_.resNum; i++) rf_.resQualityQ0[i] = num(v, i);
            else if (k == "rfrescouplingbeta") for (int i = 0; i < rf_.resNum; i++) rf_.resCouplingBeta[i] = num(v, i);
            else if (k == "rfresdetunefre") for (int i = 0; i < rf_.resNum; i++) rf_.resDetuneFre[i] = num(v, i);
            else if (k == "rfrescold") for (int i = 0; i < rf_.resNum; i++) rf_.resCold[i] = inum(v, i);
            else if (k == "rfresexciteintability")
                for (int i = 0; i < rf_.resNum; i++) rf_.resExciteIntability[i] = (int)num(v, i);
            else if (k == "rfresdirfb") for (int i = 0; i < rf_.resNum; i++) rf_.resDirFB[i] = (int)num(v, i);
            else if (k == "rfresdirfbgain") for (int i = 0; i < rf_.resNum; i++) rf_.resDirFBGain[i] = num(v, i);
            else if (k == "rfresdirfbphaseshift") for (int i = 0; i < rf_.resNum; i++) rf_.resDirFBPhase[i] = num(v, i);
            else if (k == "rfresdirfbdelay") for (int i = 0; i < rf_.resNum; i++) rf_.resDirFBDelay[i] = num(v, i);
            else if (k == "rfbunchbinnum") rf_.rfBunchBinNum = (int)num(v, 0);
            else if (k == "rftransientcavparwriteto") rf_.transResonParWriteTo = v.at(0);
            // filling pattern (sized by fptrainnumber)
            else if (k == "fptotbunchnumber") fill_.totBunchNumber = inum(v, 0);
            else if (k == "fptrainnumber") {
                fill_.trainNumber = inum(v, 0);
                fill_.bunchNumberPerTrain.assign(fill_.trainNumber, 0);
                fill_.trainGaps.assign(fill_.trainNumber, 0);
                fill_.bunchGaps.assign(fill_.trainNumber, 0);
            } else if (k == "fpbunchnumberpertrain") for (int i = 0; i < fill_.trainNumber; i++) fill_.bunchNumberPerTrain[i] = inum(v, i);
            else if (k == "fptraingaps") for (int i = 0; i < fill_.trainNumber; i++) fill_.trainGaps[i] = inum(v, i);
            else if (k == "fpbunchgaps") for (int i = 0; i < fill_.trainNumber; i++) fill_.bunchGaps[i] = inum(v, i);
            else if (k == "fpsetbunchchargenum") {
                fill_.bunchChargeNum = inum(v, 0);
                fill_.bunchChargeIndex.assign(fill_.bunchChargeNum, 0);
                fill_.bunchCharge.assign(fill_.bunchChargeNum, 1);
            } else if (k == "fpsetbunchchargeindex") for (int i = 0; i < fill_.bunchChargeNum; i++) fill_.bunchChargeIndex[i] = inum(v, i);
            else if (k == "fpsetbunchcharge") for (int i = 0; i < fill_.bunchChargeNum; i++) fill_.bunchCharge[i] = inum(v, i);
            // bunch
            else if (k == "bunchcurrent") bunch_.current = num(v, 0);
            else if (k == "bunchemittance") bunch_.emittanceX = num(v, 0);
            else if (k == "bunchkappa") bunch_.kappa = num(v, 0);
            else if (k == "bunchrmsbunchlength") bunch_.rmsBunchLength = num(v, 0);
            else if (k == "bunchrmsenergyspread") bunch_.rmsEnergySpread = num(v, 0);
            else if (k == "bunchdistributiontype") bunch_.distributionType = (int)num(v, 0);
            else if (k == "bunchmacroelenumperbunch") bunch_.macroEleNumPerBunch = inum(v, 0);
            else if (k == "bunchinitialstaticoffset") for (int i = 0; i < 6; i++) bunch_.initialStaticOffSet[i] = num(v, i);
            else if (k == "bunchinitialdynamicoffset") for (int i = 0; i < 6; i++) bunch_.initialDynamicOffSet[i] = num(v, i);
            // ions (sized by ioncalgasspec)
            else if (k == "ioncalgasspec") {
                ion_.gasSpec = inum(v, 0);
                ion_.ionMassNumber.assign(ion_.gasSpec, 0.0);
                ion_.gasPercent.assign(ion_.gasSpec, 0.0);
                ion_.corssSectionEI.assign(ion_.gasSpec, 0.0);
            } else if (k == "ioncalionmassnumber") for (int i = 0; i < ion_.gasSpec; i++) ion_.ionMassNumber[i] = num(v, i);
            else if (k == "ioncalcorsssectionei") for (int i = 0; i < ion_.gasSpec; i++) ion_.corssSectionEI[i] = num(v, i) * 1.e-22;
            else if (k == "ioncaliongaspercent") for (int i = 0; i < ion_.gasSpec; i++) ion_.gasPercent[i] = num(v, i);
            else if (k == "ioncalionmaxnumber") ion_.ionMaxNumber = num(v, 0);
            else if (k == "ioncalionlossboundary") ion_.ionLossBoundary = num(v, 0);
            else if (k == "ioncalmacroionnumbergeneratedperip") ion_.macroIonNumberGeneratedPerIP = (int)num(v, 0);
            // long-range wake (src/ReadInputSettings.cpp:688-705)
            else if (k == "lrwpipegeoinput") lrw_.pipeGeoInput = v.at(0);
            else if (k == "lrwoutput") lrw_.lrwOutput = v.at(0);
            else if (k == "lrwbbrinput") lrw_.bbrInput = v.at(0);
            else if (k == "lrwnturnswaketrunction") lrw_.nTurnswakeTrunction = inum(v, 0) + 1;
            // short-range wake
            else if (k == "srwbbrinput") srw_.SRWBBRInput = v.at(0);
            else if (k == "srwbunchbinnum") srw_.SRWBunchBinNum = (int)num(v, 0);
            // run
            else if (k == "runcalsetting") run_.calSetting = inum(v, 0);
            else if (k == "runnturns") run_.nTurns = inum(v, 0);
            else if (k == "runbunchinfoprintinterval") run_.bunchInfoPrintInterval = inum(v, 0);
            else if (k == "runsynraddampingflag") run_.synRadDampingFlag = inum(v, 0);
            else if (k == "runbeamionflag") run_.beamIonFlag = inum(v, 0);
            else if (k == "runshortrangewake") run_.sRWakeFlag = inum(v, 0);
            else if (k == "runlongrangewake") run_.lRWakeFlag = inum(v, 0);
            else if (k == "runbbiflag") run_.bBImpFlag = inum(v, 0);
            else if (k == "bbiimpedinput") bbi_.impedInput = v.at(0);
            else if (k == "bbibunchbinnumberz") bbi_.bunchBinNumberZ = inum(v, 0);
            else if (k == "bbiimpedsimlandtflag") for (int i = 0; i < 5; i++) bbi_.impedSimFlag[i] = inum(v, i);
            else if (k == "bbiimpedsimquasigreenbunchlength") bbi_.quasiWakeBunchLen = num(v, 0);
            else if (k == "bbiimpedsimtimeorfreflag") bbi_.timeDomain = inum(v, 0);
            else if (k == "bbiwakeinput") bbi_.wakeInput = v.at(0);
            else if (k == "rungrowthratefittingstart") run_.growthRateFittingStart = inum(v, 0);
            else if (k == "rungrowthratefittingend") run_.growthRateFittingEnd = inum(v, 0);
            else if (k == "runtbtbunchaverdata") run_.TBTBunchAverData = v.at(0);
            else if (k == "runtbtbunchprintnum") run_.TBTBunchPrintNum = inum(v, 0);
            else if (k == "runcbmgr") run_.runCBMGR = v.at(0);
            else if (k == "runseed") run_.seed = (uint64_t)std::stoull(v.at(0));
        } catch (const std::exception &) {
            std::cerr << path << ":" << lineno << ": bad value for " << k << std::endl;
            return 1;
        }
    }
    // the growth-rate window is checked even when unused
    // (src/ReadInputSettings.cpp:946-956)
    if (run_.nTurns > 0 && !(run_.growthRateFittingStart < run_.growthRateFittingEnd &&
                             run_.growthRateFittingEnd <= run_.nTurns)) {
        std::cerr << "growth rate fitting window has to satisfy start < end <= nTurns" << std::endl;
        return 1;
    }
    Derive();
    return 0;
}

// src/ReadInputSettings.cpp:986-1077 through cts_derive_ring
void ReadInputSettings::Derive()
{
    ion_.numberofIonBeamInterPoint = basic_.ringSectNum;
    ion_.twissInput = basic_.twissInput;
    bunch_.emittanceY = bunch_.emittanceX / (1 + bunch_.kappa) * bunch_.kappa;
    bunch_.emittanceX = bunch_.emittanceX / (1 + bunch_.kappa);
    cts_ring_input in{};
    in.circ = basic_.circRing;
    in.energy = basic_.electronBeamEnergy;
    in.u0 = basic_.u0;
    in.sdelta0 = basic_.sdelta0;
    for (int i = 0; i < 3; i++) in.alphac[i] = basic_.alphac[i];
    in.nux = basic_.workQx;
    in.nuy = basic_.workQy;
    in.chrom[0] = basic_.chrom[0];
    in.chrom[1] = basic_.chrom[1];
    for (int i = 0; i < 5; i++) { in.adtx[i] = basic_.aDTX[i]; in.adty[i] = basic_.aDTY[i]; }
    in.skew_k = basic_.skewQuadK;
    for (int i = 0; i < 3; i++) in.synrad_tau[i] = basic_.synchRadDampTime[i];
    in.natural_emit = basic_.naturalEmit;
    in.coupling = basic_.couplingfactorXY;
    in.harmonic = rf_.ringHarm;
    in.rf_volt = rf_.resVol.empty() ? 0.0 : rf_.resVol[0];
    in.ap_x = basic_.pipeAperature[1];
    in.ap_y = basic_.pipeAperature[2];
    cts_ring ring{};
    cts_ring_derived der{};
    if (cts_derive_ring(&in, &ring, &der) == CTS_OK) {
        basic_.rGamma = ring.rgamma;
        basic_.rBeta = ring.rbeta;
        basic_.t0 = ring.t0;
        basic_.f0 = ring.f0;
        basic_.harmonics = ring.harmonic;
        basic_.eta = der.eta;
        basic_.workQz = der.workqz;  // the input value is replaced (Q9)
        basic_.naturalBunchLength = der.natural_bunch_length;
        for (int i = 0; i < 3; i++) basic_.emitNat[i] = ring.emit_nat[i];
    }
}

// ----------------------------------------------------------------- train
int Train::Initial(ReadInputSettings &in)
{
    const int harmonics = in.ringParBasic->harmonics;
    trainNumber = in.ringFillPatt->trainNumber;
    totBunchNum = in.ringFillPatt->totBunchNumber;
    bunchNumberPerTrain = in.ringFillPatt->bunchNumberPerTrain;
    trainGaps = in.ringFillPatt->trainGaps;
    bunchGaps = in.ringFillPatt->bunchGaps;
    int nCheck = 0, hCheck = 0;
    for (int i = 0; i < trainNumber; i++) {
        nCheck += bunchNumberPerTrain[i];
        hCheck += bunchNumberPerTrain[i] * (bunchGaps[i] + 1) + trainGaps[i];
    }
    if (nCheck != totBunchNum || hCheck != harmonics) {
        std::cerr << "mistakes in settings for bunch filling pattern, sum of gaps + totBunchNumber has to be harmonics"
                  << std::endl;
        return 1;
    }
    trainStart.assign(trainNumber, 0);
    for (int i = 1; i < trainNumber; i++)
        for (int j = 0; j < i; j++) trainStart[i] += bunchNumberPerTrain[j] * (bunchGaps[i] + 1) + trainGaps[j];
    return 0;
}

// --------------------------------------------------------------- lattice
int LatticeInterActionPoint::Initial(const ReadInputSettings &in)
{
    const auto *B = in.ringParBasic;
    numberOfInteraction = in.ringIonEffPara->numberofIonBeamInterPoint;
    gasSpec = in.ringIonEffPara->gasSpec;
    ionMassNumber = in.ringIonEffPara->ionMassNumber;
    corssSectionEI = in.ringIonEffPara->corssSectionEI;
    gasPercent = in.ringIonEffPara->gasPercent;
    ionLossBoundary = in.ringIonEffPara->ionLossBoundary;
    ionMaxNumberOneInterPoint = in.ringIonEffPara->ionMaxNumber;
    if (gasSpec > 0) {
        double tot = 0;
        for (double p : gasPercent) tot += p;
        if (tot != 1.e0) {  // src/LatticeInterActionPoint.cpp:56-65
            std::cerr << "non 100 percent gas" << std::endl;
            return 1;
        }
    }
    std::ifstream fin(B->twissInput);
    if (!fin.is_open()) {
        std::cerr << "Error opening file " << B->twissInput << std::endl;
        return 1;
    }
    std::string str;
    int index = 0;
    twissBetaZ = B->naturalBunchLength / B->sdelta0;
    while (std::getline(fin, str)) {
        if (index++ <= 4) continue;  // five header lines
        std::istringstream ls(str);
        std::vector<double> c;
        std::string tok;
        while (ls >> tok) c.push_back(std::stod(tok));
        if (c.empty()) continue;
        if (c.size() < 16) {
            std::cerr << B->twissInput << ": 16 columns expected" << std::endl;
            return 1;
        }
        elemPos.push_back(c[0]);
        twissBetaX.push_back(c[1]); twissAlphaX.push_back(c[2]); xPhaseAdv.push_back(c[3]);
        twissDispX.push_back(c[4]); twissDispPX.push_back(c[5]);
        pipeAperatureX.push_back(std::min(c[6], B->pipeAperature[1]));
        twissBetaY.push_back(c[7]); twissAlphaY.push_back(c[8]); yPhaseAdv.push_back(c[9]);
        twissDispY.push_back(c[10]); twissDispPY.push_back(c[11]);
        pipeAperatureY.push_back(std::min(c[12], B->pipeAperature[2]));
        interactionLength.push_back(c[13]);
        vacuumPressure.push_back(c[14] * 1.0E-9 * 133.3224);  // nTorr -> Pa
        temperature.push_back(c[15]);
    }
    const int n = (int)elemPos.size();
    if (n != numberOfInteraction) {
        std::cerr << "data of rintSecNum in input.dat and InterPointParameter.dat files does not match" << std::endl;
        return 1;
    }
    phaseAdvX12.assign(n, 0.0);
    phaseAdvY12.assign(n, 0.0);
    for (int i = 0; i < n; i++) {
        if (i != n - 1) {
            interactionLength[i] = elemPos[i + 1] - elemPos[i];
            phaseAdvX12[i] = xPhaseAdv[i + 1] - xPhaseAdv[i];
            phaseAdvY12[i] = yPhaseAdv[i + 1] - yPhaseAdv[i];
        } else {
            interactionLength[i] = B->circRing - elemPos[i] + elemPos[0];
            phaseAdvX12[i] = 2 * kPi * B->workQx - xPhaseAdv[i] + xPhaseAdv[0];
            phaseAdvY12[i] = 2 * kPi * B->workQy - yPhaseAdv[i] + yPhaseAdv[0];
        }
    }
    sections.assign(n, cts_section{});
    for (int k = 0; k < n; k++) {
        const int k2 = (k + 1) % n;
        cts_twiss t1{twissBetaX[k], twissAlphaX[k], twissDispX[k], twissDispPX[k],
                     twissBetaY[k], twissAlphaY[k], twissDispY[k], twissDispPY[k], twissBetaZ};
        cts_twiss t2{twissBetaX[k2], twissAlphaX[k2], twissDispX[k2], twissDispPX[k2],
                     twissBetaY[k2], twissAlphaY[k2], twissDispY[k2], twissDispPY[k2], twissBetaZ};
        if (cts_build_section(&t1, &t2, phaseAdvX12[k], phaseAdvY12[k], &sections[k]) != CTS_OK) return 1;
    }
    return 0;
}

// ------------------------------------------------------------------ wake
int WakeFunction::BBRWakeParaReadIn(const std::string &path)
{
    std::ifstream fin(path);
    if (!fin.is_open()) {
        std::cerr << "Error opening file " << path << std::endl;
        return 1;
    }
    std::string str;
    int index = 0;
    while (std::getline(fin, str)) {
        if (index++ < 12) continue;  // twelve header lines
        std::istringstream ls(str);
        std::vector<double> c;
        std::string tok;
        while (ls >> tok) c.push_back(std::stod(tok));
        if (c.empty()) continue;
        if (c.size() < 9) {
            std::cerr << path << ": 9 columns expected (Q11)" << std::endl;
            return 1;
        }
        lRs.push_back(c[0]);
        lQ.push_back(c[1]);
        lOmega.push_back(c[2] * 2 * kPi);
        txRs.push_back(c[3]);
        txQ.push_back(c[4]);
        txOmega.push_back(c[5] * 2 * kPi);
        tyRs.push_back(c[6]);
        tyQ.push_back(c[7]);
        tyOmega.push_back(c[8] * 2 * kPi);
    }
    return 0;
}

int WakeFunction::InitialSRWake(const ReadInputSettings &in)
{
    if (in.ringSRWake->SRWBBRInput.empty()) return 0;
    return BBRWakeParaReadIn(in.ringSRWake->SRWBBRInput);
}

int WakeFunction::RWWakeParaReadIn(const std::string &path)
{
    std::ifstream fin(path);
    if (!fin.is_open()) {
        std::cerr << "Error opening file: " << path << std::endl;
        return 1;
    }
    std::string str;
    std::getline(fin, str);  // one header line
    std::vector<double> *col[7] = {&sectorRadiusX, &sectorRadiusY, &sectorLength, &sectorBetaX,
                                   &sectorBetaY,   &sectorNum,     &sectormatSigma};
    while (std::getline(fin, str)) {
        std::istringstream ls(str);
        std::vector<double> c;
        std::string tok;
        while (ls >> tok) c.push_back(std::stod(tok));
        if (c.empty()) continue;
        if (c.size() < 7) {
            std::cerr << path << ": 7 columns expected" << std::endl;
            return 1;
        }
        for (int k = 0; k < 7; k++) col[k]->push_back(c[k]);
    }
    return 0;
}

int WakeFunction::InitialLRWake(const ReadInputSettings &in, const LatticeInterActionPoint &lat)
{
    betaFunIntPoint[0] = lat.twissBetaX[0];
    betaFunIntPoint[1] = lat.twissBetaY[0];
    const auto *L = in.ringLRWake;
    if (!L->pipeGeoInput.empty() && RWWakeParaReadIn(L->pipeGeoInput)) return 1;
    if (!L->bbrInput.empty() && BBRWakeParaReadIn(L->bbrInput)) return 1;
    return 0;
}

// ---------------------------------------------------------- distribution
namespace {

// The time-domain broadband impedance's wake file, as
// MPBeam::GetQuasiWakePoten reads it (src/MPBeam.cpp:748-812): five header
// lines, then rows z, wz, wDx, wDy, wQx, wQy; every wake times -1 (the
// reference's sign convention), positions moved to start at 0, and the
// longitudinal wake folded about the row whose z is exactly 0.
int ReadQuasiWake(const std::string &path, std::vector<double> (&out)[6])
{
    std::ifstream f(path);
    if (!f) {
        std::cerr << "cannot open the wake file " << path << std::endl;
        return 1;
    }
    for (auto &v : out) v.clear();
    std::string line;
    int cen = 0;
    for (int ln = 0; std::getline(f, line); ln++) {
        if (ln < 5) continue;
        std::istringstream ss(line);
        std::vector<double> col;
        double x;
        while (ss >> x) col.push_back(x);
        if (col.empty()) continue;
        if (col.size() < 6) {
            std::cerr << path << ": row " << ln + 1 << " has " << col.size() << " columns, 6 needed" << std::endl;
            return 1;
        }
        if (col[0] == 0) cen = (int)out[0].size();
        out[0].push_back(col[0]);
        for (int c = 1; c < 6; c++) out[c].push_back(-1 * col[c]);
    }
    const int n = (int)out[0].size();
    if (n < 2 || 2 * cen > n) {
        std::cerr << path << ": at least two rows, and the z = 0 row in the first half of the table" << std::endl;
        return 1;
    }
    const double z0 = out[0][0];
    for (double &z : out[0]) z -= z0;
    const std::vector<double> t = out[1];
    std::vector<double> &wz = out[1];
    wz[0] = t[cen];
    for (int i = 1; i < cen; i++) wz[i] = t[cen + i] + t[cen - i];
    for (int i = cen; i < n - cen; i++) wz[i] = t[cen + i];
    for (int i = n - cen; i < n; i++) wz[i] = 0.0;
    return 0;
}

// Uniform [0, 1) and standard normal draws from a splitmix64 counter stream.
struct SeededStream {
    uint64_t key, ctr = 0;
    explicit SeededStream(uint64_t k) : key(k) {}
    double uniform() { return (double)(splitmix64(key + ctr++) >> 11) * 0x1p-53; }
    double normal()
    {
        const double u1 = 1.0 - uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * kPi * u2);
    }
};

// Root of 1 - f - (1 + a) e^{-a} on [0, 20] by bisection: the action of a
// Gaussian beam with cumulative fraction f (MPBunch::GSSlover,
// src/MPBunch.cpp:300-338: same bracket, tolerances and early exit).
double gaussian_action(double f)
{
    double lo = 0.0, hi = 20.0;
    while (hi - lo > 1.0e-5) {
        const double mid = 0.5 * (lo + hi);
        const double g = 1 - f - (1 + mid) * std::exp(-mid);
        if (std::abs(g) < 1.0e-6) return mid;
        if (g < 0) lo = mid;
        else if (g > 0) hi = mid;
    }
    return 0.5 * (lo + hi);
}

}  // namespace

int DistriGenerator(MPBunch &b, const LatticeInterActionPoint &lat, const ReadInputSettings &in, int bunchIndex)
{
    const auto *P = in.ringBunchPara;
    const int n = b.macroEleNumPerBunch;
    const double rBeta = in.ringParBasic->rBeta;
    b.ePositionX.assign(n, 0.0); b.eMomentumX.assign(n, 0.0); b.ePositionY.assign(n, 0.0);
    b.eMomentumY.assign(n, 0.0); b.ePositionZ.assign(n, 0.0); b.eMomentumZ.assign(n, 0.0);
    b.eSurive.assign(n, 0);
    SeededStream rng(splitmix64(in.ringRun->seed ^ ((uint64_t)(bunchIndex + 1) << 40)));
    double off[6];
    for (int i = 0; i < 6; i++) off[i] = rng.normal() * P->initialDynamicOffSet[i] + P->initialStaticOffSet[i];
    const double sz = P->rmsBunchLength, sd = P->rmsEnergySpread;
    // longitudinal: Gaussian truncated at 3 sigma in (z, dE/E)
    for (int i = 0; i < n;) {
        const double z = rng.normal() * sz, e = rng.normal() * sd;
        if ((z / sz) * (z / sz) + (e / sd) * (e / sd) < 9.0) {
            b.ePositionZ[i] = z;
            b.eMomentumZ[i] = e / (rBeta * rBeta);
            i++;
        }
    }
    // transverse: action-angle sampling at interaction point 0
    const double ax = lat.twissAlphaX[0], ay = lat.twissAlphaY[0], bx = lat.twissBetaX[0], by = lat.twissBetaY[0];
    const double sgx = std::sqrt(bx), sgy = std::sqrt(by), dsx = -ax / std::sqrt(bx), dsy = -ay / std::sqrt(by);
    const double ex = P->emittanceX, rX = std::sqrt(P->emittanceX * bx), rY = std::sqrt(P->emittanceY * by);
    const int type = P->distributionType;
    if (type < 1 || type > 3) {
        std::cerr << "no distribution type " << type << std::endl;
        return 1;
    }
    for (int i = 0; i < n;) {
        double fi;
        const double u = rng.uniform();
        if (type == 1) fi = 4 * ex;
        else if (type == 2) fi = 6 * ex * std::sqrt(u);
        else fi = gaussian_action(u) * 2 * ex;
        const double axax = fi * rng.uniform();
        const double ayay = (fi - axax) * P->kappa;
        const double a_x = std::sqrt(axax), a_y = std::sqrt(ayay);
        const double phx = 2 * kPi * rng.uniform(), phy = 2 * kPi * rng.uniform();
        const double x = a_x * sgx * std::cos(phx), y = a_y * sgy * std::cos(phy);
        if (type == 3 && (x / rX) * (x / rX) + (y / rY) * (y / rY) > 9) continue;
        b.ePositionX[i] = x;
        b.ePositionY[i] = y;
        b.eMomentumX[i] = a_x * (dsx * std::cos(phx) - std::sin(phx) / sgx);
        b.eMomentumY[i] = a_y * (dsy * std::cos(phy) - std::sin(phy) / sgy);
        i++;
    }
    // offsets and dispersion
    for (int i = 0; i < n; i++) {
        b.ePositionZ[i] += off[2];
        b.eMomentumZ[i] += off[5];
        b.ePositionX[i] += off[0] + lat.twissDispX[0] * b.eMomentumZ[i];
        b.ePositionY[i] += off[1] + lat.twissDispY[0] * b.eMomentumZ[i];
        b.eMomentumX[i] += off[3] + lat.twissDispPX[0] * b.eMomentumZ[i];
        b.eMomentumY[i] += off[4] + lat.twissDispPY[0] * b.eMomentumZ[i];
    }
    return 0;
}

// ------------------------------------------------------------------ beam
MPBeam::~MPBeam()
{
    if (gpu) cts_destroy(gpu);
}

int MPBeam::check(int rc, const char *what)
{
    if (rc == CTS_OK) return 0;
    if (rc == CTS_LOW_TRANSMISSION)
        std::cerr << "more than 50% particles are lost in a bunch (" << what << ")" << std::endl;
    else if (rc == CTS_ION_OVERFLOW)
        std::cerr << "the accumulated ions are larger than the limit; reduce the ions generated per interaction "
                     "or enlarge ionCalIonMaxNumber (" << what << ")" << std::endl;
    else
        std::cerr << what << ": " << cts_last_error() << std::endl;
    return rc;
}

int MPBeam::InitialBunches(Train &train, LatticeInterActionPoint &lat, ReadInputSettings &in)
{
    const auto *B = in.ringParBasic;
    const int nb = in.ringFillPatt->totBunchNumber;
    beamVec.assign(nb, MPBunch{});
    int counter = 0;
    for (int i = 0; i < train.trainNumber; i++)
        for (int j = 0; j < train.bunchNumberPerTrain[i]; j++)
            beamVec[counter++].bunchHarmNum = train.trainStart[i] + j * (train.bunchGaps[i] + 1);
    for (int i = 0; i < in.ringFillPatt->bunchChargeNum; i++) {
        const int idx = in.ringFillPatt->bunchChargeIndex[i];
        if (idx < 0 || idx >= nb) {
            std::cerr << "fpSetBunchChargeIndex " << idx << " is not a bunch index" << std::endl;
            return 1;
        }
        beamVec[idx].normCurrent = in.ringFillPatt->bunchCharge[i];
    }
    int totNorm = 0;
    for (auto &b : beamVec) totNorm += b.normCurrent;
    for (int i = 0; i < nb; i++) {
        MPBunch &b = beamVec[i];
        b.current = b.normCurrent * B->ringCurrent / totNorm;
        b.macroEleNumPerBunch = in.ringBunchPara->macroEleNumPerBunch;
        b.electronNumPerBunch = b.current / B->f0 / kEcharge;  // src/Bunch.cpp:69-70
        b.macroEleCharge = b.electronNumPerBunch / b.macroEleNumPerBunch;
        b.rmsBunchLength = in.ringBunchPara->rmsBunchLength;
        b.rmsEnergySpread = in.ringBunchPara->rmsEnergySpread;
        if (DistriGenerator(b, lat, in, i)) return 1;
    }
    for (int i = 0; i < nb - 1; i++) beamVec[i].bunchGap = beamVec[i + 1].bunchHarmNum - beamVec[i].bunchHarmNum;
    beamVec[nb - 1].bunchGap = B->harmonics - beamVec[nb - 1].bunchHarmNum;
    return 0;
}

int MPBeam::DeviceRing(const LatticeInterActionPoint &lat, const ReadInputSettings &in, cts_ring &ring)
{
    const auto *B = in.ringParBasic;
    cts_ring_input ri{};
    ri.circ = B->circRing; ri.energy = B->electronBeamEnergy; ri.u0 = B->u0; ri.sdelta0 = B->sdelta0;
    for (int i = 0; i < 3; i++) ri.alphac[i] = B->alphac[i];
    ri.nux = B->workQx; ri.nuy = B->workQy; ri.chrom[0] = B->chrom[0]; ri.chrom[1] = B->chrom[1];
    for (int i = 0; i < 5; i++) { ri.adtx[i] = B->aDTX[i]; ri.adty[i] = B->aDTY[i]; }
    ri.skew_k = B->skewQuadK;
    for (int i = 0; i < 3; i++) ri.synrad_tau[i] = B->synchRadDampTime[i];
    ri.natural_emit = B->naturalEmit; ri.coupling = B->couplingfactorXY;
    ri.harmonic = B->harmonics;
    ri.rf_volt = in.ringParRf->resVol.empty() ? 0.0 : in.ringParRf->resVol[0];
    ri.ap_x = lat.pipeAperatureX[0]; ri.ap_y = lat.pipeAperatureY[0];
    cts_ring_derived der{};
    if (check(cts_derive_ring(&ri, &ring, &der), "cts_derive_ring")) return 1;
    // cavities (CavityResonator::Initial, src/CavityResonator.cpp:60-139):
    // one cavity takes the phase that restores U0, several keep their
    // input phases.  The ring lists the ideal (rfMode 0) cavities; the
    // beam-loaded ones (rfMode 1) are set up by cts_set_beam_loading.
    const auto *RF = in.ringParRf;
    if (RF->resNum > CTS_MAX_RF) {
        std::cerr << "at most " << CTS_MAX_RF << " cavities" << std::endl;
        return 1;
    }
    for (int i = 0; i < RF->resNum; i++)
        if (RF->rfMode[i] != 0 && RF->rfMode[i] != 1) {
            std::cerr << "cavity " << i << ": rfMode " << RF->rfMode[i] << " (0 ideal, 1 beam loading)" << std::endl;
            return 1;
        }
    if (RF->resNum == 1 && RF->rfMode[0] == 1) ring.n_rf = 0;
    if (RF->resNum > 1) {
        ring.n_rf = 0;
        for (int i = 0; i < RF->resNum; i++) {
            if (RF->rfMode[i] != 0) continue;
            ring.rf_vr[ring.n_rf] = RF->resVol[i] * std::cos(RF->resPhase[i]);
            ring.rf_vi[ring.n_rf] = RF->resVol[i] * std::sin(RF->resPhase[i]);
            ring.rf_harm[ring.n_rf] = RF->resHarm[i];
            ring.n_rf++;
        }
    }
    return 0;
}

int MPBeam::Initial(Train &train, LatticeInterActionPoint &lat, ReadInputSettings &in)
{
    const int nb = in.ringFillPatt->totBunchNumber;
    if (InitialBunches(train, lat, in)) return 1;

    cts_ring ring{};
    if (DeviceRing(lat, in, ring)) return 1;
    // the ideal cavities' phasors in input order (zeros for beam-loaded ones)
    const auto *RF = in.ringParRf;
    cavity_.assign(2 * (size_t)RF->resNum, 0.0);
    for (int j = 0, k = 0; j < RF->resNum; j++)
        if (RF->rfMode[j] == 0 && k < ring.n_rf) {
            cavity_[2 * j] = ring.rf_vr[k];
            cavity_[2 * j + 1] = ring.rf_vi[k];
            k++;
        }
    if (check(cts_create(&ring, lat.sections.data(), (int32_t)lat.sections.size(), &gpu), "cts_create")) return 1;
    std::vector<int64_t> counts(nb);
    std::vector<double> q(nb), ne(nb);
    std::vector<int32_t> gap(nb);
    for (int i = 0; i < nb; i++) {
        counts[i] = beamVec[i].macroEleNumPerBunch;
        q[i] = beamVec[i].macroEleCharge;
        ne[i] = beamVec[i].electronNumPerBunch;
        gap[i] = beamVec[i].bunchGap;
    }
    if (check(cts_set_bunches(gpu, nb, counts.data(), q.data(), ne.data(), gap.data()), "cts_set_bunches")) return 1;
    std::vector<double> X[6];
    for (auto &b : beamVec) {
        const std::vector<double> *src[6] = {&b.ePositionX, &b.eMomentumX, &b.ePositionY,
                                             &b.eMomentumY, &b.ePositionZ, &b.eMomentumZ};
        for (int c = 0; c < 6; c++) X[c].insert(X[c].end(), src[c]->begin(), src[c]->end());
    }
    return check(cts_upload(gpu, X[0].data(), X[1].data(), X[2].data(), X[3].data(), X[4].data(), X[5].data(),
                            nullptr), "cts_upload");
}

int MPBeam::BeamTransferPerInteractionPointDueToLatticeT(int k) { return check(cts_section_map(gpu, k), "lattice"); }
int MPBeam::BeamTransferDueToSkewQuad() { return check(cts_skew_quad(gpu), "skew quad"); }
int MPBeam::BeamMomtumUpdateDueToRFTest() { return check(cts_rf_kick(gpu), "rf"); }
int MPBeam::BeamLongiPosTransferOneTurn() { return check(cts_drift(gpu), "drift"); }
int MPBeam::BeamEnergyLossOneTurn() { return check(cts_energy_loss(gpu), "energy loss"); }
int MPBeam::BeamSynRadDamping(int turn) { return check(cts_synrad(gpu, turn), "synrad"); }
int MPBeam::SSBeamIonEffectOneInteractionPoint(int turn, int k)
{
    return check(cts_ion_phase_ip(gpu, turn, k), "ions");
}

int MPBeam::SRWakeBeamIntaction()
{
    std::vector<double> sz(beamVec.size());
    for (size_t i = 0; i < beamVec.size(); i++) sz[i] = beamVec[i].rmsBunchLength;
    return check(cts_sr_wake(gpu, sz.data()), "sr wake");
}

int MPBeam::MarkParticleLostInBunch()
{
    std::vector<double> t(beamVec.size());
    const int rc = check(cts_mark_lost(gpu, t.data()), "mark lost");
    for (size_t i = 0; i < beamVec.size(); i++) beamVec[i].transmission = t[i];
    return rc;
}

int MPBeam::MPBeamRMSCal(int k)
{
    std::vector<cts_moments> m(beamVec.size());
    if (check(cts_rms(gpu, m.data()), "rms")) return 1;
    for (size_t i = 0; i < beamVec.size(); i++) {
        MPBunch &b = beamVec[i];
        cts_derived d{};
        cts_derive_moments(gpu, &m[i], k, &d);
        b.xAver = m[i].mean[0]; b.pxAver = m[i].mean[1]; b.yAver = m[i].mean[2];
        b.pyAver = m[i].mean[3]; b.zAver = m[i].mean[4]; b.pzAver = m[i].mean[5];
        for (int j = 0; j < 3; j++) { b.emittance[j] = d.emit[j]; b.effectiveEmittance[j] = d.emit_eff[j]; }
        b.rmsRx = d.rms_x; b.rmsRy = d.rms_y; b.rmsEffectiveRx = d.rms_eff_x; b.rmsEffectiveRy = d.rms_eff_y;
        b.rmsBunchLength = d.sigma_z; b.rmsEnergySpread = d.sigma_dp;
    }
    return 0;
}

int MPBeam::DownloadParticles()
{
    size_t n = 0;
    for (auto &b : beamVec) n += b.macroEleNumPerBunch;
    std::vector<double> X[6];
    for (auto &c : X) c.resize(n);
    std::vector<int32_t> lost(n);
    if (check(cts_download(gpu, X[0].data(), X[1].data(), X[2].data(), X[3].data(), X[4].data(), X[5].data(),
                           lost.data()), "download"))
        return 1;
    size_t off = 0;
    for (auto &b : beamVec) {
        std::vector<double> *dst[6] = {&b.ePositionX, &b.eMomentumX, &b.ePositionY,
                                       &b.eMomentumY, &b.ePositionZ, &b.eMomentumZ};
        for (int c = 0; c < 6; c++) dst[c]->assign(X[c].begin() + off, X[c].begin() + off + b.macroEleNumPerBunch);
        b.eSurive.assign(lost.begin() + off, lost.begin() + off + b.macroEleNumPerBunch);
        off += b.macroEleNumPerBunch;
    }
    return 0;
}

int MPBeam::GetCBMGR(ReadInputSettings &in, std::vector<double> &rates)
{
    const auto *R = in.ringRun;
    const int interval = std::max(R->bunchInfoPrintInterval, 1);
    const int nb = (int)beamVec.size();
    const int ns = (int)(samples_.size() / (size_t)std::max(nb, 1));
    const size_t N = (size_t)ns * 3 * nb;
    std::vector<int32_t> harm(nb);
    for (int b = 0; b < nb; b++) harm[b] = beamVec[b].bunchHarmNum;
    // GetCBMGR reads workQx/y/z and rmsBunchLength / rmsEnergySpread of the
    // input (src/MPBeam.cpp:1498-1508)
    const double tunes[3] = {in.ringParBasic->workQx, in.ringParBasic->workQy, in.ringParBasic->workQz};
    const double twiss[5] = {alphaX0_, betaX0_, alphaY0_, betaY0_,
                             in.ringBunchPara->rmsBunchLength / in.ringBunchPara->rmsEnergySpread};
    std::vector<double> a[8], r[4];
    for (auto &v : a) v.assign(N, 0.0);
    for (auto &v : r) v.assign(3 * (size_t)nb, 0.0);
    cts_cbm_result res{a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(), a[5].data(),
                       a[6].data(), a[7].data(), r[0].data(), r[1].data(), r[2].data(), r[3].data()};
    if (check(cts_cbm_analysis(samples_.data(), ns, nb, harm.data(), in.ringParBasic->harmonics, tunes, twiss,
                               interval, R->growthRateFittingStart, R->growthRateFittingEnd, in.ringParBasic->t0,
                               &res), "growth rates"))
        return 1;
    rates = r[0];
    // runCBMGR.sdds in the reference's layout (src/MPBeam.cpp:1735-1886),
    // its repeated columns included (ArgIQY/Z print the x phase;
    // AbsCBMHilbertAnaY and ArgCBMHilbertAnaY print x)
    std::ofstream fout(R->runCBMGR + ".sdds");
    const char *head[] = {
        "SDDS1",
        "&parameter name=dataPointsForFit                       type=long     &end",
        "&parameter name=ModeIndex                              type=long     &end",
        "&parameter name=CBMIdealGRX,                  units=1/s,    type=float    &end",
        "&parameter name=CBMIdealGRY,                  units=1/s,    type=float    &end",
        "&parameter name=CBMIdealGRZ,                  units=1/s,    type=float    &end",
        "&parameter name=CBMIQGRX,                units=1/s,    type=float    &end",
        "&parameter name=CBMIQGRY,                units=1/s,    type=float    &end",
        "&parameter name=CBMIQGRZ,                units=1/s,    type=float    &end",
        "&parameter name=CBMHTGRX,                units=1/s,    type=float    &end",
        "&parameter name=CBMHTGRY,                units=1/s,    type=float    &end",
        "&parameter name=CBMHTGRZ,                units=1/s,    type=float    &end",
        "&parameter name=BunchHarmIndex                         type=long     &end",
        "&parameter name=BGRX,                    units=1/s,    type=float    &end",
        "&parameter name=BGRY,                    units=1/s,    type=float    &end",
        "&parameter name=BGRZ,                    units=1/s,    type=float    &end",
        "&column name=Turns,                                    type=long      &end",
        "&column name=AbsCBMIdealX,                              type=float     &end",
        "&column name=AbsCBMIdealY,                              type=float     &end",
        "&column name=AbsCBMIdealZ,                              type=float     &end",
        "&column name=ArgCBMIdealX,                units=rad,    type=float     &end",
        "&column name=ArgCBMIdealY,                units=rad,    type=float     &end",
        "&column name=ArgCBMIdealZ,                units=rad,    type=float     &end",
        "&column name=AbsIQX,                                   type=float     &end",
        "&column name=AbsIQY,                                   type=float     &end",
        "&column name=AbsIQZ,                                   type=float     &end",
        "&column name=ArgIQX,                     units=rad,    type=float     &end",
        "&column name=ArgIQY,                     units=rad,    type=float     &end",
        "&column name=ArgIQZ,                     units=rad,    type=float     &end",
        "&column name=AbsCBMHilbertAnaX,         units=m,      type=float     &end",
        "&column name=AbsCBMHilbertAnaY,         units=m,      type=float     &end",
        "&column name=AbsCBMHilbertAnaZ,         units=m,      type=float     &end",
        "&column name=ArgCBMHilbertAnaX,         units=rad,    type=float     &end",
        "&column name=ArgCBMHilbertAnaY,         units=rad,    type=float     &end",
        "&column name=ArgCBMHilbertAnaZ,         units=rad,    type=float     &end",
        "&column name=AverX,                      units=m,      type=float     &end",
        "&column name=AverY,                      units=m,      type=float     &end",
        "&column name=AverZ,                      units=m,      type=float     &end",
        "&column name=AnalyticalAverX,            units=m,      type=float     &end",
        "&column name=AnalyticalAverY,            units=m,      type=float     &end",
        "&column name=AnalyticalAverZ,            units=m,      type=float     &end",
        "&column name=AbsHilbertAnaX,                 units=m,      type=float     &end",
        "&column name=AbsHilbertAnaY,                 units=m,      type=float     &end",
        "&column name=AbsHilbertAnaZ,                 units=m,      type=float     &end",
        "&data mode=ascii &end"};
    for (const char *l : head) fout << l << std::endl;
    const int dim = R->growthRateFittingEnd / interval - R->growthRateFittingStart / interval;
    auto at = [&](int n, int pl, int k) { return ((size_t)n * 3 + pl) * nb + k; };
    auto w = [&](double v) { fout << std::setw(15) << std::left << v; };
    for (int i = 0; i < nb; i++) {
        fout << "! page number " << i << std::endl;
        fout << std::setw(15) << std::left << dim << std::endl;
        fout << std::setw(15) << std::left << i << std::endl;
        for (int k = 0; k < 3; k++)
            for (int pl = 0; pl < 3; pl++) {
                w(r[k][(size_t)pl * nb + i]);
                fout << std::endl;
            }
        fout << std::setw(15) << std::left << beamVec[i].bunchHarmNum << std::endl;
        for (int pl = 0; pl < 3; pl++) {
            w(r[3][(size_t)pl * nb + i]);
            fout << std::endl;
        }
        fout << ns << std::endl;
        for (int n = 0; n < ns; n++) {
            fout << std::setw(15) << std::left << n * interval;
            for (int pl = 0; pl < 3; pl++) w(a[0][at(n, pl, i)]);
            for (int pl = 0; pl < 3; pl++) w(a[1][at(n, pl, i)]);
            for (int pl = 0; pl < 3; pl++) w(a[2][at(n, pl, i)]);
            for (int pl = 0; pl < 3; pl++) w(a[3][at(n, 0, i)]);
            w(a[6][at(n, 0, i)]); w(a[6][at(n, 0, i)]); w(a[6][at(n, 2, i)]);
            w(a[7][at(n, 0, i)]); w(a[7][at(n, 0, i)]); w(a[7][at(n, 2, i)]);
            for (int pl = 0; pl < 3; pl++) w(samples_[(size_t)n * nb + i].mean[2 * pl]);
            for (int pl = 0; pl < 3; pl++) w(a[4][at(n, pl, i)]);
            for (int pl = 0; pl < 3; pl++) w(std::hypot(a[4][at(n, pl, i)], a[5][at(n, pl, i)]));
            fout << std::endl;
        }
    }
    return 0;
}

// One page of TBTBunchAverData.sdds (MPBeam::MPBeamDataPrintPerTurn,
// src/MPBeam.cpp:2319-2411): the moments of every bunch at the end of turn
// `turn`, the ion charge its passage saw, and the cavity phasors
// (cavInfo: cts_get_cavity_info of that turn when beam loading is on).
void MPBeam::PrintResultHeader(std::ofstream &fout, const ReadInputSettings &in)
{
    const auto *B = in.ringParBasic;
    const char *head[] = {
        "SDDS1",
        "&parameter name=beamCurr,    units=mA,   type=float,  &end",
        "&parameter name=I1,          units=m,   type=float,  &end",
        "&parameter name=I2,          units=1/m, type=float,  &end",
        "&parameter name=I3,          units=1/m^2,   type=float,  &end",
        "&parameter name=I4,          units=1/m, type=float,  &end",
        "&parameter name=I5,          units=1/m, type=float,  &end",
        "&parameter name=Jx,                     type=float,  &end",
        "&parameter name=Jy,                     type=float,  &end",
        "&parameter name=Jz,                     type=float,  &end",
        "&parameter name=f1001Abs,               type=float,  &end",
        "&parameter name=f1010Abs,               type=float,  &end",
        "&parameter name=f0110Abs,               type=float,  &end",
        "&parameter name=f0101Abs,               type=float,  &end",
        "&parameter name=f1001Arg,  units=rad,   type=float,  &end",
        "&parameter name=f1010Arg,  units=rad,   type=float,  &end",
        "&parameter name=f0110Arg,  units=rad,   type=float,  &end",
        "&parameter name=f0101Arg,  units=rad,   type=float,  &end",
        "&parameter name=LCFactor,               type=float,  &end",
        "&column name=Turns,                     type=long,   &end",
        "&column name=IonCharge,      units=e,   type=float,  &end",
        "&column name=maxAverX,       units=m,   type=float,  &end",
        "&column name=maxAverY,       units=m,   type=float,  &end",
        "&column name=averAllBunchX,  units=m,   type=float,  &end",
        "&column name=averAllBunchY,  units=m,   type=float,  &end",
        "&column name=averAllBunchZ,  units=m,   type=float,  &end",
        "&column name=averAllBunchPX, units=rad, type=float,  &end",
        "&column name=averAllBunchPY, units=rad, type=float,  &end",
        "&column name=averAllBunchPZ, units=rad, type=float,  &end",
        "&column name=rmsAllBunchX,   units=m,   type=float,  &end",
        "&column name=rmsAllBunchY,   units=m,   type=float,  &end",
        "&column name=rmsAllBunchZ,   units=m,   type=float,  &end",
        "&column name=rmsAllBunchPX,  units=rad, type=float,  &end",
        "&column name=rmsAllBunchPY,  units=rad, type=float,  &end",
        "&column name=rmsAllBunchPZ,  units=rad, type=float,  &end",
        "&column name=Nux,                       type=float,  &end",
        "&column name=Nuy,                       type=float,  &end",
        "&column name=deltaNu,                   type=float,  &end",
        "&column name=f1001Abs,               type=float,  &end",
        "&column name=f1010Abs,               type=float,  &end",
        "&column name=f1001Arg,  units=rad,   type=float,  &end",
        "&column name=f1010Arg,  units=rad,   type=float,  &end",
        "&column name=LCFactor,               type=float,  &end",
        "&column name=xyAlpha,   units=rad,   type=float,  &end",
        "&column name=traceA,                 type=float,  &end",
        "&column name=traceB,                 type=float,  &end",
        "&column name=traceMSubN,             type=float,  &end",
        "&column name=gamma0,                 type=float,  &end",
        "&column name=gamma1,                 type=float,  &end",
        "&column name=detH,                   type=float,  &end"};
    for (const char *l : head) fout << l << std::endl;
    // per printed bunch j: the reference's column names and units
    const char *cols[16][2] = {{"averX_", "units=m,   "},           {"averPx_", "units=rad, "},
                               {"averY_", "units=m,   "},           {"averPy_", "units=rad, "},
                               {"averZ_", "units=m,   "},           {"averPz_", "units=rad, "},
                               {"rmsBunchZ_", "units=m,   "},       {"rmsEnergySpread_", "units=rad, "},
                               {"rmsRx_", "units=m,   "},           {"rmsRy_", "units=m,   "},
                               {"rmsEmitx_", "units=m*rad,   "},    {"rmsEmity_", "units=m*rad,   "},
                               {"eigenEmitx_", "units=m*rad,   "},  {"eigenEmity_", "units=m*rad,    "},
                               {"transmission_", "               "}, {"xyCouplingAlpha_", "units=rad,  "}};
    for (int j = 0; j < in.ringRun->TBTBunchPrintNum; j++)
        for (auto &c : cols)
            fout << "&column name=" << c[0] << j << ", " << c[1] << "type=float,  &end" << std::endl;
    fout << "&data mode=ascii, &end" << std::endl;
    // radiation integrals and damping partitions from the damping times
    // (src/ReadInputSettings.cpp:1054-1077): J_x = 2 E / (U0 tau_x),
    // J_y = 1, J_z = 3 - J_x; alpha_i = J_i U0 / (2 T0 E)
    const double E = B->electronBeamEnergy, U0 = B->u0;
    double J[3];
    J[0] = 2 * E / U0 / B->synchRadDampTime[0];
    J[1] = 1;
    J[2] = 3 - J[0];
    const double a0 = U0 / (2 * B->t0 * E);
    double al[3];
    for (int i = 0; i < 3; i++) al[i] = J[i] * a0;
    const double re = 2.81794092E-15, clight = 299792458.0, cq = 3.83E-13;  // include/Global.h:27, 34, 38
    const double g3 = std::pow(B->rGamma, 3), g2 = std::pow(B->rGamma, 2);
    double I[5];
    I[0] = B->circRing * B->alphac[0];
    I[1] = B->circRing * (al[0] + al[2]) / (re * clight * g3);
    I[3] = B->circRing * (al[2] - 2 * al[0]) / (re * clight * g3);
    I[2] = (2 * I[1] + I[3]) * std::pow(B->sdelta0, 2) / (cq * g2);
    I[4] = B->emitNat[0] * (I[1] - I[3]) / (cq * g2);
    fout << B->ringCurrent << std::endl;
    for (double v : I) fout << v << std::endl;
    for (double v : J) fout << v << std::endl;
    // resonance driving terms and the linear coupling factor of the
    // lattice-coupling analysis: not computed, written as 0
    for (int i = 0; i < 9; i++) fout << 0.0 << std::endl;
    fout << "! page number " << 1 << std::endl;
    fout << in.ringRun->nTurns << std::endl;
}

void MPBeam::PrintResultRow(std::ofstream &fout, int turn, const cts_moments *m, const double *eig, double ionCharge,
                            const ReadInputSettings &in)
{
    // MPGetBeamInfo (src/MPBeam.cpp:2135-2317): the largest centroids and the
    // means and rms over the bunches of the centroids
    const size_t nb = beamVec.size();
    double mx = m[0].mean[0], my = m[0].mean[2];
    double av[6] = {0, 0, 0, 0, 0, 0}, rm[6] = {0, 0, 0, 0, 0, 0};
    const int ord[6] = {0, 2, 4, 1, 3, 5};  // x y z px py pz
    for (size_t i = 0; i < nb; i++) {
        if (i > 0 && m[i].mean[0] > mx) mx = m[i].mean[0];
        if (i > 0 && m[i].mean[2] > my) my = m[i].mean[2];
        for (int c = 0; c < 6; c++) {
            av[c] += m[i].mean[ord[c]];
            rm[c] += std::pow(m[i].mean[ord[c]], 2);
        }
    }
    auto w = [&](double v) { fout << std::setw(15) << std::left << v; };
    fout << std::setw(15) << std::left << turn;
    w(ionCharge);
    w(mx);
    w(my);
    for (int c = 0; c < 6; c++) w(av[c] / nb);
    for (int c = 0; c < 6; c++) w(std::sqrt(rm[c] / nb));
    const double nux = in.ringParBasic->workQx - std::floor(in.ringParBasic->workQx);
    const double nuy = in.ringParBasic->workQy - std::floor(in.ringParBasic->workQy);
    w(nux);
    w(nuy);
    w(nuy - nux);
    for (int c = 0; c < 12; c++) w(0.0);  // lattice-coupling columns, not computed
    for (int j = 0; j < in.ringRun->TBTBunchPrintNum; j++) {
        cts_derived d{};
        cts_derive_moments(gpu, &m[j], 0, &d);
        w(m[j].mean[0]); w(m[j].mean[1]); w(m[j].mean[2]); w(m[j].mean[3]); w(m[j].mean[4]); w(m[j].mean[5]);
        w(d.sigma_z); w(d.sigma_dp); w(d.rms_x); w(d.rms_y); w(d.emit[0]); w(d.emit[1]);
        w(eig[4 * j]); w(eig[4 * j + 1]); w(m[j].transmission); w(eig[4 * j + 3]);
    }
    fout << std::endl;
}

void MPBeam::PrintTBTBunchAverData(std::ofstream &fout, int turn, const cts_moments *m, const double *ionCharge,
                                   const double *cavInfo,
                                   const ReadInputSettings &in, bool header)
{
    const auto *RF = in.ringParRf;
    if (header) {
        const char *head[] = {
            "SDDS1",
            "&column name=Turns,                           type=long,  &end",
            "&column name=HarmIndex,                       type=long,  &end",
            "&column name=BunchIndex,                       type=long,  &end",
            "&column name=TotIonCharge,       units=e,     type=float,  &end",
            "&column name=AverX,              units=m,     type=float,  &end",
            "&column name=AverY,              units=m,     type=float,  &end",
            "&column name=AverZ,              units=m,     type=float,  &end",
            "&column name=AverXP,             units=rad,   type=float,  &end",
            "&column name=AverYP,             units=rad,   type=float,  &end",
            "&column name=AverZP,             units=rad,   type=float,  &end",
            "&column name=RmsEmitX,           units=m*rad, type=float,  &end",
            "&column name=RmsEmitY,           units=m*rad, type=float,  &end",
            "&column name=RmsEmitZ,           units=m*rad, type=float,  &end",
            "&column name=RmsRX,              units=m*rad, type=float,  &end",
            "&column name=RmsRY,              units=m*rad, type=float,  &end",
            "&column name=RmsBunchLength,     units=m,     type=float,  &end",
            "&column name=RmsBunchEnergySpread, units=rad, type=float,  &end"};
        for (const char *l : head) fout << l << std::endl;
        for (int j = 0; j < RF->resNum; j++) {
            const std::string s = std::to_string(j);
            fout << "&column name=cavAmp_" << s << ", units=V,   type=float,  &end" << std::endl;
            fout << "&column name=cavPhase_" << s << ", units=rad,   type=float,  &end" << std::endl;
            fout << "&column name=cavReal_" << s << ", units=V,   type=float,  &end" << std::endl;
            fout << "&column name=genAmp_" << s << ", units=V,   type=float,  &end" << std::endl;
            fout << "&column name=genPhase_" << s << ", units=rad,   type=float,  &end" << std::endl;
            fout << "&column name=beamIndAmp_" << s << ", units=V,   type=float,  &end" << std::endl;
            fout << "&column name=beamIndPhase_" << s << ", units=rad,   type=float,  &end" << std::endl;
        }
        fout << "&data mode=ascii, &end" << std::endl;
    }
    fout << "! page number " << turn + 1 << std::endl;
    fout << beamVec.size() << std::endl;
    const double rBeta = in.ringParBasic->rBeta, fRF = in.ringParBasic->f0 * in.ringParRf->ringHarm;
    for (size_t i = 0; i < beamVec.size(); i++) {
        cts_derived d{};
        cts_derive_moments(gpu, &m[i], 0, &d);
        auto w = [&](double v) { fout << std::setw(24) << std::left << std::setprecision(16) << v; };
        fout << std::setw(24) << std::left << std::setprecision(16) << turn;
        fout << std::setw(24) << std::left << std::setprecision(16) << beamVec[i].bunchHarmNum;
        fout << std::setw(24) << std::left << std::setprecision(16) << i;
        w(ionCharge ? ionCharge[i] : 0.0);
        w(m[i].mean[0]); w(m[i].mean[2]); w(m[i].mean[4]);
        w(m[i].mean[1]); w(m[i].mean[3]); w(m[i].mean[5]);
        w(d.emit[0]); w(d.emit[1]); w(d.emit[2]);
        w(d.rms_x); w(d.rms_y); w(d.sigma_z); w(d.sigma_dp);
        // ideal cavities: the generator phasor, which is also the cavity
        // voltage, at the bunch's mean z; no beam-induced part
        for (int j = 0; j < RF->resNum; j++) {
            if (cavInfo && RF->rfMode[j] == 1) {
                // beam-loaded: cavity, generator and beam-induced voltages
                // the bunch saw (BunchMomentumUpdateDueToRFModeStable)
                const double *f = cavInfo + ((size_t)i * RF->resNum + j) * 8;
                const std::complex<double> cav(f[0], f[1]), ind(f[2], f[3]), gen(f[4], f[5]);
                w(std::abs(cav)); w(std::arg(cav)); w(cav.real());
                w(std::abs(gen)); w(std::arg(gen));
                w(std::abs(ind)); w(std::arg(ind));
                continue;
            }
            const std::complex<double> req(cavity_[2 * j], cavity_[2 * j + 1]);
            const double ph = -m[i].mean[4] / kClight / rBeta * 2. * kPi * (double)RF->resHarm[j] * fRF;
            const std::complex<double> v = req * std::exp(std::complex<double>(0.0, ph));
            w(std::abs(v)); w(std::arg(v)); w(v.real());
            w(std::abs(v)); w(std::arg(v));
            w(0.0); w(0.0);
        }
        fout << std::endl;
    }
}

int MPBeam::RunSettings(const LatticeInterActionPoint &lat, const ReadInputSettings &in, RunCalls &c)
{
    const auto *R = in.ringRun;
    c = RunCalls{};
    if (R->sRWakeFlag) {
        WakeFunction sr;
        if (sr.InitialSRWake(in)) return 1;
        c.wake = true;
        c.nbins = in.ringSRWake->SRWBunchBinNum;
        c.rs = sr.lRs;
        c.q = sr.lQ;
        c.fr.resize(sr.lOmega.size());
        for (size_t i = 0; i < c.fr.size(); i++) c.fr[i] = sr.lOmega[i] / (2 * kPi);
    }
    if (R->lRWakeFlag) {
        if (lat.sections.size() != 1) {
            std::cerr << "long-range wake: one interaction point (fused turn) only" << std::endl;
            return 1;
        }
        WakeFunction lw;
        if (lw.InitialLRWake(in, lat)) return 1;
        c.lrw = true;
        c.lrw_nturns = in.ringLRWake->nTurnswakeTrunction;
        for (size_t i = 0; i < lw.sectorRadiusX.size(); i++)
            c.lrw_sectors.insert(c.lrw_sectors.end(), {lw.sectorRadiusX[i], lw.sectorRadiusY[i], lw.sectorLength[i],
                                                       lw.sectorBetaX[i], lw.sectorBetaY[i], lw.sectorNum[i],
                                                       lw.sectormatSigma[i]});
        for (size_t i = 0; i < lw.lRs.size(); i++)
            c.lrw_res.insert(c.lrw_res.end(), {lw.lRs[i], lw.lQ[i], lw.lOmega[i] / (2 * kPi), lw.txRs[i], lw.txQ[i],
                                               lw.txOmega[i] / (2 * kPi), lw.tyRs[i], lw.tyQ[i],
                                               lw.tyOmega[i] / (2 * kPi)});
        c.lrw_beta[0] = lw.betaFunIntPoint[0];
        c.lrw_beta[1] = lw.betaFunIntPoint[1];
        c.lrw_output = in.ringLRWake->lrwOutput;
    }
    // broadband impedance: BoardBandImp::ReadInImp (src/BoardBandImp.cpp:42-70):
    // 5 header lines, then frequency and ten impedance columns per row;
    // Z_|| = column 2 - i column 7, Z_Qx = -column 10 - i column 5,
    // Z_Qy = -column 11 - i column 6.  In the time domain (the default) the
    // wake file instead (MPBeam::GetQuasiWakePoten, src/MPBeam.cpp:748-812)
    if (R->bBImpFlag && in.ringBBImp->timeDomain == 1) {
        if (lat.sections.size() != 1) {
            std::cerr << "broadband impedance: one interaction point (fused turn) only" << std::endl;
            return 1;
        }
        const auto *B = in.ringBBImp;
        if (ReadQuasiWake(B->wakeInput, c.bbi_wake)) return 1;
        c.bbi = true;
        for (int i = 0; i < 5; i++) c.bbi_flags[i] = B->impedSimFlag[i];
        c.bbi_time_domain = 1;
        c.bbi_beta[0] = lat.twissBetaX[0];
        c.bbi_beta[1] = lat.twissBetaY[0];
    } else if (R->bBImpFlag) {
        if (lat.sections.size() != 1) {
            std::cerr << "broadband impedance: one interaction point (fused turn) only" << std::endl;
            return 1;
        }
        const auto *B = in.ringBBImp;
        std::ifstream f(B->impedInput);
        if (!f) {
            std::cerr << "cannot open the impedance file " << B->impedInput << std::endl;
            return 1;
        }
        std::string line;
        double fmax = 0;
        for (int ln = 0; std::getline(f, line); ln++) {
            if (ln < 5) continue;
            std::istringstream ss(line);
            std::vector<double> col;
            double x;
            while (ss >> x) col.push_back(x);
            if (col.empty()) continue;
            if (col.size() < 11) {
                std::cerr << B->impedInput << ": row " << ln + 1 << " has " << col.size() << " columns, 11 needed"
                          << std::endl;
                return 1;
            }
            fmax = col[0];
            c.bbi_re.push_back(col[1]);
            c.bbi_im.push_back(-col[6]);
            c.bbi_qx_re.push_back(-col[9]);
            c.bbi_qx_im.push_back(-col[4]);
            c.bbi_qy_re.push_back(-col[10]);
            c.bbi_qy_im.push_back(-col[5]);
            c.bbi_dx_re.push_back(-col[7]);
            c.bbi_dx_im.push_back(-col[2]);
            c.bbi_dy_re.push_back(-col[8]);
            c.bbi_dy_im.push_back(-col[3]);
        }
        c.bbi = true;
        c.bbi_fmax = fmax;
        for (int i = 0; i < 5; i++) c.bbi_flags[i] = B->impedSimFlag[i];
        c.bbi_time_domain = B->timeDomain;
        c.bbi_beta[0] = lat.twissBetaX[0];
        c.bbi_beta[1] = lat.twissBetaY[0];
    }
    // rfMode-1 cavities: the reference's stable beam-loading model
    // (MPBeam::BeamMomtumUpdateDueToRFTest, src/MPBeam.cpp:2624-2706)
    const auto *RF = in.ringParRf;
    for (int i = 0; i < RF->resNum; i++) c.bl = c.bl || RF->rfMode[i] == 1;
    if (c.bl) {
        if (lat.sections.size() != 1) {
            std::cerr << "beam loading (rfMode 1): one interaction point (fused turn) only" << std::endl;
            return 1;
        }
        cts_beam_loading_config &b = c.blc;
        b.nres = RF->resNum;
        for (int i = 0; i < RF->resNum; i++) {
            b.rf_mode[i] = RF->rfMode[i];
            b.res_type[i] = RF->resType[i];
            b.res_harm[i] = RF->resHarm[i];
            b.rs[i] = RF->resShuntImpRs[i];
            b.q0[i] = RF->resQualityQ0[i];
            b.coupling[i] = RF->resCouplingBeta[i];
            b.detune_hz[i] = RF->resDetuneFre[i];
            b.vol[i] = RF->resVol[i];
            b.phase[i] = RF->resPhase[i];
            b.cold[i] = RF->resCold[i];
            b.excite_instability[i] = RF->resExciteIntability[i];
            b.dir_fb[i] = RF->resDirFB[i];
            b.dir_fb_gain[i] = RF->resDirFBGain[i];
            b.dir_fb_phase[i] = RF->resDirFBPhase[i];
            b.dir_fb_delay[i] = RF->resDirFBDelay[i];
        }
        b.nbins = RF->rfBunchBinNum;
        b.nturns = R->nTurns;
    }
    c.synrad = R->synRadDampingFlag != 0;
    c.seed = R->seed;
    if (R->beamIonFlag) {
        if (lat.gasSpec > CTS_MAX_SPECIES) {
            std::cerr << "at most " << CTS_MAX_SPECIES << " gas species" << std::endl;
            return 1;
        }
        // a cloud at every interaction point, with that point's vacuum
        // pressure, temperature, interaction length and aperture
        // (LatticeInterActionPoint::InitialLattice, src/LatticeInterActionPoint.cpp:232-250)
        c.ions = true;
        c.ion.assign(lat.sections.size(), cts_ion_config{});
        for (size_t k = 0; k < lat.sections.size(); k++) {
            cts_ion_config &ic = c.ion[k];
            ic.nspec = lat.gasSpec;
            ic.n_new = in.ringIonEffPara->macroIonNumberGeneratedPerIP;
            ic.max_ions = (int64_t)lat.ionMaxNumberOneInterPoint;
            for (int p = 0; p < ic.nspec; p++) {
                ic.mass_number[p] = lat.ionMassNumber[p];
                ic.cross_section[p] = lat.corssSectionEI[p];
                ic.pressure[p] = lat.gasPercent[p] * lat.vacuumPressure[k];
            }
            ic.temperature = lat.temperature[k];
            ic.length = lat.interactionLength[k];
            ic.loss_boundary = lat.ionLossBoundary;
            ic.ap_x = lat.pipeAperatureX[k];
            ic.ap_y = lat.pipeAperatureY[k];
            // each point draws from its own stream
            ic.seed = R->seed + 0x9E3779B97F4A7C15ull * k;
        }
    }
    return 0;
}

int MPBeam::PrintLRWakeTable(const std::string &path)
{
    int32_t np = 0;
    (void)cts_get_lrwake_table(gpu, &np, nullptr, nullptr, nullptr, nullptr);  // np
    std::vector<double> t(np), w(3 * (size_t)np), rw(3 * (size_t)np), bbr(3 * (size_t)np);
    if (check(cts_get_lrwake_table(gpu, &np, t.data(), w.data(), rw.data(), bbr.data()), "wake table")) return 1;
    std::ofstream fout(path);
    // the SDDS layout of the reference's file, column header lines verbatim
    fout << "SDDS1" << std::endl;
    fout << "&column name=time,                    units=s,      type=float,  &end" << std::endl;
    const char *names[9] = {"rw_wx", "rw_wy", "rw_wz", "bbr_wx", "bbr_wy", "bbr_wz", "tot_wx", "tot_wy", "tot_wz"};
    for (int c = 0; c < 9; c++)
        fout << "&column name=" << names[c] << ",              units=" << (c % 3 == 2 ? "V/C,    " : "V/C/m,  ")
             << "type=float,  &end" << std::endl;
    fout << "&data mode=ascii, &end" << std::endl;
    fout << "! page number 1" << std::endl;
    fout << np << std::endl;
    for (int i = 0; i < np; i++) {
        // the reference prints the table's times before flipping their sign
        fout << std::setw(15) << std::left << (t[i] == 0.0 ? 0.0 : -t[i]);
        const double *src[3] = {rw.data(), bbr.data(), w.data()};
        for (int g = 0; g < 3; g++)
            for (int k = 0; k < 3; k++) fout << std::setw(15) << std::left << src[g][(size_t)k * np + i];
        fout << std::endl;
    }
    return fout.good() ? 0 : 1;
}

// MPBeam::Run, multi-particle tracking (src/MPBeam.cpp:1010-1173).  With one
// section the whole turn is one cts_turn call and the moments of the print
// turns come from the device history after the loop (no host round trip
// per turn); with several sections the seams run in the reference's order.
// Samples are the moments at the end of turns n with n % interval == 0, as
// MPBeamDataPrintPerTurn and GetCBMGR take them (src/MPBeam.cpp:1161-1172).
int MPBeam::Run(Train &, LatticeInterActionPoint &lat, ReadInputSettings &in)
{
    const auto *R = in.ringRun;
    const int nsec = (int)lat.sections.size();
    const int nb = (int)beamVec.size();
    alphaX0_ = lat.twissAlphaX[0];
    betaX0_ = lat.twissBetaX[0];
    alphaY0_ = lat.twissAlphaY[0];
    betaY0_ = lat.twissBetaY[0];
    RunCalls c;
    if (RunSettings(lat, in, c)) return 1;
    if (c.wake && check(cts_set_wake(gpu, c.nbins, 0, 0, 0, (int32_t)c.rs.size(), c.rs.data(), c.q.data(),
                                     c.fr.data()), "wake"))
        return 1;
    if (check(cts_set_synrad(gpu, c.synrad, 2, c.seed), "synrad")) return 1;
    for (size_t k = 0; c.ions && k < c.ion.size(); k++)
        if (check(cts_set_ions_ip(gpu, (int32_t)k, &c.ion[k]), "ions")) return 1;
    if (c.lrw) {
        cts_lrwake_config lc{};
        lc.nturns = c.lrw_nturns;
        lc.n_sectors = (int32_t)(c.lrw_sectors.size() / 7);
        lc.sectors = c.lrw_sectors.data();
        lc.n_res = (int32_t)(c.lrw_res.size() / 9);
        lc.res = c.lrw_res.data();
        lc.beta_x = c.lrw_beta[0];
        lc.beta_y = c.lrw_beta[1];
        if (check(cts_set_lrwake(gpu, &lc), "long-range wake")) return 1;
        if (!c.lrw_output.empty() && PrintLRWakeTable(c.lrw_output)) return 1;
    }
    if (c.bl && check(cts_set_beam_loading(gpu, &c.blc), "beam loading")) return 1;
    if (c.bbi) {
        cts_bbimp_config bc{};
        bc.nimp = (int32_t)c.bbi_re.size();
        bc.f_max = c.bbi_fmax;
        bc.zz_re = c.bbi_re.data();
        bc.zz_im = c.bbi_im.data();
        for (int i = 0; i < 5; i++) bc.flags[i] = c.bbi_flags[i];
        bc.time_domain = c.bbi_time_domain;
        bc.zqx_re = c.bbi_qx_re.data();
        bc.zqx_im = c.bbi_qx_im.data();
        bc.zqy_re = c.bbi_qy_re.data();
        bc.zqy_im = c.bbi_qy_im.data();
        bc.zdx_re = c.bbi_dx_re.data();
        bc.zdx_im = c.bbi_dx_im.data();
        bc.zdy_re = c.bbi_dy_re.data();
        bc.zdy_im = c.bbi_dy_im.data();
        bc.beta_x = c.bbi_beta[0];
        bc.beta_y = c.bbi_beta[1];
        bc.nwake = (int32_t)c.bbi_wake[0].size();
        const double **wk[6] = {&bc.wake_pos, &bc.wake_z, &bc.wake_dx, &bc.wake_dy, &bc.wake_qx, &bc.wake_qy};
        for (int i = 0; i < 6; i++) *wk[i] = c.bbi_wake[i].empty() ? nullptr : c.bbi_wake[i].data();
        if (check(cts_set_bbimp(gpu, &bc), "broadband impedance")) return 1;
    }
    const int interval = std::max(R->bunchInfoPrintInterval, 1);
    samples_.clear();
    // output files start empty, as the reference removes *.sdds at set-up
    // (RMOutPutFiles, src/Global.cpp:48-55)
    std::ofstream tbt;
    if (!R->TBTBunchAverData.empty()) tbt.open(R->TBTBunchAverData + ".sdds");
    const int nsamp = (R->nTurns + interval - 1) / interval;
    std::vector<double> ionq((size_t)nsamp * nb, 0.0);
    std::vector<double> pas((size_t)nb * CTS_MAX_SPECIES * 6);
    // the ion number every bunch saw, summed over the interaction points
    auto record_ions = [&](int sample) -> int {
        if (!c.ions) return 0;
        for (int k = 0; k < nsec; k++) {
            if (check(cts_get_ion_passages_ip(gpu, k, pas.data()), "ion passages")) return 1;
            for (int b = 0; b < nb; b++) {
                double q = 0;
                for (int p = 0; p < lat.gasSpec; p++) q += pas[((size_t)b * CTS_MAX_SPECIES + p) * 6 + 5];
                ionq[(size_t)sample * nb + b] += q;
            }
        }
        return 0;
    };
    std::vector<cts_moments> m(nb);
    if (R->TBTBunchPrintNum > nb) {
        std::cerr << "wrong settings: Total Bunch Number is " << nb << std::endl;
        return 1;
    }
    // result.sdds: a row at the end of every turn, from the moments after
    // the turn's kicks (MPBeamRMSCal(0)), the eigen-emittances and the ion
    // charge; the turns then run one by one
    std::ofstream result;
    std::vector<double> eig((size_t)nb * 4);
    auto result_row = [&](int n) -> int {
        double ionc = 0.0;
        if (check(cts_rms(gpu, m.data()), "rms") || check(cts_eigen_emit(gpu, eig.data()), "eigen-emittances") ||
            check(cts_get_ion_charge(gpu, &ionc), "ion charge"))
            return 1;
        PrintResultRow(result, n, m.data(), eig.data(), ionc, in);
        return 0;
    };
    if (writeResult) {
        result.open("result.sdds");
        PrintResultHeader(result, in);
    }
    if (nsec == 1 && writeResult) {
        for (int n = 0; n < R->nTurns; n++) {
            if (check(cts_turn(gpu, n), "turn") || result_row(n)) return 1;
            if (n % interval == 0) {
                samples_.insert(samples_.end(), m.begin(), m.end());
                if (record_ions(n / interval)) return 1;
            }
        }
    } else if (nsec == 1) {
        // the turn loop in stretches that end on the print turns; HIP-graph
        // replays inside a stretch
        if (check(cts_set_history(gpu, R->nTurns), "history")) return 1;
        for (int s = 0; s < nsamp; s++) {
            const int t0 = s == 0 ? 0 : (s - 1) * interval + 1, t1 = s * interval;
            if (check(cts_run_turns(gpu, t0, t1 - t0 + 1, 1), "turns")) return 1;
            if (record_ions(s)) return 1;
        }
        if (R->nTurns > (nsamp - 1) * interval + 1) {
            const int t0 = (nsamp - 1) * interval + 1;
            if (check(cts_run_turns(gpu, t0, R->nTurns - t0, 1), "turns")) return 1;
        }
        for (int s = 0; s < nsamp; s++) {
            if (check(cts_get_history(gpu, s * interval, 1, m.data()), "history")) return 1;
            samples_.insert(samples_.end(), m.begin(), m.end());
        }
    } else {
        for (int n = 0; n < R->nTurns; n++) {
            // section by section: the ion phase at point k (on the moments
            // after map k - 1), then map k (src/MPBeam.cpp:1059-1070)
            for (int k = 0; k < nsec; k++) {
                if (c.ions && SSBeamIonEffectOneInteractionPoint(n, k)) return 1;
                if (BeamTransferPerInteractionPointDueToLatticeT(k)) return 1;
            }
            if (in.ringParBasic->skewQuadK != 0 && BeamTransferDueToSkewQuad()) return 1;
            if (BeamMomtumUpdateDueToRFTest() || BeamLongiPosTransferOneTurn() || BeamEnergyLossOneTurn()) return 1;
            if (R->synRadDampingFlag && BeamSynRadDamping(n)) return 1;
            if (R->sRWakeFlag && check(cts_sr_wake(gpu, nullptr), "sr wake")) return 1;
            if (MarkParticleLostInBunch()) return 1;
            if (writeResult && result_row(n)) return 1;
            if (n % interval == 0) {
                if (!writeResult && check(cts_rms(gpu, m.data()), "rms")) return 1;
                samples_.insert(samples_.end(), m.begin(), m.end());
                if (record_ions(n / interval)) return 1;
            }
        }
    }
    if (check(cts_sync(gpu), "sync")) return 1;
    if (tbt.is_open()) {
        std::vector<double> cav(c.bl ? (size_t)nb * in.ringParRf->resNum * 8 : 0);
        for (int s = 0; s < nsamp; s++) {
            if (c.bl && check(cts_get_cavity_info(gpu, s * interval, cav.data()), "cavity info")) return 1;
            PrintTBTBunchAverData(tbt, s * interval, samples_.data() + (size_t)s * nb, ionq.data() + (size_t)s * nb,
                                  c.bl ? cav.data() : nullptr, in, s == 0);
        }
    }
    if (MPBeamRMSCal(0)) return 1;
    // GetCBMGR fits and writes its file at the turn with turn + interval ==
    // nTurns (src/MPBeam.cpp:1651), i.e. when nTurns is a multiple of the
    // print interval
    if (!R->runCBMGR.empty() && R->nTurns % interval == 0 && nsamp >= 2) {
        std::vector<double> rates;
        if (GetCBMGR(in, rates)) return 1;
    }
    return 0;
}

}  // namespace cetasim
