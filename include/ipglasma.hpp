// ipglasma.hpp -- C++ call surface of the reference (schenke/ipglasma) over
// the C ABI of include/ipg.h.  Header only.
//
// The reference drives one event through (main.cpp:167-360):
//     Setup setup; readInput(&setup, param, ...);          Setup.cpp:24-104
//     Random random; random.init_genrand64(seed);          Random.cpp:67-73
//     Init init(nn); Group group(Nc); Glauber glauber;     Init.h:51
//     glauber.initGlauber(...);                            Glauber.cpp:1135
//     Evolution evolution(nn); Lattice lat(param, Nc, N);  Evolution.h:23
//     while (!param->getSuccess()) {
//         init.init(&lat, &group, param, &random, &glauber, 0);  Init.cpp:2303
//         if (param->getSuccess()) evolution.run(&lat, &group, param);
//     }
// The classes below keep those names, argument order and meaning.  The
// differences are the error convention (an ipg::Error exception instead of
// exit(1)) and that the event state lives on the GPU inside Lattice.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "ipg.h"

namespace ipglasma {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

inline void check(int rc) {
    if (rc != IPG_OK) throw Error(ipg_last_error());
}

// Parameters: the parsed input file and the per-event results, with the
// reference's accessors (Parameters.h:215-493): every setX/getX pair that
// readInput (main.cpp:448-578) and the event loop (main.cpp:129-360) call.
// Keys the hot path reads live in the ipg_params the library takes (raw());
// the others are kept for callers and ignored by the library.  b, Npart, the
// <Q_s> averages, alpha_s, T_pp, the area, eps_2 and Psi_2 are filled by
// Init::init and Evolution::run (setObservables).  setSize, setL and
// setMaxtime re-derive dtau and itmax (main.cpp:531-534).
class Parameters {
   public:
    Parameters() { ipg_default_params(&p_); }
    ipg_params& raw() { return p_; }
    const ipg_params& raw() const { return p_; }
    const ipg_observables& observables() const { return obs_; }
    void setObservables(const ipg_observables& o) { obs_ = o; }
    int getItmax() const { return p_.itmax; }
    // the posterior sub-nucleon parameter sets (Parameters.cpp) belong to the
    // constituent-quark proton, which the library does not support
    void loadPosteriorParameterSets(const int itype) {
        if (itype > 0) throw Error("SubNucleonParamType > 0 (sub-nucleon posterior sets) is not supported");
    }
    void setParamsWithPosteriorParameterSet(const int itype, int iset) {
        (void)iset;
        loadPosteriorParameterSets(itype);
    }

    void setSubNucleonParamType(int x) { p_.SubNucleonParamType = x; }
    int getSubNucleonParamType() const { return p_.SubNucleonParamType; }
    void setSubNucleonParamSet(int x) { subNucleonParamSet_ = x; }
    int getSubNucleonParamSet() const { return subNucleonParamSet_; }
    void setSeed(unsigned long long x) { p_.seed = x; }
    unsigned long long getSeed() const { return p_.seed; }
    void setA(int x) { a_ = x; }
    int getA() const { return a_; }
    void setNc(int x) { p_.Nc = x; }
    int getNc() const { return p_.Nc; }
    void setNy(int x) { p_.Ny = x; }
    int getNy() const { return p_.Ny; }
    void setSize(int x) { p_.size = x; ipg_derive(&p_); }
    int getSize() const { return p_.size; }
    void setProtonAnisotropy(double x) { p_.protonAnisotropy = x; }
    double getProtonAnisotropy() const { return p_.protonAnisotropy; }
    void setSizeOutput(int x) { p_.sizeOutput = x; }
    int getSizeOutput() const { return p_.sizeOutput; }
    void setEtaSizeOutput(int x) { p_.etaSizeOutput = x; }
    int getEtaSizeOutput() const { return p_.etaSizeOutput; }
    void setDetaOutput(double x) { p_.detaOutput = x; }
    double getDetaOutput() const { return p_.detaOutput; }
    void setAverageOverNuclei(int x) { p_.averageOverThisManyNuclei = x; }
    int getAverageOverNuclei() const { return p_.averageOverThisManyNuclei; }
    void setR(double x) { r_ = x; }
    double getR() const { return r_; }
    void setDs(double x) { ds_ = x; }
    double getDs() const { return ds_; }
    void setg2mu(double x) { p_.g2mu = x; }
    double getg2mu() const { return p_.g2mu; }
    void setQs(double x) { qs_ = x; }
    double getQs() const { return qs_; }
    void setSteps(int x) { steps_ = x; }
    int getSteps() const { return steps_; }
    void setMeasureSteps(int x) { measureSteps_ = x; }
    int getMeasureSteps() const { return measureSteps_; }
    void setMode(int x) { mode_ = x; }
    int getMode() const { return mode_; }
    void setRunningCoupling(int x) { p_.runningCoupling = x; }
    int getRunningCoupling() const { return p_.runningCoupling; }
    void setMu0(double x) { mu0_ = x; }
    double getMu0() const { return mu0_; }
    void setg(double x) { p_.g = x; }
    double getg() const { return p_.g; }
    void setLambdaQCD(double x) { lambdaQCD_ = x; }
    double getLambdaQCD() const { return lambdaQCD_; }
    void setkappa4Factor(double x) { kappa4Factor_ = x; }
    double getkappa4Factor() const { return kappa4Factor_; }
    void setSigmaNN(double x) { p_.SigmaNN = x; }
    double getSigmaNN() const { return p_.SigmaNN; }
    void setb(double x) { obs_.b = x; }
    double getb() const { return obs_.b; }
    void setbmin(double x) { p_.bmin = x; }
    double getbmin() const { return p_.bmin; }
    void setbmax(double x) { p_.bmax = x; }
    double getbmax() const { return p_.bmax; }
    void setTarget(std::string x) { copy_name(p_.Target, sizeof(p_.Target), x); }
    std::string getTarget() const { return std::string(p_.Target); }
    void setProjectile(std::string x) { copy_name(p_.Projectile, sizeof(p_.Projectile), x); }
    std::string getProjectile() const { return std::string(p_.Projectile); }
    void setL(double x) { p_.L = x; ipg_derive(&p_); }
    double getL() const { return p_.L; }
    void setLOutput(double x) { p_.LOutput = x; }
    double getLOutput() const { return p_.LOutput; }
    void setm(double x) { p_.m = x; }
    double getm() const { return p_.m; }
    void setJacobianm(double x) { p_.Jacobianm = x; }
    double getJacobianm() const { return p_.Jacobianm; }
    void setQsmuRatio(double x) { p_.QsmuRatio = x; }
    double getQsmuRatio() const { return p_.QsmuRatio; }
    void setQsmuRatioB(double x) { qsmuRatioB_ = x; }
    double getQsmuRatioB() const { return qsmuRatioB_; }
    void setRapidity(double x) { p_.Rapidity = x; }
    double getRapidity() const { return p_.Rapidity; }
    void setMaxtime(double x) { p_.maxtime = x; ipg_derive(&p_); }
    double getMaxtime() const { return p_.maxtime; }
    void setdtau(double x) { p_.dtau = x; }
    double getdtau() const { return p_.dtau; }
    void setNpart(int x) { obs_.Npart = x; }
    int getNpart() const { return obs_.Npart; }
    void setAverageQs(double x) { obs_.averageQs = x; }
    double getAverageQs() const { return obs_.averageQs; }
    void setAverageQsAvg(double x) { obs_.averageQsAvg = x; }
    double getAverageQsAvg() const { return obs_.averageQsAvg; }
    void setAverageQsmin(double x) { obs_.averageQsmin = x; }
    double getAverageQsmin() const { return obs_.averageQsmin; }
    void setalphas(double x) { obs_.alphas = x; }
    double getalphas() const { return obs_.alphas; }
    void setxExponent(double x) { xExponent_ = x; }
    double getxExponent() const { return xExponent_; }
    void setRandomSeed(unsigned long long x) { randomSeed_ = x; }
    unsigned long long getRandomSeed() const { return randomSeed_; }
    void setUseTimeForSeed(int x) { p_.useTimeForSeed = x; }
    int getUseTimeForSeed() const { return p_.useTimeForSeed; }
    void setUseSeedList(int x) { p_.useSeedList = x; }
    int getUseSeedList() const { return p_.useSeedList; }
    void setNucleusQsTableFileName(std::string x) { copy_name(p_.NucleusQsTableFileName, sizeof(p_.NucleusQsTableFileName), x); }
    std::string getNucleusQsTableFileName() const { return std::string(p_.NucleusQsTableFileName); }
    void setA1FromFile(int x) { a1FromFile_ = x; }
    int getA1FromFile() const { return a1FromFile_; }
    void setA2FromFile(int x) { a2FromFile_ = x; }
    int getA2FromFile() const { return a2FromFile_; }
    void setBG(double x) { p_.BG = x; }
    double getBG() const { return p_.BG; }
    void setBGq(double x) { bGq_ = x; }
    double getBGq() const { return bGq_; }
    void setBGqVar(double x) { bGqVar_ = x; }
    double getBGqVar() const { return bGqVar_; }
    void setDqmin(double x) { dqmin_ = x; }
    double getDqmin() const { return dqmin_; }
    void setMuZero(double x) { p_.muZero = x; }
    double getMuZero() const { return p_.muZero; }
    void setc(double x) { p_.c = x; }
    double getc() const { return p_.c; }
    void setRoots(double x) { p_.roots = x; }
    double getRoots() const { return p_.roots; }
    void setUseFluctuatingx(int x) { p_.useFluctuatingx = x; }
    int getUseFluctuatingx() const { return p_.useFluctuatingx; }
    void setRunWithThisFactorTimesQs(double x) { p_.runWithThisFactorTimesQs = x; }
    double getRunWithThisFactorTimesQs() const { return p_.runWithThisFactorTimesQs; }
    void setxFromThisFactorTimesQs(double x) { p_.xFromThisFactorTimesQs = x; }
    double getxFromThisFactorTimesQs() const { return p_.xFromThisFactorTimesQs; }
    void setTpp(double x) { obs_.Tpp = x; }
    double getTpp() const { return obs_.Tpp; }
    void setNu(double x) { nu_ = x; }
    double getNu() const { return nu_; }
    void setUseFixedNpart(int x) { p_.useFixedNpart = x; }
    int getUseFixedNpart() const { return p_.useFixedNpart; }
    void setArea(double x) { obs_.area = x; }
    double getArea() const { return obs_.area; }
    void setEccentricity2(double x) { obs_.eccentricity[1] = x; }
    double getEccentricity2() const { return obs_.eccentricity[1]; }
    void setRnp(double x) { rnp_ = x; }
    double getRnp() const { return rnp_; }
    void setPsi(double x) { obs_.Psi[1] = x; }
    double getPsi() const { return obs_.Psi[1]; }
    void setSmearingWidth(double x) { p_.smearingWidth = x; }
    double getSmearingWidth() const { return p_.smearingWidth; }
    void setMPIRank(int x) { mPIRank_ = x; }
    int getMPIRank() const { return mPIRank_; }
    void setEventId(int x) { eventId_ = x; }
    int getEventId() const { return eventId_; }
    void setMPISize(int x) { mPISize_ = x; }
    int getMPISize() const { return mPISize_; }
    void setSuccess(int x) { success_ = x; }
    int getSuccess() const { return success_; }
    void setRmax(double x) { p_.rmax = x; }
    double getRmax() const { return p_.rmax; }
    void setUVdamp(double x) { p_.UVdamp = x; }
    double getUVdamp() const { return p_.UVdamp; }
    void setSetWSDeformParams(int x) { p_.setWSDeformParams = x != 0; }
    bool getSetWSDeformParams() const { return p_.setWSDeformParams != 0; }
    void setR_WS(double x) { p_.R_WS = x; }
    double getR_WS() const { return p_.R_WS; }
    void setA_WS(double x) { p_.a_WS = x; }
    double getA_WS() const { return p_.a_WS; }
    void setBeta2(double x) { p_.beta2 = x; }
    double getBeta2() const { return p_.beta2; }
    void setBeta3(double x) { p_.beta3 = x; }
    double getBeta3() const { return p_.beta3; }
    void setBeta4(double x) { p_.beta4 = x; }
    double getBeta4() const { return p_.beta4; }
    void setGamma(double x) { p_.gamma = x; }
    double getGamma() const { return p_.gamma; }
    void setDmin(double x) { p_.d_min = x; }
    double getDmin() const { return p_.d_min; }
    void setForceDmin(int x) { p_.force_dmin_flag = x != 0; }
    bool getForceDmin() const { return p_.force_dmin_flag != 0; }
    void setWSdR_np(double x) { p_.dR_np = x; }
    double getWSdR_np() const { return p_.dR_np; }
    void setWSda_np(double x) { p_.da_np = x; }
    double getWSda_np() const { return p_.da_np; }
    void setRotateReactionPlane(int x) { p_.rotateReactionPlane = x != 0; }
    bool getRotateReactionPlane() const { return p_.rotateReactionPlane != 0; }
    void setInitMethod(int x) { initMethod_ = x; }
    int getInitMethod() const { return initMethod_; }
    void setUseNucleus(int x) { p_.useNucleus = x; }
    int getUseNucleus() const { return p_.useNucleus; }
    void setUseGaussian(int x) { useGaussian_ = x; }
    int getUseGaussian() const { return useGaussian_; }
    void setlightNucleusOption(int x) { lightNucleusOption_ = x; }
    int getlightNucleusOption() const { return lightNucleusOption_; }
    void setRunWithQs(int x) { runWithQs_ = x; }
    int getRunWithQs() const { return runWithQs_; }
    void setRunWithkt(int x) { runWithkt_ = x; }
    int getRunWithkt() const { return runWithkt_; }
    void setRunWithLocalQs(int x) { runWithLocalQs_ = x; }
    int getRunWithLocalQs() const { return runWithLocalQs_; }
    void setLinearb(int x) { p_.samplebFromLinearDistribution = x; }
    int getLinearb() const { return p_.samplebFromLinearDistribution; }
    void setWriteOutputs(int x) { p_.writeOutputs = x; }
    int getWriteOutputs() const { return p_.writeOutputs; }
    void setWriteOutputsToHDF5(int x) { writeOutputsToHDF5_ = x; }
    int getWriteOutputsToHDF5() const { return writeOutputsToHDF5_; }
    void setWriteEvolution(int x) { writeEvolution_ = x; }
    int getWriteEvolution() const { return writeEvolution_; }
    void setWriteInitialWilsonLines(int x) { p_.writeInitialWilsonLines = x; }
    int getWriteInitialWilsonLines() const { return p_.writeInitialWilsonLines; }
    void setReadInitialWilsonLines(int x) { p_.readInitialWilsonLines = x; }
    int getReadInitialWilsonLines() const { return p_.readInitialWilsonLines; }
    void setNucleonPositionsFromFile(int x) { p_.nucleonPositionsFromFile = x; }
    int getNucleonPositionsFromFile() const { return p_.nucleonPositionsFromFile; }
    void setInverseQsForMaxTime(int x) { p_.inverseQsForMaxTime = x; }
    int getInverseQsForMaxTime() const { return p_.inverseQsForMaxTime; }
    void setUseFatTails(int x) { useFatTails_ = x; }
    int getUseFatTails() const { return useFatTails_; }
    void setSmearQs(int x) { p_.smearQs = x; }
    int getSmearQs() const { return p_.smearQs; }
    void setReadMultFromFile(int x) { readMultFromFile_ = x; }
    int getReadMultFromFile() const { return readMultFromFile_; }
    void setGaussianWounding(int x) { p_.gaussianWounding = x; }
    int getGaussianWounding() const { return p_.gaussianWounding; }
    void setUsePseudoRapidity(int x) { p_.usePseudoRapidity = x; }
    int getUsePseudoRapidity() const { return p_.usePseudoRapidity; }
    void setUseConstituentQuarkProton(int x) { p_.useConstituentQuarkProton = x; }
    int getUseConstituentQuarkProton() const { return p_.useConstituentQuarkProton; }
    void setNqBase(double x) { nqBase_ = x; }
    double getNqBase() const { return nqBase_; }
    void setNqFluc(double x) { nqFluc_ = x; }
    double getNqFluc() const { return nqFluc_; }
    void setUseSmoothNucleus(int x) { p_.useSmoothNucleus = x; }
    int getUseSmoothNucleus() const { return p_.useSmoothNucleus; }
    void setShiftConstituentQuarkProtonOrigin(int x) { shiftConstituentQuarkProtonOrigin_ = x; }
    int getShiftConstituentQuarkProtonOrigin() const { return shiftConstituentQuarkProtonOrigin_; }
    void setMinimumQs2ST(int x) { p_.minimumQs2ST = x; }
    int getMinimumQs2ST() const { return static_cast<int>(p_.minimumQs2ST); }
    void setComputeGluonMultiplicity(int x) { p_.computeGluonMultiplicity = x != 0; }
    bool getComputeGluonMultiplicity() const { return p_.computeGluonMultiplicity != 0; }

   private:
    static void copy_name(char* dst, size_t n, const std::string& v) {
        if (v.size() + 1 > n) throw Error("name too long: " + v);
        std::memcpy(dst, v.c_str(), v.size() + 1);
    }
    ipg_params p_;
    ipg_observables obs_{};
    int subNucleonParamSet_ = 0;
    int a_ = 0;
    double r_ = 0.;
    double ds_ = 0.;
    double qs_ = 0.;
    int steps_ = 0;
    int measureSteps_ = 0;
    int mode_ = 0;
    double mu0_ = 0.;
    double lambdaQCD_ = 0.;
    double kappa4Factor_ = 0.;
    double qsmuRatioB_ = 0.;
    double xExponent_ = 0.;
    unsigned long long randomSeed_ = 0;
    int a1FromFile_ = 0;
    int a2FromFile_ = 0;
    double bGq_ = 0.;
    double bGqVar_ = 0.;
    double dqmin_ = 0.;
    double nu_ = 0.;
    double rnp_ = 0.;
    int mPIRank_ = 0;
    int eventId_ = 0;
    int mPISize_ = 1;
    int success_ = 0;
    int initMethod_ = 0;
    int useGaussian_ = 0;
    int lightNucleusOption_ = 0;
    int runWithQs_ = 0;
    int runWithkt_ = 0;
    int runWithLocalQs_ = 0;
    int writeOutputsToHDF5_ = 0;
    int writeEvolution_ = 0;
    int useFatTails_ = 0;
    int readMultFromFile_ = 0;
    double nqBase_ = 0.;
    double nqFluc_ = 0.;
    int shiftConstituentQuarkProtonOrigin_ = 0;
};

// Setup: the reference's key/value input reader (Setup.cpp:24-104).  A file
// is whitespace-separated "key value" pairs up to the word EndOfFile.
// StringFind returns the value after the first occurrence of a key; DFind
// converts with atof, IFind truncates (static_cast<int>), ULLIFind rounds
// (+0.5).  A missing file or key throws ipglasma::Error where the reference
// prints and calls exit(1).
class Setup {
   public:
    Setup() {}
    std::string StringFind(std::string file_name, std::string st) {
        std::ifstream in(file_name);
        if (!in) throw Error("The input file named " + file_name + " is absent.");
        std::string key, value;
        while (in >> key && key != "EndOfFile") {
            if (!(in >> value)) break;
            if (key == st) return value;
        }
        throw Error(st + " not found in " + file_name);
    }
    double DFind(std::string file_name, std::string st) {
        return std::atof(StringFind(file_name, st).c_str());
    }
    int IFind(std::string file_name, std::string st) { return static_cast<int>(DFind(file_name, st)); }
    unsigned long long int ULLIFind(std::string file_name, std::string st) {
        return static_cast<unsigned long long int>(DFind(file_name, st) + 0.5);
    }
    int IsFile(std::string file_name) {
        std::ifstream in(file_name);
        return in.good() ? 1 : 0;
    }
    // the whole file at once (ipg_read_input: keys absent from the file keep
    // their current values instead of being an error)
    void read(const std::string& file_name, Parameters* param) {
        check(ipg_read_input(file_name.c_str(), &param->raw()));
    }
};

// Random: the event's MT19937-64 stream (Random.cpp:67-73); the generator
// itself runs inside the library so that the draw order of Init::init is
// kept exactly
class Random {
   public:
    void init_genrand64(unsigned long long seed) {
        seed_ = seed;
        seeded_ = true;
        host_.seed(seed);
    }
    // the GSL generator of the reference feeds only samplers outside the
    // library's path (fat tails, Poisson smearing)
    void gslRandomInit(unsigned long long seed) { (void)seed; }
    // draws from a host copy of the stream; the library's Init::init keeps
    // its own stream from the same seed, so draws taken here do not shift it
    // (in the reference they would)
    unsigned long long genrand64_int64() { return host_(); }
    long long genrand64_int63() { return (long long)(host_() >> 1); }
    unsigned long long seed() const { return seed_; }
    bool seeded() const { return seeded_; }

   private:
    unsigned long long seed_ = 5489ULL;
    bool seeded_ = false;
    std::mt19937_64 host_{5489ULL};  // the MT19937-64 of Random.cpp:67-120
};

// Group and Glauber keep their places in the call sequence; SU(3) and the
// nucleus table are part of the library
class Group {
   public:
    explicit Group(int Nc) {
        if (Nc != 3) throw Error("only Nc = 3 is supported");
    }
};
class Glauber {
   public:
    void initGlauber(double, std::string, std::string, double, bool, double, double, double, double, double,
                     double, bool, double, double, double, int) {}
};

// Lattice: owns the device state of one event (Lattice.cpp:4-30)
class Lattice {
   public:
    Lattice(Parameters* param, int Nc, int N) {
        (void)Nc;
        n_ = N;
        param->raw().size = N;
        ctx_ = ipg_create(&param->raw());
        if (!ctx_) throw Error(ipg_last_error());
        check(ipg_load_qs_table(ctx_, param->raw().NucleusQsTableFileName));
    }
    ~Lattice() { ipg_destroy(ctx_); }
    Lattice(const Lattice&) = delete;
    Lattice& operator=(const Lattice&) = delete;
    ipg_ctx* ctx() { return ctx_; }
    int N() const { return n_; }
    // cell fields as [S][18] (3x3 complex row-major) or [S]
    std::vector<double> field(int which, size_t S) {
        std::vector<double> v(which < IPG_F_G2MU2A ? 18 * S : S);
        check(ipg_get_field(ctx_, which, v.data()));
        return v;
    }
    std::vector<double> tmunu(size_t S) {
        std::vector<double> v(10 * S);
        check(ipg_get_tmunu(ctx_, v.data()));
        return v;
    }
    // Cell epsilon, utau, ux, uy, ueta, pi^{mu nu} after Evolution::run
    // (MyEigen.cpp:25-321), [15][S]
    std::vector<double> landau(size_t S) {
        std::vector<double> v(15 * S);
        check(ipg_get_landau(ctx_, v.data()));
        return v;
    }
    bool seeded = false;

   private:
    ipg_ctx* ctx_ = nullptr;
    int n_ = 0;
};

// Init: Init::init (Init.cpp:2303-2696)
class Init {
   public:
    explicit Init(const int nn[]) { (void)nn; }
    void init(Lattice* lat, Group* group, Parameters* param, Random* random, Glauber* glauber,
              int READFROMFILE) {
        (void)group;
        (void)glauber;
        if (!lat->seeded) {
            check(ipg_seed(lat->ctx(), random->seed()));
            lat->seeded = true;
        }
        int ok = 0;
        if (READFROMFILE > 0) {
            readV(lat, param, READFROMFILE);
            ok = 1;
        } else if (param->getWriteInitialWilsonLines() > 0) {
            // setV writes V_A, V_B before the links (Init.cpp:1857-1980)
            check(ipg_init_nucleons(lat->ctx()));
            check(ipg_init_color_charge(lat->ctx(), &ok));
            if (ok) {
                check(ipg_init_wilson_lines(lat->ctx()));
                const int f = param->getWriteInitialWilsonLines();
                check(ipg_write_wilson_lines(lat->ctx(), vname(param, 0, f).c_str(), vname(param, 1, f).c_str(), f));
                check(ipg_init_fields(lat->ctx()));
            }
        } else {
            check(ipg_init(lat->ctx(), &ok));
        }
        ipg_observables o;
        check(ipg_get_observables(lat->ctx(), &o));
        param->setObservables(o);
        param->setSuccess(ok);
    }
    // Init::readV (Init.cpp:1988-2301) with the b draw of Init::init before it
    void readV(Lattice* lat, Parameters* param, int format) {
        check(ipg_read_wilson_lines(lat->ctx(), vname(param, 0, format).c_str(), vname(param, 1, format).c_str(),
                                    format));
        check(ipg_init_fields(lat->ctx()));
    }
    // V-<eventId + 2 seed mpiSize> (A), V-<eventId + (1 + 2 seed) mpiSize> (B),
    // ".txt" for text (Init.cpp:1863-1877)
    static std::string vname(const Parameters* param, int nucleus, int format) {
        const unsigned long long n = (unsigned long long)param->getEventId() +
                                     (nucleus + 2ULL * param->getSeed()) * (unsigned long long)param->getMPISize();
        return "V-" + std::to_string(n) + (format == 1 ? ".txt" : "");
    }
};

// Evolution: Evolution::run (Evolution.cpp:366-657)
class Evolution {
   public:
    explicit Evolution(const int nn[]) { (void)nn; }
    void run(Lattice* lat, Group* group, Parameters* param) {
        (void)group;
        const ipg_params& p = param->raw();
        if (p.computeGluonMultiplicity == 1) {
            // Evolution.cpp:366-657 step by step: after the updates of steps 1
            // and itmax the gauge fix and the spectra (Evolution.cpp:644-651);
            // NpartdNdy-t<tau>-<id>.dat at itmax; a step whose dN/dy is zero
            // ends the evolution as the reference's success = 0 does
            for (int it = 0; it <= p.itmax; it++) {
                check(ipg_evolve_step(lat->ctx(), it));
                if (it >= 1 && (it == 1 || it == p.itmax)) {
                    check(ipg_gluon_multiplicity(lat->ctx(), it, &mult_));
                    if (mult_.dNdy == 0.) {
                        param->setSuccess(0);
                        return;
                    }
                    if (it == p.itmax) {
                        char name[96];
                        std::snprintf(name, sizeof name, "NpartdNdy-t%g-%d.dat", it * p.dtau * (p.L / p.size),
                                      param->getEventId());
                        check(ipg_write_multiplicity(lat->ctx(), name, &mult_));
                    }
                }
            }
        } else {
            check(ipg_evolve(lat->ctx()));
        }
        ipg_observables o;
        check(ipg_get_observables(lat->ctx(), &o));
        param->setObservables(o);
        // the eccentricity line at itmax (the it = 1 line reads a Landau
        // frame that does not exist before itmax and is not written)
        if (p.computeGluonMultiplicity == 1)
            check(ipg_write_eccentricities(
                lat->ctx(), ("eccentricities" + std::to_string(param->getEventId()) + ".dat").c_str(), p.itmax));
        if (p.writeOutputs <= 0) return;
        // MyEigen::flowVelocity4D's files at itmax (MyEigen.cpp:326-1293)
        char tau[32];
        std::snprintf(tau, sizeof tau, "%g", p.itmax * p.dtau * (p.L / p.size));
        const std::string id = std::to_string(param->getEventId());
        double etot = 0.;
        if (p.writeOutputs % 2 == 1)
            check(ipg_write_hydro(lat->ctx(), ("epsilon-u-Hydro-TauHydro-" + id + ".dat").c_str(), p.itmax,
                                  p.sizeOutput, p.etaSizeOutput, p.LOutput, p.detaOutput, &etot));
        if (p.writeOutputs / 4 == 1)
            check(ipg_write_tmunu(lat->ctx(), ("Tmunu-t" + std::string(tau) + "-" + id + ".dat").c_str(), p.itmax,
                                  p.sizeOutput, p.etaSizeOutput, p.LOutput, p.detaOutput));
        if ((p.writeOutputs % 4) / 2 == 1)
            check(ipg_write_jazma(lat->ctx(), ("Jazma-Hydro-t" + std::string(tau) + "-" + id + ".dat").c_str(),
                                  p.itmax, p.sizeOutput, p.etaSizeOutput, p.LOutput, p.detaOutput, etot));
    }
    // the spectra of the last multiplicity step (dN/dy, dE/dy, k_T bins)
    const ipg_multiplicity& lastMultiplicity() const { return mult_; }
    // In place of run(): the loss at tau_max (IPG_LOSS_DEDETA, IPG_LOSS_ECC2,
    // or IPG_LOSS_TMUNU with ten weights) and its gradient with respect to
    // rho ([2][Ny][S][8]), the tau = 0 fields ([6][S][18]) or T^{mu nu} at
    // itmax ([10][S]).  No counterpart in the reference.
    double adjoint(Lattice* lat, int loss, int wrt, std::vector<double>& grad, size_t S, int Ny,
                   const double* weights = nullptr) {
        grad.assign(wrt == IPG_WRT_RHO ? 2 * (size_t)Ny * S * 8 : (wrt == IPG_WRT_TMUNU ? 10 * S : 6 * S * 18), 0.);
        double v = 0.;
        check(ipg_adjoint_loss(lat->ctx(), loss, weights, wrt, &v, grad.data()));
        return v;
    }

   private:
    ipg_multiplicity mult_{};
};

}  // namespace ipglasma
