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

// Parameters: the parsed input file (Parameters.h getters for the hot path)
class Parameters {
   public:
    Parameters() { ipg_default_params(&p_); }
    ipg_params& raw() { return p_; }
    const ipg_params& raw() const { return p_; }
    int getSize() const { return p_.size; }
    int getNc() const { return p_.Nc; }
    int getNy() const { return p_.Ny; }
    double getL() const { return p_.L; }
    double getg() const { return p_.g; }
    double getdtau() const { return p_.dtau; }
    double getMaxtime() const { return p_.maxtime; }
    unsigned long long getSeed() const { return p_.seed; }
    int getUseTimeForSeed() const { return p_.useTimeForSeed; }
    std::string getNucleusQsTableFileName() const { return p_.NucleusQsTableFileName; }
    int getReadInitialWilsonLines() const { return p_.readInitialWilsonLines; }
    int getWriteInitialWilsonLines() const { return p_.writeInitialWilsonLines; }
    // set by the driver per event (main.cpp:71,148); they name the
    // Wilson-line files
    int getEventId() const { return event_id_; }
    void setEventId(int id) { event_id_ = id; }
    int getMPISize() const { return mpi_size_; }
    void setMPISize(int n) { mpi_size_ = n; }
    // filled by Init::init (Init.cpp:1614-1644)
    int getSuccess() const { return success_; }
    void setSuccess(int s) { success_ = s; }
    double getb() const { return obs_.b; }
    int getNpart() const { return obs_.Npart; }
    double getAverageQs() const { return obs_.averageQs; }
    double getAverageQsAvg() const { return obs_.averageQsAvg; }
    double getAverageQsmin() const { return obs_.averageQsmin; }
    double getTpp() const { return obs_.Tpp; }
    double getalphas() const { return obs_.alphas; }
    const ipg_observables& observables() const { return obs_; }
    void setObservables(const ipg_observables& o) { obs_ = o; }

   private:
    ipg_params p_;
    int event_id_ = 0, mpi_size_ = 1;
    int success_ = 0;
    ipg_observables obs_{};
};

// Setup: the reference's key/value input reader (Setup.cpp:24-104)
class Setup {
   public:
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
    }
    unsigned long long seed() const { return seed_; }
    bool seeded() const { return seeded_; }

   private:
    unsigned long long seed_ = 5489ULL;
    bool seeded_ = false;
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
        check(ipg_evolve(lat->ctx()));
        ipg_observables o;
        check(ipg_get_observables(lat->ctx(), &o));
        param->setObservables(o);
        const ipg_params& p = param->raw();
        // Evolution.cpp:644-651 at itmax: the eccentricity line (no gluon
        // spectra, no it = 1 line)
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
};

}  // namespace ipglasma
