// ipglasma_main.cpp -- drop-in driver: `ipglasma [inputfile] [nev]`.
//
// Reproduces the control flow of the reference driver (main.cpp:47-360):
// input file (default "input"), seed = seed (+1000*rank, rank 0 here), event
// loop with the success/retry loop around Init::init, then Evolution::run.
// Each event prints one line of observables instead of the reference's
// per-event files.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../include/ipg.h"

int main(int argc, char** argv) {
    const char* file = argc > 1 ? argv[1] : "input";
    const int nev = argc == 3 ? std::atoi(argv[2]) : 1;
    ipg_params p;
    ipg_default_params(&p);
    if (ipg_read_input(file, &p) != IPG_OK) {
        std::fprintf(stderr, "%s\n", ipg_last_error());
        return 1;
    }
    unsigned long long seed = p.seed;
    if (p.useSeedList == 0 && p.useTimeForSeed == 1) {
        std::random_device rd;
        seed = rd();
    }
    std::printf("Random seed = %llu\n", seed);
    ipg_ctx* c = ipg_create(&p);
    if (!c) {
        std::fprintf(stderr, "%s\n", ipg_last_error());
        return 1;
    }
    if (ipg_load_qs_table(c, p.NucleusQsTableFileName) != IPG_OK || ipg_seed(c, seed) != IPG_OK) {
        std::fprintf(stderr, "%s\n", ipg_last_error());
        ipg_destroy(c);
        return 1;
    }
    const int readV = p.readInitialWilsonLines, writeV = p.writeInitialWilsonLines;
    for (int iev = 0; iev < nev; iev++) {
        // Wilson-line files (Init.cpp:1863-1877): event id iev, one process
        const char* ext = (readV > 0 ? readV : writeV) == 1 ? ".txt" : "";
        char fa[64], fb[64];
        std::snprintf(fa, sizeof fa, "V-%llu%s", (unsigned long long)iev + 2ULL * seed, ext);
        std::snprintf(fb, sizeof fb, "V-%llu%s", (unsigned long long)iev + 1ULL + 2ULL * seed, ext);
        int ok = 0, rc = IPG_OK;
        while (!ok) {
            if (readV > 0) {  // Init::init with READFROMFILE (Init.cpp:2367-2369)
                rc = ipg_read_wilson_lines(c, fa, fb, readV);
                if (rc == IPG_OK) rc = ipg_init_fields(c);
                ok = 1;
            } else if (writeV > 0) {  // setV's output before the links
                rc = ipg_init_nucleons(c);
                if (rc == IPG_OK) rc = ipg_init_color_charge(c, &ok);
                if (rc == IPG_OK && ok) rc = ipg_init_wilson_lines(c);
                if (rc == IPG_OK && ok) rc = ipg_write_wilson_lines(c, fa, fb, writeV);
                if (rc == IPG_OK && ok) {
                    std::printf("wrote %s and %s\n", fa, fb);
                    rc = ipg_init_fields(c);
                }
            } else {
                rc = ipg_init(c, &ok);
            }
            if (rc != IPG_OK) break;
            if (!ok) std::printf("No collision happened. Restarting with new random number...\n");
        }
        if (rc == IPG_OK && p.computeGluonMultiplicity == 1) {
            // Evolution::run step by step with the gauge fix and the spectra
            // after steps 1 and itmax (Evolution.cpp:644-651, 2284-3188)
            ipg_multiplicity m;
            for (int it = 0; rc == IPG_OK && it <= p.itmax; it++) {
                rc = ipg_evolve_step(c, it);
                if (rc != IPG_OK || it < 1 || (it != 1 && it != p.itmax)) continue;
                rc = ipg_gluon_multiplicity(c, it, &m);
                if (rc != IPG_OK) break;
                std::printf("tau = %g fm: %d gauge-fixing steps (residual %g)\n", m.tau_fm, m.gauge_iterations,
                            m.gauge_residual);
                std::printf("dN/dy 1 = %g, dE/dy 1 = %g\n", m.dNdy, m.dEdy);
                std::printf("dN/dy 2 = %g, dE/dy 2 = %g\n", m.dNdy_kt, m.dEdy_kt);
                std::printf("gluon <p_T> = %g\n", m.mean_pt);
                if (m.dNdy == 0.) {
                    std::printf("No collision happened. Stopping the evolution.\n");
                    break;
                }
                if (it == p.itmax) {
                    char nname[96];
                    std::snprintf(nname, sizeof nname, "NpartdNdy-t%g-%d.dat", m.tau_fm, iev);
                    rc = ipg_write_multiplicity(c, nname, &m);
                }
            }
        } else if (rc == IPG_OK) {
            rc = ipg_evolve(c);
        }
        if (rc != IPG_OK) {
            std::fprintf(stderr, "event %d: %s\n", iev, ipg_last_error());
            ipg_destroy(c);
            return 1;
        }
        if (p.computeGluonMultiplicity == 1) {
            // the eccentricity line at itmax (the it = 1 line reads a Landau
            // frame that does not exist before itmax and is not written)
            char ename[64];
            std::snprintf(ename, sizeof ename, "eccentricities%d.dat", iev);
            rc = ipg_write_eccentricities(c, ename, p.itmax);
            if (rc != IPG_OK) {
                std::fprintf(stderr, "event %d: %s\n", iev, ipg_last_error());
                ipg_destroy(c);
                return 1;
            }
        }
        if (p.writeOutputs > 0) {
            // MyEigen::flowVelocity4D's files at itmax (finalFlag set,
            // MyEigen.cpp:326-1293; Evolution.cpp:406-409)
            const int it = p.itmax;
            const double tau = it * p.dtau * (p.L / p.size);
            char name[160];
            double etot = 0.;
            if (rc == IPG_OK && p.writeOutputs % 2 == 1) {
                std::snprintf(name, sizeof name, "epsilon-u-Hydro-TauHydro-%d.dat", iev);
                rc = ipg_write_hydro(c, name, it, p.sizeOutput, p.etaSizeOutput, p.LOutput, p.detaOutput, &etot);
                if (rc == IPG_OK) std::printf("Etot = %g GeV \n", etot);
            }
            if (rc == IPG_OK && p.writeOutputs / 4 == 1) {
                std::snprintf(name, sizeof name, "Tmunu-t%g-%d.dat", tau, iev);
                rc = ipg_write_tmunu(c, name, it, p.sizeOutput, p.etaSizeOutput, p.LOutput, p.detaOutput);
            }
            if (rc == IPG_OK && (p.writeOutputs % 4) / 2 == 1) {
                std::snprintf(name, sizeof name, "Jazma-Hydro-t%g-%d.dat", tau, iev);
                rc = ipg_write_jazma(c, name, it, p.sizeOutput, p.etaSizeOutput, p.LOutput, p.detaOutput, etot);
            }
            if (rc != IPG_OK) {
                std::fprintf(stderr, "event %d: %s\n", iev, ipg_last_error());
                ipg_destroy(c);
                return 1;
            }
        }
        ipg_observables o;
        ipg_get_observables(c, &o);
        std::printf("event %d: b=%.6g fm Npart=%d Ncoll=%d <Qs>=%.6g GeV tau=%.6g fm dE/deta=%.12g GeV "
                    "Gauss violation=%.6g\n",
                    iev, o.b, o.Npart, o.Ncoll, o.averageQs, o.tau_fm, o.dEdeta, o.gauss_max);
    }
    ipg_destroy(c);
    return 0;
}
