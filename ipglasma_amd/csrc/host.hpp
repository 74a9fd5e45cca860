// host.hpp -- host-side C++ of the solver core: input parsing, MT19937-64,
// nucleus data, nucleon sampling and wounding.  These are the parts of the
// hot path that are sequential by definition (one random stream per event)
// and bit-exact targets (nucleon positions, N_coll, N_part).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ipg.h"

namespace ipg {

// MT19937-64 with the reference's transforms (Random.cpp:67-204)
class MT64 {
   public:
    void seed(uint64_t s);
    uint64_t next();
    double real1();  // [0,1]
    double real2();  // [0,1)
    double real3();  // (0,1)
    double gauss();  // Marsaglia polar, second deviate cached

   private:
    uint64_t mt_[312];
    int mti_ = 313;
    int iset_ = 0;
    double gset_ = 0.;
};

struct Nucleon {
    double x, y, z, phi;
    int collided;
    int proton;
};

struct NucleusData {
    int A = 0, Z = 0;
    double R = 0, a = 0, beta2 = 0, beta3 = 0, beta4 = 0, gamma = 0, dR = 0, da = 0;
    double w = 0;  // the second Hulthen range of the deuteron (w_WS)
    int force_dmin = 0;
    double d_min = 0;
};

// Glauber::FindNucleusData2 (Glauber.cpp:19-287) with the setWSDeformParams
// override; returns false for an unknown name
bool find_nucleus(const std::string& name, const ipg_params& p, NucleusData* out);

struct EventInfo {
    double b = 0, phiRP = 0;
    int Ncoll = 0, Npart = 0;
};

// Init::init's draws up to the colour-charge density: b (Init.cpp:2320-2345),
// sampleTA (65-439), reaction plane and shift (868-901), wounding
// (1117-1216).  Returns IPG_OK or IPG_ERR_ARG for unsupported options.
int sample_event(MT64& rng, const ipg_params& p, const NucleusData& proj, const NucleusData& targ,
                 std::vector<Nucleon>& A, std::vector<Nucleon>& B, EventInfo& info, std::string& err);

// The deuteron's nucleon thickness (Glauber.cpp:513-576, 847-1100): the
// Hulthen density rho(r) = rho0 (e^{-a r} - e^{-w r})^2 / r^2 integrated
// along the beam axis as the reference integrates it (integrand in
// xi = e^{-z a}, adaptive 7-point Newton-Cotes to 1e-6), tabulated at 101
// points on [0, 24] fm and read back by 4-point interpolation; units
// sigma_NN [fm^2] / fm^2 as in the reference.  R0 is the density's starting
// normalisation (R_WS), sigma the nucleon cross section in fm^2.
double deuteron_thickness(double s, double a, double w, double R0, double sigma);
// Glauber::SampleTARejection (Glauber.cpp:1188-1243): proton-neutron
// separation (x, y) by rejection against r e^{-r^2/40} envelopes
void sample_deuteron_separation(MT64& rng, double a, double w, double R0, double sigma, double* x, double* y);

// maxX, maxY of Evolution::eccentricity (Evolution.cpp:1492-1640): over
// 2 <= ix, iy < N-2, the last x = -L/2 + a ix - avx on the row
// iy = N/2 + yshift, and the last y on the column ix = N/2 + xshift, whose
// weight eps u^tau exceeds cutoff (weight 0 where eps < cutoff);
// shift = floor(<x>/a + 1e-11).  eps, utau: [N*N], site ix*N + iy.
void ecc_extents(const double* eps, const double* utau, int N, double L, double avx, double avy, double cutoff,
                 double* maxX, double* maxY);
// one line appended to eccentricities<id>.dat (Evolution.cpp:1797-1809):
// tau, (eps_n, Psi_n) for n = 1..6, cutoff, sqrt<r^2>, maxX, maxY, b, Tpp,
// area, Rbar, <eps>, <Q_sA^2 Q_sB^2> hbarc, as an ostream prints them
int append_eccentricity_line(const char* path, double tau, const double* ecc, const double* psi, double cutoff,
                             double rms_r, double maxX, double maxY, double b, double Tpp, double area, double Rbar,
                             double avg_eps, double avg_qs, std::string& err);

// epsilon-u-Hydro file (MyEigen.cpp:326-414, 440-800; writeOutputs odd,
// gfactor 1): the Landau frame land [15][N*N] (eps, u^tau, u^x, u^y, u^eta,
// pi in T's order) interpolated onto hx x hx x heta hydro cells of size
// hL/hx; *Etot as the reference sums it
int write_hydro(const char* path, const double* land, int N, double L, int it, double dtau, int hx, int heta,
                double hL, double deta, double* Etot, std::string& err);

// Tmunu-t file (MyEigen.cpp:801-1046; writeOutputs / 4 == 1, gfactor 1):
// T [10][N*N] in T's order and physical units, interpolated as for the
// hydro file onto the hx x hx grid (y rows outer, x inner)
int write_tmunu(const char* path, const double* T, int N, double L, int it, double dtau, int hx, int heta, double hL,
                double deta, std::string& err);

// Jazma-Hydro file (MyEigen.cpp:1047-1293; (writeOutputs % 4) / 2 == 1):
// g2mu2A g2mu2B [N*N] interpolated onto the hydro grid, normalised so that
// it integrates to Etot (the hydro file's sum, 0 when that file was not
// written, as in the reference)
int write_jazma(const char* path, const double* g2A, const double* g2B, int N, double L, int it, double dtau, int hx,
                int heta, double hL, double deta, double Etot, std::string& err);

// the NpartdNdy-t<tau>-<id>.dat line (Evolution.cpp:3153-3177), as an
// ostream prints it: Npart dN/dy Tpp b dE/dy seed N/A N/A N/A cut3 cut3
// cut6 cut6 and g^2 / (4 pi 4 pi / (9 log((mu0/0.2)^(2/c) + (f <Q_s>/0.2)^(2/c))^c))
int write_npart_dndy(const char* path, int Npart, double dNdy, double Tpp, double b, double dEdy,
                     unsigned long long seed, double n3, double e3, double n6, double e6, double g, double muZero,
                     double c, double fQs, double avgQs, std::string& err);

// b from the first draw of Init::init (Init.cpp:2320-2347)
double sample_b(MT64& rng, const ipg_params& p);

int read_input(const char* path, ipg_params* p, std::string& err);
void default_params(ipg_params* p);
void derive(ipg_params* p);

constexpr double HBARC = 0.1973269718;  // Phys_consts.h:6

// Wilson-line files.  V is [N*N][18]: site s = i*N + j (i along x), entry
// r*3 + c of the 3x3 matrix as (re, im).  format 1 = text, 2 = binary.
// write_wilson_lines: Init.cpp:1857-1980 (one nucleus per file).
int write_wilson_lines(const char* path, int N, double L, double rapidity, const double* V, int format,
                       std::string& err);
// read_wilson_lines: Init.cpp:1988-2301; nucleus 0 (A, rows shifted by -b/2)
// or 1 (B, +b/2); sites the file does not reach keep what V holds.
int read_wilson_lines(const char* path, int N, double L, double b, int nucleus, int format, double* V,
                      std::string& err);

}  // namespace ipg
