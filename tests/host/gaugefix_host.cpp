// Host build of ipglasma_amd/csrc/gaugefix.hpp (the per-site code of the
// gauge-fixing kernels) for CPU comparison with the oracle.
#include "../../ipglasma_amd/csrc/gaugefix.hpp"

static M3 load18(const double* p) {
    M3 m;
    for (int k = 0; k < 9; k++) m.e[k] = cd{p[2 * k], p[2 * k + 1]};
    return m;
}
static void store18(double* p, const M3& m) {
    for (int k = 0; k < 9; k++) {
        p[2 * k] = m.e[k].r;
        p[2 * k + 1] = m.e[k].i;
    }
}

extern "C" int gfh_expm(const double* A, double* out) {
    bool ok = true;
    M3 r = ipggf::expm_pade6(load18(A), &ok);
    store18(out, r);
    return ok ? 0 : 1;
}

extern "C" void gfh_reunitarize(const double* M, double* out) { store18(out, ipggf::reunitarize(load18(M))); }

extern "C" void gfh_rotation(const double* chi, double* out) {
    double c[8];
    for (int a = 0; a < 8; a++) c[a] = chi[a];
    store18(out, ipggf::gauge_rotation(c));
}
