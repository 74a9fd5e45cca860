// Host build of ipglasma_amd/csrc/landau.hpp (the code the Landau kernels run)
// for CPU comparison with the oracle.  The per-site routine and the carried
// (eps, u^mu) scan are composed here in the same way as k_landau and
// k_landau_carry compose them on the device.
#include <vector>

#include "../../ipglasma_amd/csrc/landau.hpp"

extern "C" int ipgl_eig4_host(const double* A, double* wr, double* wi, double* vr, double* vi) {
    double M[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) M[i][j] = A[4 * i + j];
    ipgl::Eig4 r;
    if (!ipgl::eig4(M, r)) return 1;
    for (int i = 0; i < 4; i++) {
        wr[i] = r.wr[i];
        wi[i] = r.wi[i];
        for (int j = 0; j < 4; j++) {
            vr[4 * i + j] = r.vr[i][j];
            vi[4 * i + j] = r.vi[i][j];
        }
    }
    return 0;
}

extern "C" void ipgl_landau_host(int N, double L, double dtau, int it, const double* T, double* out, double* avg,
                                 int* nfail) {
    const long long S = (long long)N * N;
    const double a = L / N;
    std::vector<double> epsv(S);
    std::vector<int> flag(S);
    avg[0] = avg[1] = avg[2] = avg[3] = 0.;
    *nfail = 0;
    for (long long s = 0; s < S; s++) {
        double t[10];
        for (int k = 0; k < 10; k++) t[k] = T[k * S + s];
        ipgl::LandauSite o;
        ipgl::landau_site(t, N, a, dtau, it, (int)(s / N), (int)(s % N), o);
        out[s] = o.eps;
        for (int k = 0; k < 4; k++) out[(1 + k) * S + s] = o.u[k];
        epsv[s] = o.eps_var;
        flag[s] = o.flags;
        for (int k = 0; k < 4; k++) avg[k] += o.acc[k];
        if (o.flags & 4) (*nfail)++;
    }
    long long re = -1, ru = -1;
    for (long long s = 0; s < S; s++) {
        if (flag[s] & 1) re = s;
        if (flag[s] & 2) ru = s;
        double t[10], pi[10], u[4] = {0., 0., 0., 0.};
        for (int k = 0; k < 10; k++) t[k] = T[k * S + s];
        if (ru >= 0)
            for (int k = 0; k < 4; k++) u[k] = out[(1 + k) * S + ru];
        ipgl::landau_pi(t, re >= 0 ? epsv[re] : 0., u[0], u[1], u[2], u[3], it, dtau, a, pi);
        for (int k = 0; k < 10; k++) out[(5 + k) * S + s] = pi[k];
    }
}
