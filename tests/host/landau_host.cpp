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

// cell values (eps, u^tau) of one site and the reverse mode of
// eps_b eps + ut_b u^tau
extern "C" void ipgl_landau_cell_host(const double* T, int N, double a, double dtau, int it, int si, int sj,
                                      double* eps, double* ut) {
    double t[10];
    for (int k = 0; k < 10; k++) t[k] = T[k];
    ipgl::LandauSite o;
    ipgl::landau_site(t, N, a, dtau, it, si, sj, o);
    *eps = o.eps;
    *ut = o.u[0];
}
extern "C" void ipgl_landau_adj_host(const double* T, int N, double a, double dtau, int it, int si, int sj,
                                     double eps_b, double ut_b, double* Tb) {
    double t[10], tb[10] = {0.};
    for (int k = 0; k < 10; k++) t[k] = T[k];
    ipgl::landau_site_adj(t, N, a, dtau, it, si, sj, eps_b, ut_b, tb);
    for (int k = 0; k < 10; k++) Tb[k] = tb[k];
}

// d eps_2 / d eps and d eps_2 / d u^tau at every site (the sums of the
// reverse kernels done serially; weights eps u^tau where eps >= cutoff)
extern "C" void ipgl_ecc2_adj_host(int N, double L, double cutoff, const double* eps, const double* ut,
                                   double* eps_b, double* ut_b) {
    const double a = L / N;
    const long long S = (long long)N * N;
    double W = 0., X = 0., Y = 0.;
    for (long long s = 0; s < S; s++) {
        const int i = (int)(s / N), j = (int)(s % N);
        const double w = eps[s] < cutoff ? 0. : eps[s] * ut[s];
        W += w;
        X += (-L / 2. + a * i) * w;
        Y += (-L / 2. + a * j) * w;
    }
    const double xb = X / W, yb = Y / W;
    double q[5] = {0., 0., 0., 0., 0.};
    for (long long s = 0; s < S; s++) {
        const int i = (int)(s / N), j = (int)(s % N);
        if (i < 2 || i >= N - 2 || j < 2 || j >= N - 2) continue;
        const double w = eps[s] < cutoff ? 0. : eps[s] * ut[s];
        const double dx = -L / 2. + a * i - xb, dy = -L / 2. + a * j - yb;
        q[0] += (dx * dx - dy * dy) * w;
        q[1] += 2. * dx * dy * w;
        q[2] += (dx * dx + dy * dy) * w;
        q[3] += dx * w;
        q[4] += dy * w;
    }
    for (long long s = 0; s < S; s++) {
        const int i = (int)(s / N), j = (int)(s % N);
        const double dx = -L / 2. + a * i - xb, dy = -L / 2. + a * j - yb;
        const bool in = i >= 2 && i < N - 2 && j >= 2 && j < N - 2;
        const double wb = ipgl::ecc2_wbar(q[0], q[1], q[2], q[3], q[4], W, dx, dy, in, 1.);
        eps_b[s] = eps[s] < cutoff ? 0. : wb * ut[s];
        ut_b[s] = eps[s] < cutoff ? 0. : wb * eps[s];
    }
}
