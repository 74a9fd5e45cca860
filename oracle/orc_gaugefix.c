/* orc_gaugefix.c -- oracle Coulomb gauge fixing and gluon spectra.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   orc_expm_pade     Matrix::expm (Matrix.cpp:431-560): order-6 diagonal
 *                     Pade approximant with the reference's rounded
 *                     coefficients, scaling and squaring, the 3x3 inverse by
 *                     cofactors
 *   orc_reunitarize   Matrix::reu (Matrix.h:67-90) with Spinor::normalize,
 *                     GramSchmidt and operator% (Spinor.cpp:52-60, 128-137)
 *   orc_fftn_shifted  FFT::fftn (FFT.cpp:260-372) on [S][18] matrix fields:
 *                     per entry, quadrant swap in, DFT, 1/N^2 on the inverse,
 *                     quadrant swap out
 *   orc_gauge_fix     GaugeFix::FFTChi + gaugeTransform (GaugeFix.cpp:17-210)
 *                     on full 3x3 matrices in the reference's order: chi as
 *                     a matrix, the shifted momenta, the site-by-site
 *                     transform of the links
 *   orc_multiplicity  Evolution::multiplicity (Evolution.cpp:2284-3188) for
 *                     runningCoupling 0, runWithkt 0, usePseudoRapidity 0:
 *                     three passes (E1, E2, pi) over shifted momenta with
 *                     the partner (N-i, N-j), the 100-bin k_T histogram
 *
 * This formulation (shifted transforms of the nine matrix entries) is a
 * different computation from the library's (unshifted transforms of packed
 * real planes), so agreement between the two also checks that the
 * momentum relabelling is right.
 *
 * Pins (tests/test_oracle_gaugefix.py): expm against scipy.linalg.expm;
 * reunitarize against a QR factorisation and det = 1; the shifted FFT
 * against numpy's fftshift(fft2(ifftshift(.))); one linearised gauge-fixing
 * step divides div A by 4 (residual by 16), the closed form of the
 * -1.5 / k^2 filter; gauge-invariant traces unchanged site by site;
 * single-mode E and pi fields against the closed-form multiplicity.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ipg_oracle.h"
#include "orc_internal.h"

/* Matrix::expm (Matrix.cpp:431-560), t = 1, p = 6 */
static int expm_pade6(M3 A, M3 *out) {
    static const double c[7] = {1., 0.5, 0.1136363636, 0.01515151515, 0.001262626263, 6.313131313e-05,
                                1.503126503e-06};
    const M3 Id = m_one();
    double norm = 0.0;
    for (int i = 0; i < 3; ++i) {
        double temp = 0.0;
        for (int j = 0; j < 3; j++) temp += cabs(A.e[3 * i + j]);
        norm = fmax(norm, temp);
    }
    if (norm == 0.0) {
        for (int k = 0; k < 9; k++)
            if (A.e[k] != 0.0) return 1; /* NaN or infinity */
        *out = Id;
        return 0;
    }
    int s = 0;
    M3 U = A;
    if (norm > 0.5) {
        s = (int)(log(norm) / log(2.0) + 2.0);
        if (s < 0) s = 0;
        U = m_scale(1.0 / pow(2.0, s), A);
    }
    const M3 H2 = m_mul(U, U);
    M3 Q = m_scale(c[6], Id), P = m_scale(c[5], Id);
    int odd = 1;
    for (int k = 5; k > 0; --k) {
        if (odd == 1)
            Q = m_add(m_mul(Q, H2), m_scale(c[k - 1], Id));
        else
            P = m_add(m_mul(P, H2), m_scale(c[k - 1], Id));
        odd = 1 - odd;
    }
    if (odd == 1)
        Q = m_mul(Q, U);
    else
        P = m_mul(P, U);
    Q = m_sub(Q, P);
    /* cofactor inverse of Q, row by row as Matrix.cpp:530-541 lays it out */
    const cplx *q = Q.e;
    M3 Qi;
    Qi.e[0] = q[4] * q[8] - q[5] * q[7];
    Qi.e[1] = q[2] * q[7] - q[1] * q[8];
    Qi.e[2] = q[1] * q[5] - q[2] * q[4];
    Qi.e[3] = q[5] * q[6] - q[3] * q[8];
    Qi.e[4] = q[0] * q[8] - q[2] * q[6];
    Qi.e[5] = q[2] * q[3] - q[0] * q[5];
    Qi.e[6] = q[3] * q[7] - q[4] * q[6];
    Qi.e[7] = q[1] * q[6] - q[0] * q[7];
    Qi.e[8] = q[0] * q[4] - q[1] * q[3];
    const cplx det = q[0] * q[4] * q[8] + q[1] * q[5] * q[6] + q[2] * q[3] * q[7] - q[2] * q[4] * q[6] -
                     q[1] * q[3] * q[8] - q[5] * q[7] * q[0];
    const cplx inv = 1.0 / det;
    for (int k = 0; k < 9; k++) Qi.e[k] *= inv;
    M3 R = m_mul(m_scale(2.0, Qi), P);
    for (int k = 0; k < 9; k += 4) R.e[k] += 1.0;
    if (odd == 1) R = m_scale(-1.0, R);
    for (int i = 0; i < s; ++i) R = m_mul(R, R);
    *out = R;
    return 0;
}

static void sp_normalize(cplx v[3]) {
    const double n = sqrt(creal(v[0] * conj(v[0]) + v[1] * conj(v[1]) + v[2] * conj(v[2])));
    for (int k = 0; k < 3; k++) v[k] = v[k] / n;
}

/* Matrix::reu (Matrix.h:67-90) */
static M3 reu(M3 m) {
    cplx e1[3] = {m.e[0], m.e[1], m.e[2]};
    cplx a2[3] = {m.e[3], m.e[4], m.e[5]};
    sp_normalize(e1);
    /* Spinor::GramSchmidt: c = a2 - <a1n, a2> a1n with a1n = e1 normalised
     * again, <a, b> = sum conj(a_i) b_i (Spinor.cpp:54-71) */
    cplx u[3] = {e1[0], e1[1], e1[2]};
    sp_normalize(u);
    const cplx p = conj(u[0]) * a2[0] + conj(u[1]) * a2[1] + conj(u[2]) * a2[2];
    cplx e2[3];
    for (int k = 0; k < 3; k++) e2[k] = a2[k] - p * u[k];
    sp_normalize(e2);
    /* operator% (Spinor.cpp:128-137): conjugated cross product */
    cplx e3[3] = {conj(e1[1] * e2[2] - e1[2] * e2[1]), conj(e1[2] * e2[0] - e1[0] * e2[2]),
                  conj(e1[0] * e2[1] - e1[1] * e2[0])};
    sp_normalize(e3);
    M3 r;
    for (int k = 0; k < 3; k++) {
        r.e[k] = e1[k];
        r.e[3 + k] = e2[k];
        r.e[6 + k] = e3[k];
    }
    return r;
}

int orc_expm_pade(const double *A, double *out) {
    M3 r;
    int bad = expm_pade6(m_load(A), &r);
    if (bad) return bad;
    m_store(out, r);
    return 0;
}

void orc_reunitarize(const double *M, double *out) { m_store(out, reu(m_load(M))); }

/* FFT::fftn (FFT.cpp:260-372) on a [S][18] field; isign +1 forward,
 * -1 inverse (divided by N^2) */
void orc_fftn_shifted(int N, double *f, int isign) {
    const size_t S = (size_t)N * N;
    const int h = N / 2;
    double *w = (double *)malloc(sizeof(double) * 2 * S);
    for (int k = 0; k < 9; k++) {
        /* input[(i - N/2) mod N][(j - N/2) mod N] = data[i][j] */
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                const size_t np = (size_t)((i + h) % N) * N + (size_t)((j + h) % N);
                const size_t p = (size_t)i * N + j;
                w[2 * np] = f[18 * p + 2 * k];
                w[2 * np + 1] = f[18 * p + 2 * k + 1];
            }
        orc_fft2d(N, w, isign > 0 ? 1 : -1);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                const size_t np = (size_t)((i + h) % N) * N + (size_t)((j + h) % N);
                const size_t p = (size_t)i * N + j;
                f[18 * p + 2 * k] = w[2 * np];
                f[18 * p + 2 * k + 1] = w[2 * np + 1];
            }
    }
    free(w);
}

/* GaugeFix::FFTChi (GaugeFix.cpp:17-165) followed, per step, by
 * gaugeTransform at every site (GaugeFix.cpp:167-210). */
int orc_gauge_fix(int N, double *Ux, double *Uy, double *E1, double *E2, double *phi, double *pi,
                  int max_iter, int *iterations, double *residual, double *history) {
    const size_t S = (size_t)N * N;
    M3 t[8];
    for (int a = 0; a < 8; a++) orc_generator(a, &t[a]);
    double *chi = (double *)malloc(sizeof(double) * 18 * S);
    double *G = (double *)malloc(sizeof(double) * 18 * S);
    double prev = 10000., res = 0.;
    int it = 0;
    for (it = 0; it < max_iter; it++) {
        res = 0.;
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                const size_t pos = (size_t)i * N + j;
                const size_t mx = (size_t)(i == 0 ? N - 1 : i - 1) * N + j;
                const size_t my = (size_t)i * N + (j == 0 ? N - 1 : j - 1);
                const M3 divA = m_add(m_sub(m_load(Ux + 18 * pos), m_load(Ux + 18 * mx)),
                                      m_sub(m_load(Uy + 18 * pos), m_load(Uy + 18 * my)));
                M3 g = m_zero();
                for (int a = 0; a < 8; a++) g = m_add(g, m_scale(cimag(m_tr(m_mul(divA, t[a]))), t[a]));
                m_store(chi + 18 * pos, g);
                res += creal(m_tr(m_mul(m_dag(g), g))) / 3.0;
            }
        res /= (double)N * (double)N;
        if (history) history[it] = res;
        if (it % 10 == 0) prev = res;
        if (res < 1e-9) break;
        if (res > prev && res < 1e-6) break;
        orc_fftn_shifted(N, chi, 1);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                const double kx = sin(M_PI * (-0.5 + (double)i / (double)N));
                const double ky = sin(M_PI * (-0.5 + (double)j / (double)N));
                const double kt2 = 4. * (kx * kx + ky * ky);
                const double f = -1.5 * (1. / (kt2 + 1e-9));
                double *c = chi + 18 * ((size_t)i * N + j);
                for (int k = 0; k < 18; k++) c[k] *= f;
            }
        orc_fftn_shifted(N, chi, -1);
        for (size_t p = 0; p < S; p++) {
            M3 x = m_load(chi + 18 * p), e;
            for (int k = 0; k < 9; k++) x.e[k] = I * x.e[k];
            if (expm_pade6(x, &e)) e = m_one();
            e = reu(e);
            if (isnan(creal(e.e[2])) || isnan(cimag(e.e[2]))) e = m_one();
            m_store(G + 18 * p, e);
        }
        /* gaugeTransform site by site, in the reference's order */
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                const size_t pos = (size_t)i * N + j;
                const size_t mx = (size_t)(i == 0 ? N - 1 : i - 1) * N + j;
                const size_t my = (size_t)i * N + (j == 0 ? N - 1 : j - 1);
                const M3 g = m_load(G + 18 * pos);
                m_store(Ux + 18 * pos, m_mul(g, m_load(Ux + 18 * pos)));
                m_store(Uy + 18 * pos, m_mul(g, m_load(Uy + 18 * pos)));
                m_store(Ux + 18 * mx, m_mul_bd(m_load(Ux + 18 * mx), g));
                m_store(Uy + 18 * my, m_mul_bd(m_load(Uy + 18 * my), g));
                double *fl[4] = {E1, E2, phi, pi};
                for (int q = 0; q < 4; q++)
                    if (fl[q]) m_store(fl[q] + 18 * pos, m_mul_bd(m_mul(g, m_load(fl[q] + 18 * pos)), g));
            }
    }
    free(chi);
    free(G);
    if (iterations) *iterations = it;
    if (residual) *residual = res;
    return 0;
}

/* one pass of Evolution::multiplicity over field F (Evolution.cpp:2455-2535)
 * with coefficient c; counts only on the first pass */
static void mult_pass(int N, double a, const double *F, double c, int count, double *dN, double *dE, double *n,
                      double *E, double *cnt) {
    const size_t S = (size_t)N * N;
    const int bins = 100;
    const double dkt = 2.83 / (double)bins;
    const double hbarc = 0.1973269718; /* Phys_consts.h:6 */
    double *w = (double *)malloc(sizeof(double) * 18 * S);
    memcpy(w, F, sizeof(double) * 18 * S);
    orc_fftn_shifted(N, w, 1);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            if (i == 0 || j == 0) continue;
            const size_t pos = (size_t)i * N + j, npos = (size_t)(N - i) * N + (N - j);
            const double kx = 2. * M_PI * (-0.5 + (double)i / (double)N);
            const double ky = 2. * M_PI * (-0.5 + (double)j / (double)N);
            const double kt2 = 4. * (sin(kx / 2.) * sin(kx / 2.) + sin(ky / 2.) * sin(ky / 2.));
            const double omega2 = kt2;
            double nkt = 0.;
            if (omega2 != 0)
                nkt = 2. / sqrt(omega2) / (double)((long long)N * N) *
                      (c * creal(m_tr(m_mul(m_load(w + 18 * pos), m_load(w + 18 * npos)))));
            *dN += nkt;
            *dE += nkt * sqrt(omega2) * hbarc / a;
            for (int ik = 0; ik < bins; ik++) {
                if (fabs(sqrt(kt2)) > ik * dkt && fabs(sqrt(kt2)) <= (ik + 1) * dkt) {
                    n[ik] += nkt / dkt / 2 / M_PI / sqrt(kt2) * 2 * M_PI * sqrt(kt2) * dkt * N * N / M_PI / M_PI /
                             2. / 2.;
                    E[ik] += sqrt(omega2) * hbarc / a * nkt / dkt / 2 / M_PI / sqrt(kt2) * 2 * M_PI * sqrt(kt2) *
                             dkt * N * N / M_PI / M_PI / 2. / 2.;
                    if (count) cnt[ik] += 1.;
                }
            }
        }
    free(w);
}

/* Evolution::multiplicity (Evolution.cpp:2284-3188) on gauge-fixed fields. */
void orc_multiplicity(int N, double L, double g, double dtau, int it, const double *E1, const double *E2,
                      const double *pi, double *out) {
    const int bins = 100;
    const double a = L / N;
    const double hbarc = 0.1973269718; /* Phys_consts.h:6 */
    const double dkt = 2.83 / (double)bins;
    double dN = 0., dE = 0., n[100], E[100], cnt[100];
    for (int k = 0; k < bins; k++) n[k] = E[k] = cnt[k] = 0.;
    mult_pass(N, a, E1, g * g / ((it - 0.5) * dtau), 1, &dN, &dE, n, E, cnt);
    mult_pass(N, a, E2, g * g / ((it - 0.5) * dtau), 0, &dN, &dE, n, E, cnt);
    mult_pass(N, a, pi, (it - 0.5) * dtau, 0, &dN, &dE, n, E, cnt);
    double dN2 = 0., dE2 = 0., n3 = 0., e3 = 0., n6 = 0., e6 = 0.;
    for (int ik = 0; ik < bins; ik++) {
        if (cnt[ik] > 0) {
            n[ik] = n[ik] / cnt[ik];
            E[ik] = E[ik] / cnt[ik];
            dN2 += n[ik] * (ik + 0.5) * dkt * dkt * 2. * M_PI;
            dE2 += E[ik] * (ik + 0.5) * dkt * dkt * 2. * M_PI;
            if (ik * dkt / a * hbarc > 3.) {
                n3 += n[ik] * (ik + 0.5) * dkt * dkt * 2. * M_PI;
                e3 += E[ik] * (ik + 0.5) * dkt * dkt * 2. * M_PI;
            }
            if (ik * dkt / a * hbarc > 6.) {
                n6 += n[ik] * (ik + 0.5) * dkt * dkt * 2. * M_PI;
                e6 += E[ik] * (ik + 0.5) * dkt * dkt * 2. * M_PI;
            }
        }
    }
    out[0] = dN;
    out[1] = dE;
    out[2] = dN2;
    out[3] = dE2;
    out[4] = n3;
    out[5] = e3;
    out[6] = n6;
    out[7] = e6;
    for (int k = 0; k < bins; k++) {
        out[8 + k] = n[k];
        out[108 + k] = E[k];
        out[208 + k] = cnt[k];
    }
}

/* Evolution::run (Evolution.cpp:366-657) with computeGluonMultiplicity: the
 * orc_run schedule, and after the updates of steps 1 and itmax the gauge fix
 * (4000 steps at most) and the spectra.  m1, mmax receive orc_multiplicity's
 * 308 values; gf = iterations and residual at step 1, then at itmax. */
int orc_run_multiplicity(int N, double g, double L, double dtau, int itmax, double *Ux, double *Uy, double *E1,
                         double *E2, double *phi, double *pi, double *T, double obs[3], double *m1, double *mmax,
                         double gf[4]) {
    const double a = L / N;
    orc_evolve_phi(N, dtau, 0., pi, phi);
    orc_evolve_u(N, g, dtau, 0., E1, E2, Ux, Uy);
    for (int it = 1; it <= itmax; it++) {
        if (it == itmax) orc_tmunu(N, g, a, dtau, it, Ux, Uy, E1, E2, phi, pi, T);
        if (it < itmax) {
            orc_evolve_pi(N, dtau, it * dtau, Ux, Uy, phi, pi);
            orc_evolve_e(N, g, dtau, it * dtau, Ux, Uy, phi, E1, E2);
            orc_evolve_phi(N, dtau, it * dtau, pi, phi);
            orc_evolve_u(N, g, dtau, it * dtau, E1, E2, Ux, Uy);
        } else {
            orc_evolve_pi(N, dtau / 2., it * dtau, Ux, Uy, phi, pi);
            orc_evolve_e(N, g, dtau / 2., it * dtau, Ux, Uy, phi, E1, E2);
        }
        if (it == 1 || it == itmax) {
            int iters = 0;
            double res = 0.;
            orc_gauge_fix(N, Ux, Uy, E1, E2, phi, pi, 4000, &iters, &res, NULL);
            orc_multiplicity(N, L, g, dtau, it, E1, E2, pi, it == itmax ? mmax : m1);
            gf[it == itmax ? 2 : 0] = iters;
            gf[it == itmax ? 3 : 1] = res;
        }
    }
    const double tau_fm = itmax * dtau * a;
    obs[0] = itmax >= 1 ? orc_dedeta(N, a, tau_fm, T) : 0.;
    obs[1] = orc_gauss_law(N, Ux, Uy, E1, E2, phi, pi);
    obs[2] = tau_fm;
    return 0;
}
