/* orc_init.c -- oracle Wilson lines and tau=0 fields.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   orc_setv          Init::setV (Init.cpp:1678-1850): per nucleus and slice,
 *                     rho^a = g sqrt(g2mu2/Ny) z^a, FFT, x 1/(k^2+m^2)
 *                     exp(-|k| UVdamp) (zero mode dropped when m=0), inverse
 *                     FFT, V <- exp(-i Re(A^a) t^a) V
 *   orc_newton_link   Init::findUInForwardLightconeChun (Init.cpp:3344-3481)
 *                     with the forward-difference Jacobian, its analytic
 *                     fallback, the 1e5 step guard and restarts; the LU solve
 *                     is Init::solveAxb (Init.cpp:47-63): Gaussian elimination
 *                     with partial pivoting (first maximal pivot), unit-lower
 *                     forward and upper back substitution.  Restart draws use
 *                     the per-link counter map instead of the shared MT stream
 *                     (which the reference reads from inside an OpenMP loop).
 *   orc_init_fields   Init::init (Init.cpp:2437-2690): NaN-trace guard, links
 *                     U^(1,2)_{x,y} with clamped neighbours, root finder per
 *                     link, A^eta, pi = -(2i/g) A^eta, E = phi = 0.
 */
#include <math.h>
#include <stdlib.h>

#include "ipg_oracle.h"
#include "orc_internal.h"

int orc_setv(int N, int Ny, double g, double m, double uvdamp, const double *g2mu2A,
             const double *g2mu2B, const double *z, orc_mt *mt, uint64_t seed, double *V) {
    const int S = N * N;
    double *rho = malloc(sizeof(double) * 2 * (size_t)S * 8); /* [8][S] complex */
    double *zz = malloc(sizeof(double) * (size_t)S * 8);
    double *mult = malloc(sizeof(double) * (size_t)S);
    if (!rho || !zz || !mult) return -1;
    /* k-space multiplier at unshifted index (q_x, q_y) */
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) {
            double kx = 2. * M_PI * (double)i / (double)N;
            double ky = 2. * M_PI * (double)j / (double)N;
            double kt2 = 4. * (sin(kx / 2.) * sin(kx / 2.) + sin(ky / 2.) * sin(ky / 2.));
            double f;
            if (m == 0)
                f = kt2 != 0 ? 1. / kt2 : 0.;
            else
                f = (1. / (kt2 + m * m)) * exp(-sqrt(kt2) * uvdamp);
            mult[i * N + j] = f;
        }
    }
    for (int nuc = 0; nuc < 2; nuc++) {
        const double *g2 = nuc == 0 ? g2mu2A : g2mu2B;
        double *Vn = V + (size_t)nuc * S * 18;
        for (int s = 0; s < S; s++) m_store(&Vn[18 * (size_t)s], m_one());
        for (int k = 0; k < Ny; k++) {
            if (z) {
                const double *src = z + (((size_t)nuc * Ny + k) * S) * 8;
                for (size_t t = 0; t < (size_t)S * 8; t++) zz[t] = src[t];
            } else if (mt) {
                for (size_t t = 0; t < (size_t)S * 8; t++) zz[t] = orc_mt_gauss(mt);
            } else {
                orc_rho_normals_slice(seed, nuc, k, S, zz);
            }
            for (int s = 0; s < S; s++) {
                double amp = g * sqrt(g2[s] / (double)Ny);
                for (int n = 0; n < 8; n++) {
                    rho[2 * ((size_t)n * S + s)] = amp * zz[8 * (size_t)s + n];
                    rho[2 * ((size_t)n * S + s) + 1] = 0.;
                }
            }
            for (int n = 0; n < 8; n++) orc_fft2d(N, rho + 2 * (size_t)n * S, +1);
            for (int n = 0; n < 8; n++) {
                double *r = rho + 2 * (size_t)n * S;
                for (int s = 0; s < S; s++) {
                    r[2 * s] *= mult[s];
                    r[2 * s + 1] *= mult[s];
                }
            }
            for (int n = 0; n < 8; n++) orc_fft2d(N, rho + 2 * (size_t)n * S, -1);
#pragma omp parallel for
            for (int s = 0; s < S; s++) {
                double Q[8];
                for (int n = 0; n < 8; n++) Q[n] = -rho[2 * ((size_t)n * S + s)];
                M3 e = orc_exp_iQt(Q);
                M3 v = m_load(&Vn[18 * (size_t)s]);
                m_store(&Vn[18 * (size_t)s], m_mul(e, v));
            }
        }
    }
    free(rho);
    free(zz);
    free(mult);
    return 0;
}

/* Gaussian elimination with partial pivoting, then forward/back substitution */
static void solve8(double *J /* row-major 8x8, destroyed */, const double *F, double *x) {
    int perm[8];
    for (int i = 0; i < 8; i++) perm[i] = i;
    for (int j = 0; j < 7; j++) {
        double mx = fabs(J[8 * j + j]);
        int ip = j;
        for (int i = j + 1; i < 8; i++) {
            double v = fabs(J[8 * i + j]);
            if (v > mx) {
                mx = v;
                ip = i;
            }
        }
        if (ip != j) {
            for (int k = 0; k < 8; k++) {
                double t = J[8 * j + k];
                J[8 * j + k] = J[8 * ip + k];
                J[8 * ip + k] = t;
            }
            int t = perm[j];
            perm[j] = perm[ip];
            perm[ip] = t;
        }
        double ajj = J[8 * j + j];
        if (ajj != 0.0) {
            for (int i = j + 1; i < 8; i++) {
                double aij = J[8 * i + j] / ajj;
                J[8 * i + j] = aij;
                for (int k = j + 1; k < 8; k++) J[8 * i + k] = J[8 * i + k] - aij * J[8 * j + k];
            }
        }
    }
    for (int i = 0; i < 8; i++) x[i] = F[perm[i]];
    for (int i = 1; i < 8; i++) {
        double t = x[i];
        for (int j = 0; j < i; j++) t -= J[8 * i + j] * x[j];
        x[i] = t;
    }
    x[7] = x[7] / J[63];
    for (int i = 6; i >= 0; i--) {
        double t = x[i];
        for (int j = i + 1; j < 8; j++) t -= J[8 * i + j] * x[j];
        x[i] = t / J[8 * i + i];
    }
}

/* extra Newton iterations after the stopping rule holds (the library's
 * newton_polish option); 0 is the reference's rule.  They are exact Newton
 * steps about the current iterate: U <- exp(i beta^a t^a) U with
 * J_ab = -dF_a/dbeta_b = -2 Re tr(t^a W U^+ t^b) at beta = 0, which
 * converges quadratically (the reference's forward-difference Jacobian in
 * alpha, with steps of at least 1e-3, does not converge past ~1e-9). */
static int newton_polish = 0;
void orc_set_newton_polish(int n) { newton_polish = n < 0 ? 0 : n; }

int orc_newton_link(const double *pU1, const double *pU2, uint64_t seed, uint32_t link,
                    double *pUsol, orc_newton_stat *st) {
    const int maxIterations = 2000, maxRetrys = 200;
    M3 U1 = m_load(pU1), U2 = m_load(pU2);
    M3 U0 = m_mul(U1, U2);
    M3 W = m_add(U1, U2);
    M3 Wd = m_dag(W);
    M3 t[8];
    cplx trc[8];
    for (int a = 0; a < 8; a++) {
        orc_generator(a, &t[a]);
        trc[a] = m_tr(m_mul(t[a], m_sub(W, Wd)));
    }
    double Jab[64], Fa[8], dal[8];
    double Fzero = 10., FzeroMin = 1e6;
    M3 best = m_one();
    double alpha[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    M3 Usol = m_mul(orc_exp_iQt(alpha), U0);
    M3 Ud = m_dag(Usol);
    int iter = 0, nRestart = 0, total = 0, analytic = 0;
    uint32_t ndraw = 0; /* restart draws taken so far on this link */
    int npol = 0;
    for (;;) {
        if (!(Fzero > 1e-6 && iter < maxIterations && nRestart < maxRetrys)) {
            if (!(nRestart != maxRetrys && (Fzero <= 1e-6 || npol > 0) && npol < newton_polish)) break;
            npol++;
        }
        if (npol > 0) {
            /* exact Newton step about the current iterate */
            iter++;
            total++;
            M3 Mt = m_sub(m_mul(W, Ud), m_mul(Usol, Wd));
            Fzero = 0.;
            for (int a = 0; a < 8; a++) {
                Fa[a] = cimag((-1.) * (trc[a] + m_trprod(t[a], Mt)));
                Fzero += fabs(Fa[a]);
            }
            M3 WUd = m_mul(W, Ud);
            for (int b = 0; b < 8; b++) {
                M3 X = m_mul(WUd, t[b]);
                for (int a = 0; a < 8; a++) Jab[8 * a + b] = -2. * creal(m_trprod(t[a], X));
            }
            solve8(Jab, Fa, dal);
            Usol = m_mul(orc_exp_iQt(dal), Usol);
            Ud = m_dag(Usol);
            continue;
        }
        iter++;
        total++;
        Fzero = 0.;
        M3 Mt = m_sub(m_mul(W, Ud), m_mul(Usol, Wd));
        for (int a = 0; a < 8; a++) {
            cplx tl = m_trprod(t[a], Mt);
            cplx res = (-1.) * (trc[a] + tl);
            Fa[a] = cimag(res);
            Fzero += fabs(Fa[a]);
        }
        int good = 1;
        for (int b = 0; b < 8; b++) {
            double d = fmax(0.001, fmin(10., 0.01 * fabs(alpha[b])));
            alpha[b] = alpha[b] + d;
            M3 Mp = m_dag(m_mul(orc_exp_iQt(alpha), U0));
            Mp = m_mul(W, m_sub(Mp, Ud));
            double chk = 0.;
            for (int a = 0; a < 8; a++) {
                cplx tl = m_trprod(t[a], Mp);
                Jab[8 * a + b] = 2. * cimag(tl) / d;
                chk += fabs(Jab[8 * a + b]);
                if (chk < 1e-15) good = 0;
            }
            alpha[b] = alpha[b] - d;
        }
        if (!good) {
            analytic = 1;
            for (int b = 0; b < 8; b++) {
                M3 Mb = m_mul(t[b], Ud);
                for (int a = 0; a < 8; a++) {
                    cplx tl = m_trprod(m_mul(t[a], W), Mb);
                    Jab[8 * a + b] = -2. * creal(tl);
                }
            }
        }
        solve8(Jab, Fa, dal);
        int ok = 1;
        for (int a = 0; a < 8; a++)
            if (fabs(dal[a]) > 1e5) {
                ok = 0;
                break;
            }
        if (ok) {
            for (int a = 0; a < 8; a++) alpha[a] = alpha[a] + dal[a];
        } else {
            orc_newton_restart_normals(seed, link, ndraw++, alpha);
        }
        Usol = m_mul(orc_exp_iQt(alpha), U0);
        Ud = m_dag(Usol);
        if (Fzero < FzeroMin) {
            FzeroMin = Fzero;
            best = Usol;
        }
        if (iter == maxIterations) {
            orc_newton_restart_normals(seed, link, ndraw++, alpha);
            Usol = m_mul(orc_exp_iQt(alpha), U0);
            Ud = m_dag(Usol);
            nRestart++;
            iter = 0;
        }
    }
    int conv = 1;
    if (nRestart == maxRetrys) {
        Usol = best;
        conv = 0;
    }
    m_store(pUsol, Usol);
    if (st) {
        st->iterations = total;
        st->restarts = nRestart;
        st->converged = conv;
        st->used_analytic_jacobian = analytic;
    }
    return conv;
}

static int has_nan_trace(M3 m) {
    cplx t = m_tr(m);
    return isnan(creal(t)) || isnan(cimag(t));
}

int orc_init_fields(int N, double g, double *VA, double *VB, uint64_t seed, double *Ux,
                    double *Uy, double *E1, double *E2, double *phi, double *pi,
                    double stats[5]) {
    const int S = N * N;
    double *Ux1 = malloc(sizeof(double) * 18 * (size_t)S), *Uy1 = malloc(sizeof(double) * 18 * (size_t)S);
    double *Ux2 = malloc(sizeof(double) * 18 * (size_t)S), *Uy2 = malloc(sizeof(double) * 18 * (size_t)S);
    for (int s = 0; s < S; s++) {
        if (has_nan_trace(m_load(&VA[18 * (size_t)s]))) m_store(&VA[18 * (size_t)s], m_one());
        if (has_nan_trace(m_load(&VB[18 * (size_t)s]))) m_store(&VB[18 * (size_t)s], m_one());
    }
#define PX(s) ((((s) / N) < N - 1 ? (s) + N : (s)))
#define PY(s) ((((s) % N) < N - 1 ? (s) + 1 : (s)))
#define MX(s) ((((s) / N) > 0 ? (s) - N : (s)))
#define MY(s) ((((s) % N) > 0 ? (s) - 1 : (s)))
#pragma omp parallel for
    for (int s = 0; s < S; s++) {
        M3 va = m_load(&VA[18 * (size_t)s]), vb = m_load(&VB[18 * (size_t)s]);
        m_store(&Ux1[18 * (size_t)s], m_mul(va, m_dag(m_load(&VA[18 * (size_t)PX(s)]))));
        m_store(&Uy1[18 * (size_t)s], m_mul(va, m_dag(m_load(&VA[18 * (size_t)PY(s)]))));
        m_store(&Ux2[18 * (size_t)s], m_mul(vb, m_dag(m_load(&VB[18 * (size_t)PX(s)]))));
        m_store(&Uy2[18 * (size_t)s], m_mul(vb, m_dag(m_load(&VB[18 * (size_t)PY(s)]))));
    }
    long sum_it = 0;
    int max_it = 0, restarts = 0, unconv = 0, analytic = 0;
#pragma omp parallel for reduction(+ : sum_it, restarts, unconv, analytic) reduction(max : max_it)
    for (int s = 0; s < S; s++) {
        orc_newton_stat st;
        orc_newton_link(&Ux1[18 * (size_t)s], &Ux2[18 * (size_t)s], seed, 2u * (uint32_t)s,
                        &Ux[18 * (size_t)s], &st);
        sum_it += st.iterations;
        if (st.iterations > max_it) max_it = st.iterations;
        restarts += st.restarts;
        unconv += !st.converged;
        analytic += st.used_analytic_jacobian;
        orc_newton_link(&Uy1[18 * (size_t)s], &Uy2[18 * (size_t)s], seed, 2u * (uint32_t)s + 1u,
                        &Uy[18 * (size_t)s], &st);
        sum_it += st.iterations;
        if (st.iterations > max_it) max_it = st.iterations;
        restarts += st.restarts;
        unconv += !st.converged;
        analytic += st.used_analytic_jacobian;
    }
    /* A^eta from the four links around the site (Init.cpp:2521-2584) */
#pragma omp parallel for
    for (int s = 0; s < S; s++) {
        int nbx = MX(s), nby = MY(s);
        M3 D, Dd, U, Udg, acc;
        /* +x link at s */
        D = m_sub(m_load(&Ux1[18 * (size_t)s]), m_load(&Ux2[18 * (size_t)s]));
        Dd = m_sub(m_dag(m_load(&Ux1[18 * (size_t)s])), m_dag(m_load(&Ux2[18 * (size_t)s])));
        U = m_load(&Ux[18 * (size_t)s]);
        Udg = m_dag(U);
        acc = m_add(m_sub(m_sub(m_mul(D, Udg), D), m_mul(U, Dd)), Dd);
        /* x link at s - x */
        D = m_sub(m_load(&Ux1[18 * (size_t)nbx]), m_load(&Ux2[18 * (size_t)nbx]));
        Dd = m_sub(m_dag(m_load(&Ux1[18 * (size_t)nbx])), m_dag(m_load(&Ux2[18 * (size_t)nbx])));
        U = m_load(&Ux[18 * (size_t)nbx]);
        Udg = m_dag(U);
        acc = m_sub(m_add(m_add(m_sub(acc, m_mul(Udg, D)), D), m_mul(Dd, U)), Dd);
        /* +y link at s */
        D = m_sub(m_load(&Uy1[18 * (size_t)s]), m_load(&Uy2[18 * (size_t)s]));
        Dd = m_sub(m_dag(m_load(&Uy1[18 * (size_t)s])), m_dag(m_load(&Uy2[18 * (size_t)s])));
        U = m_load(&Uy[18 * (size_t)s]);
        Udg = m_dag(U);
        acc = m_add(m_sub(m_sub(m_add(acc, m_mul(D, Udg)), D), m_mul(U, Dd)), Dd);
        /* y link at s - y */
        D = m_sub(m_load(&Uy1[18 * (size_t)nby]), m_load(&Uy2[18 * (size_t)nby]));
        Dd = m_sub(m_dag(m_load(&Uy1[18 * (size_t)nby])), m_dag(m_load(&Uy2[18 * (size_t)nby])));
        U = m_load(&Uy[18 * (size_t)nby]);
        Udg = m_dag(U);
        acc = m_sub(m_add(m_add(m_sub(acc, m_mul(Udg, D)), D), m_mul(Dd, U)), Dd);
        M3 Aeta = m_scale(1. / 8., acc);
        M3 p = m_cscale(CMPLX(0., -2. / g * orc_mut[ORC_MUT_PI_INIT]), Aeta);
        m_store(&pi[18 * (size_t)s], p);
        m_store(&E1[18 * (size_t)s], m_zero());
        m_store(&E2[18 * (size_t)s], m_zero());
        m_store(&phi[18 * (size_t)s], m_zero());
    }
#undef PX
#undef PY
#undef MX
#undef MY
    free(Ux1);
    free(Uy1);
    free(Ux2);
    free(Uy2);
    if (stats) {
        stats[0] = (double)sum_it;
        stats[1] = max_it;
        stats[2] = restarts;
        stats[3] = unconv;
        stats[4] = analytic;
    }
    return unconv;
}
