/*
 * ipg_oracle.h -- CPU restatement of the IP-Glasma hot path (schenke/ipglasma)
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (ipglasma_amd/, include/ipg.h) never links, loads or calls it.
 *
 * Every function restates the reference algorithm in plain C99 (strict IEEE:
 * built with -O2 -fno-fast-math -ffp-contract=off) and cites the reference
 * file:line it follows.  Host-visible layouts:
 *   site index      s = ix*N + iy  (ix <-> x, stride N; iy <-> y, contiguous)
 *                   (Lattice.cpp:19-28)
 *   3x3 complex     9 complex = 18 doubles per site, row-major e[3*i+j]
 *                   (Matrix.h:45), interleaved re,im
 *   field arrays    [S][18] per field ("site-major")
 *
 * Parity pinning (what checks each part, see tests/test_oracle_*.py):
 *   orc_mt_*         std::mt19937_64 known-answer + the C++ standard's 10000th
 *                    output KAT; real1/2/3 closed forms
 *   orc_philox4x32   Random123 published known-answer vectors
 *   orc_gauss        Marsaglia polar closed form from recorded uniforms
 *   orc_fft2d        numpy.fft.fft2/ifft2; plane-wave eigenfunction
 *   orc_expm_coeff   scipy.linalg.expm(i Q^a t^a); unitarity; det=1
 *   orc_newton_link  residual F_a = 0; U1=1 => U=U2, U2=1 => U=U1; abelian
 *                    closed form
 *   orc_evolve_*     vacuum stays vacuum; gauge covariance; Gauss law kept
 *   orc_tmunu        tracelessness T^tt - T^xx - T^yy - tau^2 T^ee = 0;
 *                    vacuum T = 0; gauge invariance
 *   orc_sample_*     Woods-Saxon KS test; d_min; recentring; rotation norm
 *   orc_tp_sum       single-nucleon Gaussian closed form
 *   orc_wounding     brute-force pair loop / hard-sphere closed form
 *   orc_g2mu2        table-node exactness; fixed-point residual
 *   orc_eig4         numpy.linalg.eig (LAPACK dgeev) eigenvalues and
 *                    eigenvector directions; A v = lambda v residual
 *   orc_landau       boosted perfect fluid and anisotropic fluid closed forms
 *                    (eps, u^mu, pi^{mu nu}); u.u = 1; sign u^tau > 0
 *   orc_eccentricity elliptic and rotated Gaussian closed forms; moments of
 *                    point sources; translation invariance
 *   orc_expm_pade    scipy.linalg.expm
 *   orc_reunitarize  QR factorisation of the first two rows; det = 1
 *   orc_fftn_shifted numpy fftshift(fft2(ifftshift(.)))
 *   orc_gauge_fix    linearised step divides div A by 4 (the -1.5/k^2
 *                    filter's closed form); gauge-invariant traces kept;
 *                    Coulomb residual below 1e-9
 *   orc_multiplicity single-mode E and pi closed forms; excluded momentum
 *                    lines; Parseval over the included modes
 * Functions with no check of their own are marked "parity unpinned" below.
 */
#ifndef IPG_ORACLE_H
#define IPG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mutation tests ------------------------------------------------------------
 * A test plants a deliberate mistake by scaling one coefficient of the
 * restatement (factor != 1) and checks that its pin then fails; factors are
 * 1 otherwise (multiplication by 1 is exact, so results are unchanged).
 * Not thread-safe with oracle calls in flight. */
enum {
    ORC_MUT_EVOLVE_U = 0,   /* i g^2 dtau/(tau + dtau/2) of evolveU            */
    ORC_MUT_EVOLVE_PHI = 1, /* (tau + dtau/2) dtau of evolvePhi                */
    ORC_MUT_EVOLVE_PI = 2,  /* dtau/tau of evolvePi                            */
    ORC_MUT_E_PLAQ = 3,     /* i tau dtau/(2 g^2) of evolveE's plaquettes      */
    ORC_MUT_E_COMM = 4,     /* i dtau/tau of evolveE's [phi~, phi]             */
    ORC_MUT_T_ELEC = 5,     /* transverse electric part of T^{mu nu}           */
    ORC_MUT_T_LONG = 6,     /* longitudinal pi part                            */
    ORC_MUT_T_MAG = 7,      /* plaquette part                                  */
    ORC_MUT_T_PHIGRAD = 8,  /* transverse phi-gradient part                    */
    ORC_MUT_TX_F = 9,       /* T^{tau x}: E^y x plaquette clover term          */
    ORC_MUT_TX_K = 10,      /* T^{tau x}: pi x D_x phi term                    */
    ORC_MUT_TY_F = 11,
    ORC_MUT_TY_K = 12,
    ORC_MUT_TE = 13,        /* T^{tau eta}                                     */
    ORC_MUT_XY_E = 14,      /* T^{xy}: E^x E^y term                            */
    ORC_MUT_XY_B = 15,      /* T^{xy}: D_x phi D_y phi terms                   */
    ORC_MUT_XE_E = 16,      /* T^{x eta}: E^x pi term                          */
    ORC_MUT_XE_B = 17,      /* T^{x eta}: plaquette x D_y phi term             */
    ORC_MUT_YE_E = 18,
    ORC_MUT_YE_B = 19,
    ORC_MUT_PI_INIT = 20,   /* -2i/g of pi = -(2i/g) A^eta at tau = 0          */
    ORC_MUT_GAUSS_COMM = 21 /* i [phi, pi] of the Gauss-law check              */
};
void orc_set_mutation(int k, double factor);
void orc_clear_mutations(void);
/* extra Newton iterations of orc_newton_link after its stopping rule holds
 * (the library's newton_polish option; 0, the default, is the reference's) */
void orc_set_newton_polish(int n);
/* OpenMP thread count of the oracle's parallel loops */
void orc_set_threads(int n);
int orc_get_threads(void);

/* ---- random numbers (Random.cpp:67-204) ---------------------------------- */
typedef struct {
    uint64_t mt[312];
    int mti;
    int iset;
    double gset;
} orc_mt;

void orc_mt_init(orc_mt *r, uint64_t seed);
uint64_t orc_mt_u64(orc_mt *r);
double orc_mt_real1(orc_mt *r);
double orc_mt_real2(orc_mt *r);
double orc_mt_real3(orc_mt *r);
double orc_mt_gauss(orc_mt *r);
/* fill n draws of kind k (0:u64 as double bits, 1:real1, 2:real2, 3:real3,
 * 4:gauss) -- convenience for tests */
void orc_mt_fill(orc_mt *r, int kind, int64_t n, double *out);

/* Philox4x32-10 (Salmon et al., SC'11) */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4]);
/* the 8 colour normals of colour-charge slice (nucleus, k) at site s: the
 * project's counter map, shared with the device generator */
void orc_rho_normals_site(uint64_t seed, int nucleus, int k, uint32_t s,
                          double z[8]);
void orc_rho_normals_slice(uint64_t seed, int nucleus, int k, int S,
                           double *z /* [S][8] */);
/* deterministic restart draws for the tau=0 root finder (per link) */
void orc_newton_restart_normals(uint64_t seed, uint32_t link, uint32_t restart,
                                double alpha[8]);

/* ---- SU(3) algebra (Matrix.cpp:310-426, Init.cpp:3483-3499) -------------- */
void orc_expm_coeff(const double Q[8], double coeff[18]);
void orc_u_from_exponent(const double Q[8], double U[18]);
void orc_matmul(const double *A, const double *B, double *C);

/* ---- FFT (FFT.cpp:492-610 semantics, unshifted form) --------------------- */
/* in-place 2D DFT of complex data[N][N] (interleaved). sign=+1 forward
 * (exp(-i...)), sign=-1 inverse with 1/N^2 */
int orc_fft2d(int N, double *data, int sign);

/* ---- nucleons and colour charge (Init.cpp, Glauber.cpp) ----------------- */
typedef struct {
    double x, y, z, phi;
    int collided;
    int proton;
} orc_nucleon;

typedef struct {
    /* projectile (A) and target (B) nucleus data after FindNucleusData2 */
    int A1, Z1, A2, Z2;
    double R1, a1, beta2_1, beta3_1, beta4_1, gamma_1, dR1, da1;
    double R2, a2, beta2_2, beta3_2, beta4_2, gamma_2, dR2, da2;
    int force_dmin;
    double d_min;
    double bmin, bmax;
    int linear_b;
    int rotate_rp;
    double xi; /* proton anisotropy */
    double sigma_nn_mb;
    int gaussian_wounding;
} orc_glauber_cfg;

typedef struct {
    double b, phiRP;
    int Ncoll, Npart;
} orc_event_info;

/* b draw, sampleTA (A then B, rotations), reaction plane shift, wounding.
 * nucleus arrays must hold A1 / A2 entries.  Returns 0, or <0 on an
 * unsupported configuration (deuteron, file-based nuclei). */
/* the deuteron's nucleon thickness at transverse distance s [fm] for the
 * Hulthen ranges a, w [fm^-1], starting normalisation R0 and cross section
 * sigma [fm^2] (Glauber.cpp:513-576, 847-1100) */
double orc_deuteron_thickness(double s, double a, double w, double R0, double sigma_fm2);
int orc_sample_event(orc_mt *r, const orc_glauber_cfg *cfg, orc_nucleon *nucA,
                     orc_nucleon *nucB, orc_event_info *info);

/* T_p sums (Init.cpp:1009-1115) */
void orc_tp_sum(int N, double L, double BG, double xi, const orc_nucleon *nuc,
                int A, double *Tp);

/* Q_s^2 table lookup (Init.cpp:686-749); table Qs2[iT*44+iy], Tlist[200] */
double orc_nuclear_qs2(const double *Tlist, const double *Qs2, double T,
                       double y);

/* g2mu2 for both nuclei (Init.cpp:1219-1421) with the serial-chain rapidity
 * start of a single-threaded run */
typedef struct {
    int N;
    double L, g, BG, rmax, QsmuRatio, QsmuRatioB;
    double xFactor, roots, rapidity;
    int useFluctuatingx;
} orc_qs_cfg;
void orc_g2mu2(const orc_qs_cfg *c, const double *Tlist, const double *Qs2,
               const double *TpA, const double *TpB, double *g2mu2A,
               double *g2mu2B);

/* statistics + success (Init.cpp:1423-1618). out[0..7] = averageQs (max),
 * averageQsAvg, averageQsmin, Tpp, averageQs2min2*a^2/hbarc^2, count, alphas,
 * success */
void orc_stats(const orc_qs_cfg *c, double sigma_nn_mb, int Npart,
               double minimumQs2ST, const orc_nucleon *nucA, int A1,
               const orc_nucleon *nucB, int A2, const double *g2mu2A,
               const double *g2mu2B, const double *TpA, const double *TpB,
               double out[8]);

/* ---- Wilson lines (Init.cpp:1678-1850) ---------------------------------- */
/* rho normals: if z != NULL it holds [2][Ny][S][8]; else if mt != NULL the
 * reference MT Gauss stream is drawn (pos-major, colour-minor, nucleus A
 * then B); else the project Philox map with `seed` is used.  V[2][S][18]. */
int orc_setv(int N, int Ny, double g, double m_lat, double uvdamp_lat,
             const double *g2mu2A, const double *g2mu2B, const double *z,
             orc_mt *mt, uint64_t seed, double *V);

/* ---- tau=0 fields (Init.cpp:2303-2696, 3344-3481) ----------------------- */
typedef struct {
    int iterations; /* total Newton iterations over all restarts */
    int restarts;
    int converged;
    int used_analytic_jacobian;
} orc_newton_stat;

int orc_newton_link(const double *U1, const double *U2, uint64_t seed,
                    uint32_t link, double *Usol, orc_newton_stat *st);

/* from V_A, V_B: NaN guard, links, Newton per link, pi.  Output state (all
 * [S][18]): Ux, Uy, E1=0, E2=0, phi=0, pi.  stats[0]=sum iterations,
 * stats[1]=max iterations, stats[2]=restarts, stats[3]=unconverged links,
 * stats[4]=links with analytic Jacobian */
int orc_init_fields(int N, double g, double *VA, double *VB, uint64_t seed,
                    double *Ux, double *Uy, double *E1, double *E2,
                    double *phi, double *pi, double stats[5]);

/* ---- evolution (Evolution.cpp:35-293, 366-431) -------------------------- */
void orc_evolve_u(int N, double g, double dtau, double tau, const double *E1,
                  const double *E2, double *Ux, double *Uy);
void orc_evolve_phi(int N, double dtau, double tau, const double *pi,
                    double *phi);
void orc_evolve_pi(int N, double dtau, double tau, const double *Ux,
                   const double *Uy, const double *phi, double *pi);
void orc_evolve_e(int N, double g, double dtau, double tau, const double *Ux,
                  const double *Uy, const double *phi, double *E1,
                  double *E2);

/* full Evolution::run schedule through itmax; T^{mu nu} (lattice-scaled to
 * fm units) at it=itmax before the final half step into T[10][S] in the
 * order tautau, xx, yy, etaeta, taux, tauy, taueta, xy, xeta, yeta.
 * obs[0] = dE/deta [GeV], obs[1] = Gauss-law max, obs[2] = tau [fm] */
int orc_run(int N, double g, double L, double dtau, int itmax, double *Ux,
            double *Uy, double *E1, double *E2, double *phi, double *pi,
            double *T, double obs[3]);

/* T^{mu nu} at step it (Evolution.cpp:659-1246), physical units */
void orc_tmunu(int N, double g, double a, double dtau, int it,
               const double *Ux, const double *Uy, const double *E1,
               const double *E2, const double *phi, const double *pi,
               double *T);
double orc_gauss_law(int N, const double *Ux, const double *Uy,
                     const double *E1, const double *E2, const double *phi,
                     const double *pi);

/* dE/deta = tau * a^2 * hbarc * sum T^tautau  (GeV) */
double orc_dedeta(int N, double a, double tau_fm, const double *Ttautau);

/* real nonsymmetric 4x4 eigenproblem (row-major A): eigenvalues wr + i wi,
 * eigenvector k in column k of vr + i vi (row-major), unit Euclidean norm;
 * a complex pair occupies adjacent columns, positive imaginary part first.
 * Returns nonzero when the QR iteration does not converge. */
int orc_eig4(const double *A, double *wr, double *wi, double *vr, double *vi);

/* Landau frame at step it (MyEigen::flowVelocity4D): T[10][S] as above ->
 * out[15][S] = eps, u^tau, u^x, u^y, u^eta, pi^{mu nu} (T's order);
 * avg = sum u_x^2 eps, sum u_y^2 eps, sum (tau u^eta)^2 eps, sum eps over
 * the timelike eigenvectors found */
void orc_landau(int N, double L, double dtau, int it, const double *T,
                double *out, double avg[4], int *nfail);

/* Evolution::eccentricity with gfactor = 1: res[0..5] = eps_1..eps_6,
 * res[6..11] = Psi_1..Psi_6, res[12] = area, res[13] = Rbar,
 * res[14] = <eps> hbarc, res[15] = <Q_sA^2 Q_sB^2> hbarc, res[16] = sqrt<r^2>,
 * res[17], res[18] = centroid */
void orc_eccentricity(int N, double L, double QsmuRatio, double QsmuRatioB,
                      double cutoff, const double *eps, const double *utau,
                      const double *g2A, const double *g2B, double *res);

/* ---- Coulomb gauge and gluon spectra (GaugeFix.cpp, Evolution.cpp) ----- */
/* Matrix::expm, Pade order 6 (Matrix.cpp:431-560); nonzero on NaN/inf */
int orc_expm_pade(const double *A, double *out);
/* Matrix::reu (Matrix.h:67-90) */
void orc_reunitarize(const double *M, double *out);
/* FFT::fftn (FFT.cpp:260-372) on a [S][18] field: quadrant-swapped DFT,
 * isign +1 forward (exp(-i...)), -1 inverse divided by N^2 */
void orc_fftn_shifted(int N, double *f, int isign);
/* GaugeFix::FFTChi + gaugeTransform (GaugeFix.cpp:17-210) in place; any of
 * E1, E2, phi, pi may be NULL; history (max_iter doubles, may be NULL)
 * receives the residual of every step */
int orc_gauge_fix(int N, double *Ux, double *Uy, double *E1, double *E2, double *phi, double *pi,
                  int max_iter, int *iterations, double *residual, double *history);
/* Evolution::multiplicity (Evolution.cpp:2284-3188) on gauge-fixed fields:
 * out[0..7] = dN/dy 1, dE/dy 1, dN/dy 2, dE/dy 2, dN/dy and dE/dy above
 * k_T = 3 GeV, above 6 GeV; out[8..107] n[ik], out[108..207] E[ik],
 * out[208..307] counts */
void orc_multiplicity(int N, double L, double g, double dtau, int it, const double *E1, const double *E2,
                      const double *pi, double *out);
/* orc_run with the gauge fix and spectra after steps 1 and itmax;
 * gf = iterations, residual at step 1 then at itmax */
int orc_run_multiplicity(int N, double g, double L, double dtau, int itmax, double *Ux, double *Uy, double *E1,
                         double *E2, double *phi, double *pi, double *T, double obs[3], double *m1, double *mmax,
                         double gf[4]);

/* lattice-unit derived step count and dtau (main.cpp:531-534,
 * Evolution.cpp:396) */
void orc_time_steps(double L, int N, double maxtime, double *dtau, int *itmax);

#ifdef __cplusplus
}
#endif
#endif
