/*
 * ipg.h -- C ABI of the MI355X-native IP-Glasma solver core.
 *
 * The reference (schenke/ipglasma) has no FFI: its drop-in seam is the C++
 * call surface driven by main.cpp:129-360 plus the `key value ... EndOfFile`
 * input file.  Each entry point below names the reference interface it
 * replaces.  All functions are extern "C", take plain pointers and sizes,
 * never exit(), and return IPG_OK (0) or a negative status; the message of
 * the last failure on the calling thread is ipg_last_error().
 *
 * Host-visible layouts (caller-owned buffers):
 *   site index   s = ix*N + iy            (Lattice.cpp:19-28: pos = i*N + j)
 *   3x3 complex  18 doubles per site, row-major, interleaved re,im
 *                (Matrix.h:45 e[j+3i]), i.e. field arrays are [S][18]
 *   T^{mu nu}    [10][S] doubles, order tautau, xx, yy, etaeta, taux, tauy,
 *                taueta, xy, xeta, yeta (Cell.h:42-51), physical units fm^-4
 *                (taueta, xeta, yeta: fm^-5; etaeta: fm^-6)
 * Device memory is structure-of-arrays FP64 and never exposed.
 */
#ifndef IPG_H
#define IPG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPG_OK 0
#define IPG_ERR_ARG -1       /* bad argument / unsupported option          */
#define IPG_ERR_IO -2        /* input or table file missing or malformed   */
#define IPG_ERR_DEVICE -3    /* HIP runtime error or no device             */
#define IPG_ERR_STATE -4     /* call out of order                          */
#define IPG_ERR_NOMEM -5

/* Parsed input file: the keys of readInput (main.cpp:448-578) that the hot
 * path reads, with the same names.  dtau and itmax are derived exactly as
 * main.cpp:531-534 and Evolution.cpp:396 derive them. */
typedef struct ipg_params {
    int size, Nc, Ny;
    double L, g, m, UVdamp, BG, QsmuRatio, rmax, roots, Rapidity;
    double xFromThisFactorTimesQs, SigmaNN, maxtime;
    int inverseQsForMaxTime, useFluctuatingx, usePseudoRapidity, useNucleus;
    int useSmoothNucleus, useConstituentQuarkProton, smearQs, runningCoupling;
    int gaussianWounding, nucleonPositionsFromFile, rotateReactionPlane;
    int samplebFromLinearDistribution, setWSDeformParams, force_dmin_flag;
    int useFixedNpart, averageOverThisManyNuclei, useSeedList, useTimeForSeed;
    int computeGluonMultiplicity, writeOutputs, readInitialWilsonLines;
    int writeInitialWilsonLines, SubNucleonParamType;
    double R_WS, a_WS, beta2, beta3, beta4, gamma, d_min, dR_np, da_np;
    double bmin, bmax, protonAnisotropy, minimumQs2ST, smearingWidth;
    double Jacobianm, g2mu;
    unsigned long long seed;
    char Projectile[16], Target[16];
    char NucleusQsTableFileName[512];
    /* derived */
    double dtau; /* lattice units */
    int itmax;
    /* project options (not in the reference input; default 0) */
    int rho_rng; /* 0: device Philox counter map, 1: the reference's
                    MT19937-64 Gauss stream drawn on the host in its order */
    int newton_polish; /* extra Newton iterations of the tau = 0 root finder
                          after the reference's stopping rule holds (0, the
                          default, stops where Init.cpp:3384 stops); a
                          test-only knob that takes every link to round-off
                          so that difference quotients see no stopping noise */
    /* the hydro output grid (main.cpp:477-487): cells per side, rapidity
     * cells, side length [fm], rapidity step; used by the drop-in driver's
     * writeOutputs files */
    int sizeOutput, etaSizeOutput;
    double LOutput, detaOutput;
    /* the running-coupling keys (main.cpp:476-477, 540-541) that the last
     * column of the NpartdNdy file evaluates */
    double muZero, c, runWithThisFactorTimesQs;
} ipg_params;

typedef struct ipg_observables {
    int success;       /* Init::init success flag (Init.cpp:1614-1618)      */
    int Npart, Ncoll;
    double b, phiRP;   /* impact parameter [fm], reaction plane angle       */
    double averageQs, averageQsAvg, averageQsmin, Tpp, Qs2minST, alphas;
    double dEdeta;     /* tau a^2 hbarc sum_x T^tautau(x)  [GeV]             */
    double gauss_max;  /* checkGaussLaw maximum (Evolution.cpp:295-364)     */
    double tau_fm;     /* time of the T^{mu nu} evaluation [fm]             */
    long long newton_iterations; /* summed over links                      */
    int newton_max_iterations, newton_restarts, newton_unconverged;
    int newton_analytic_jacobian;
    int qs_serial_fallback;
    uint64_t rho_seed;    /* key of the device colour-charge normals        */
    uint64_t newton_seed; /* key of the root-finder restart draws           */
    /* Evolution::eccentricity(lat, param, itmax, 0.0, 0) on the Landau-frame
     * eps and u^tau (Evolution.cpp:1280-1810, gfactor = 1): eps_1..eps_6 and
     * Psi_1..Psi_6 as written to eccentricities<id>.dat; param->setArea and
     * the Rbar, <eps>, <Q_sA^2 Q_sB^2>, sqrt<r^2> columns of that file */
    double eccentricity[6];
    double Psi[6];
    double area;          /* fm^2 of the sites with eps >= 0               */
    double Rbar;
    double avg_eps;       /* GeV/fm^3                                      */
    double avg_qs2A_qs2B; /* GeV^4 * fm (the reference's units)            */
    double rms_r;         /* fm                                            */
    double centroid[2];   /* eps u^tau weighted <x>, <y> [fm]              */
    /* MyEigen::flowVelocity4D printout: sqrt(sum u_x^2 eps / sum eps),
     * sqrt(sum u_y^2 eps / sum eps), sqrt(sum (tau u^eta)^2 eps / sum eps) */
    double u_rms[3];
    int landau_failures;    /* sites whose 4x4 QR iteration did not converge */
    int landau_no_timelike; /* sites without a timelike eigenvector          */
} ipg_observables;

typedef struct ipg_ctx ipg_ctx;
typedef struct ipg_batch ipg_batch;

/* ---- configuration (Setup.cpp:24-104, main.cpp:448-578) ----------------- */
/* Fill defaults (the values of the reference's `input` file). */
void ipg_default_params(ipg_params *p);
/* Parse a reference-format input file: whitespace separated key/value pairs
 * up to EndOfFile; integer keys truncate (Setup::IFind), the seed rounds
 * (Setup::ULLIFind).  Keys absent from the file keep their current value in
 * *p, unlike the reference, which exits on a missing key. */
int ipg_read_input(const char *path, ipg_params *p);
/* Recompute dtau and itmax from L, size, maxtime (main.cpp:531-534). */
void ipg_derive(ipg_params *p);
const char *ipg_last_error(void);
int ipg_version(void);
/* Per-event device memory, the reverse-mode workspace and a batch's
 * gradients are kept, per device and size, when a context or batch is
 * destroyed (up to four fifths of the device's memory; an allocation that
 * fails hands them back first) and handed, cleared, to the next one that
 * needs that size.  This frees the kept blocks.  No counterpart in the
 * reference. */
int ipg_release_cached_memory(void);
/* sizeof(ipg_params) (which=0), sizeof(ipg_observables) (1) or
 * sizeof(ipg_multiplicity) (2), so that bindings can check their struct
 * layout */
int ipg_struct_size(int which);

/* ---- one event: Init(nn) + Evolution(nn) + Lattice (main.cpp:167-187) --- */
ipg_ctx *ipg_create(const ipg_params *p);
void ipg_destroy(ipg_ctx *c);
/* Q_s^2(T_p, y) table (Init::readNuclearQs, Init.cpp:441-480): 200x44 rows
 * of "y T Qs2" */
int ipg_load_qs_table(ipg_ctx *c, const char *path);
/* Random::init_genrand64(seed) of the event's host generator (main.cpp:98);
 * also keys the device counter-based streams */
int ipg_seed(ipg_ctx *c, uint64_t seed);

/* Init::init (Init.cpp:2303-2696): b, nucleons, colour charge, success
 * test, Wilson lines, links, tau=0 root finder, pi.  *success receives the
 * reference's success flag; on 0 the caller retries (main.cpp:339-360). */
int ipg_init(ipg_ctx *c, int *success);

/* ---- Wilson-line files: the writeInitialWilsonLines output of setV
 * (Init.cpp:1857-1980) and Init::readV (Init.cpp:1988-2301) ------------
 * V is [N*N][18]: site s = i*N + j (i along x), entry r*3 + c as (re, im).
 * format 1, text: "ix iy" and the matrix column by column (%.15g), one site
 * per line, an empty line after each x row.  format 2, binary: int N,
 * int Nc, double L, double a, double rapidity, then (re, im) doubles with
 * the site index transposed (file position p*N + q holds site (q, p)).
 * A text file read back gives the transposed matrices (the reader takes
 * them row by row, as the reference's does); binary files round-trip
 * exactly.  The reference names the files V-<eventId + 2 seed mpiSize>
 * (nucleus A) and V-<eventId + (1 + 2 seed) mpiSize> (B), ".txt" for text. */
int ipg_write_wilson_line_file(const char *path, int N, double L, double rapidity, const double *V,
                               int format);
/* nucleus 0 (A: rows shifted by -b/2) or 1 (B: +b/2); sites the file does
 * not reach keep what V holds.  Binary: N must match, |L - L_file| <= 1e-5. */
int ipg_read_wilson_line_file(const char *path, int N, double L, double b, int nucleus, int format,
                              double *V);
/* the event's V_A and V_B after ipg_init_wilson_lines */
int ipg_write_wilson_lines(ipg_ctx *c, const char *path_a, const char *path_b, int format);
/* Init::init with READFROMFILE > 0 (Init.cpp:2319-2369), in place of
 * ipg_init_nucleons, the colour charge and setV: b from the event's first
 * draw, no nucleons, V_A and V_B from the files (sites the shifted rows do
 * not reach hold the identity); then ipg_init_fields.  ipg_init refuses a
 * context whose parameters set readInitialWilsonLines > 0. */
int ipg_read_wilson_lines(ipg_ctx *c, const char *path_a, const char *path_b, int format);
/* Init::init split into its stages (same order, for drop-in callers and
 * parity checks) */
int ipg_init_nucleons(ipg_ctx *c);      /* b, sampleTA, reaction plane, wounding */
int ipg_init_color_charge(ipg_ctx *c, int *success); /* T_p, g2mu2, stats */
int ipg_init_wilson_lines(ipg_ctx *c);  /* setV */
int ipg_init_fields(ipg_ctx *c);        /* links, root finder, pi */

/* Evolution::run (Evolution.cpp:366-657): leapfrog to itmax, T^{mu nu} at
 * itmax, final half step, Gauss law, dE/deta */
int ipg_evolve(ipg_ctx *c);
/* single schedule pieces: step 0 is the tau=0 phi/U kick, steps 1..itmax-1
 * are full leapfrog steps, step itmax evaluates T^{mu nu} and then takes the
 * half step and the Gauss-law check */
int ipg_evolve_step(ipg_ctx *c, int it);

int ipg_get_observables(ipg_ctx *c, ipg_observables *o);

/* Reverse mode: dE/deta at tau_max and its gradient with respect to the
 * tau = 0 fields that ipg_init leaves (the state Evolution::run starts from,
 * Evolution.cpp:366-431).  Runs the tau = 0 kick and the itmax-1 leapfrog
 * steps forward with checkpoints every ceil(sqrt(itmax-1)) steps, evaluates
 * T^{mu nu} at itmax, and walks the steps back (adjoint of each update,
 * forward states recomputed per segment).  grad (or NULL) receives
 * [6][S][18]: the cotangents dL/dRe + i dL/dIm of U_x, U_y, E1, E2, phi, pi,
 * the last four projected on Hermitian matrices.  Call after ipg_init in
 * place of ipg_evolve: the state is consumed.  No counterpart in the
 * reference (its evolution has no derivative); the chain from the tau = 0
 * fields back to rho (root finder, Wilson lines) is not part of it. */
int ipg_adjoint_dedeta(ipg_ctx *c, double *dedeta, double *grad);

/* Reverse mode to the colour charges: d(dE/deta)/d rho^a_k(x) for both
 * nuclei, grad[2][Ny][S][8] (nucleus, slice, site, colour), through the
 * evolution (as ipg_adjoint_dedeta), pi at tau = 0, the root finder by
 * implicit differentiation (Init.cpp:3344-3481), the links, the ordered
 * product of slice exponentials and the Poisson solve (Init.cpp:1678-1850).
 * rho = g sqrt(g2mu2/Ny) z with z the normals.  Call after ipg_init in place
 * of ipg_evolve: the state is consumed.  With grad NULL the gradient stays in
 * device memory owned by the context (ipg_get_rho_grad copies it out); the
 * context keeps its reverse-mode workspace (checkpoints and cotangents) for
 * the next call.  No counterpart in the reference. */
int ipg_adjoint_rho(ipg_ctx *c, double *dedeta, double *grad);

/* Reverse mode with a choice of loss (all at tau_max, from T^{mu nu} at
 * itmax as Evolution.cpp:406-431 takes it):
 *   IPG_LOSS_DEDETA  dE/deta (as ipg_adjoint_dedeta / ipg_adjoint_rho)
 *   IPG_LOSS_ECC2    eps_2 of the Landau-frame weights eps u^tau
 *                    (Evolution.cpp:1280-1660 with MyEigen.cpp:25-321)
 *   IPG_LOSS_TMUNU   sum_s sum_k weights[k] T_k(s) over the ten stored
 *                    components (T's order: tautau, xx, yy, etaeta, taux,
 *                    tauy, taueta, xy, xeta, yeta; weights has 10 entries)
 * wrt IPG_WRT_FIELDS: grad[6][S][18] as ipg_adjoint_dedeta; IPG_WRT_RHO:
 * grad[2][Ny][S][8] as ipg_adjoint_rho (NULL keeps it on the device);
 * IPG_WRT_TMUNU: grad[10][S], the loss's gradient with respect to the ten
 * stored T^{mu nu} components at itmax (no reverse evolution).
 * value receives the loss.  Call after ipg_init in place of ipg_evolve.
 * Landau-frame branches that set a constant (rest frame, eps = T^tautau
 * fallbacks) are differentiated as constants.  No counterpart in the
 * reference. */
#define IPG_LOSS_DEDETA 0
#define IPG_LOSS_ECC2 1
#define IPG_LOSS_TMUNU 2
#define IPG_WRT_FIELDS 0
#define IPG_WRT_RHO 1
#define IPG_WRT_TMUNU 2
int ipg_adjoint_loss(ipg_ctx *c, int loss, const double *weights, int wrt, double *value, double *grad);

/* Steps between the reverse mode's checkpoints.  0 (the default) chooses by
 * memory: every state is stored (K = 1, the full tape, nothing recomputed)
 * when that fits in the free device memory with 4 GiB to spare, else
 * K = ceil(sqrt(itmax-1)) (about 2 sqrt(T) states, each segment recomputed
 * once), else the K in between that fits.  K > 0 is used as given.
 * Gradients do not depend on it.  No counterpart in the reference. */
int ipg_set_checkpoint_interval(ipg_ctx *c, int K);

/* on != 0: the Wilson lines of the following ipg_init (or
 * ipg_init_wilson_lines) keep the filtered colour fields of every slice
 * (2 Ny x 4 complex planes of N^2, in a device block of their own), so that
 * the next reverse mode to rho (ipg_adjoint_rho, ipg_adjoint_loss with
 * IPG_WRT_RHO) takes them instead of recomputing the forward row and
 * column transforms of setV; that reverse pass overwrites them.  Setting a
 * field with ipg_set_field or ipg_read_wilson_lines drops them.  on = 0
 * releases the block.  Gradients do not depend on it.  No counterpart in
 * the reference. */
int ipg_set_keep_slices(ipg_ctx *c, int on);
/* the checkpoint interval the last reverse pass of the context used */
int ipg_get_checkpoint_interval(ipg_ctx *c, int *K);

/* Host copy of the last ipg_adjoint_rho gradient, grad[2][Ny][S][8].
 * IPG_ERR_STATE when there is none.  No counterpart in the reference. */
int ipg_get_rho_grad(ipg_ctx *c, double *grad);

/* Autotuner state: one line per tuned kernel ("key: cand=best time ...
 * chosen=k").  The variants of a key compute identical results (launch shape
 * or register budget only); each is timed on the real workload until the
 * fastest is known (IPG_AUTOTUNE=0 in the environment pins the first one).
 * Writes at most len bytes (NUL-terminated) and returns the full length.
 * No counterpart in the reference. */
int ipg_autotune_report(char *buf, int len);

/* Landau frame after ipg_evolve (MyEigen::flowVelocity4D, MyEigen.cpp:25-321,
 * called from Evolution.cpp:409 at itmax): out[15][S] = eps [fm^-4], u^tau,
 * u^x, u^y, u^eta [fm^-1], pi^{mu nu} in T's order (tautau, xx, yy, etaeta,
 * taux, tauy, taueta, xy, xeta, yeta) -- the cell fields epsilon, utau, ux,
 * uy, ueta, pi* of Cell.h */
int ipg_get_landau(ipg_ctx *c, double *out);
/* Recompute the Landau frame and the eccentricities at step it >= 1
 * (tau = it dtau a) from T[10][S] (or from the event's current T^{mu nu} when
 * T is NULL); updates the observables.  For callers that hold T^{mu nu}
 * themselves (Evolution::u + Evolution::eccentricity on given cells). */
int ipg_landau(ipg_ctx *c, const double *T, int it);

/* ipg_landau without a context and without device allocation: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  T [10][S] (T^{mu nu} in ipg_get_tmunu's order and units),
 * g2mu2A and g2mu2B [S] (or both NULL: <Q_sA^2 Q_sB^2> is then 0) and land
 * [15][S] (ipg_get_landau's layout) are DEVICE pointers; p gives size, L,
 * maxtime (for dtau) and QsmuRatio; it >= 1 is the step (tau = it dtau a).
 * Writes the Landau frame to land and the Landau-frame and eccentricity
 * fields of *o (eccentricity, Psi, area, Rbar, avg_eps, avg_qs2A_qs2B, rms_r,
 * centroid, u_rms, landau_failures, landau_no_timelike, tau_fm); leaves its
 * other fields as they are.  The work space (per-block partial sums, per-site
 * flags and running eps, 13.6 MB) is one static device array per GPU, so
 * size <= 1024 (IPG_ERR_ARG beyond that), and calls that overlap on one GPU
 * race.  Runs on the default stream and returns when the results are there. */
int ipg_landau_noalloc(const ipg_params *p, const double *T, const double *g2mu2A, const double *g2mu2B, int it,
                       double *land, ipg_observables *o);

/* eccentricities<id>.dat (Evolution.cpp:1280-1810, cutoff 0, doAniso 0):
 * appends one line for the Landau frame of step it (tau = it a dtau): tau,
 * eps_n and Psi_n for n = 1..6, cutoff, sqrt<r^2>, maxX, maxY, b, Tpp, area,
 * Rbar, <eps>, <Q_sA^2 Q_sB^2> hbarc, printed as an ostream prints them
 * (%g).  The reference writes it when computeGluonMultiplicity is set. */
int ipg_write_eccentricities(ipg_ctx *c, const char *path, int it);
/* epsilon-u-Hydro-t<tau>-<id>.dat (MyEigen.cpp:326-800; writeOutputs odd,
 * gfactor 1): the Landau frame of step it bilinearly interpolated onto a
 * size_out x size_out x eta_size_out hydro grid of length L_out and rapidity
 * step deta_out (the input keys sizeOutput, etaSizeOutput, LOutput,
 * detaOutput); columns eta x y eps u^tau u^x u^y u^eta and the ten pi^{mu nu}
 * (tau tau, tau x, tau y, tau eta, xx, xy, x eta, yy, y eta, eta eta), as an
 * ostream prints them.  *Etot (may be NULL): the reference's printed Etot. */
int ipg_write_hydro(ipg_ctx *c, const char *path, int it, int size_out, int eta_size_out, double L_out,
                    double deta_out, double *Etot);
/* Tmunu-t<tau>-<id>.dat (MyEigen.cpp:801-1046; writeOutputs / 4 == 1):
 * T^{mu nu} at itmax interpolated as above onto the size_out^2 grid, y rows
 * outer; columns ix iy T^tt T^xx T^yy tau^2 T^etaeta -T^tx -T^ty -tau T^teta
 * -T^xy -tau T^yeta -tau T^xeta in GeV/fm^3 (a 1e-16 floor line where
 * T^tt <= 1e-16 GeV/fm^3 or outside the lattice) */
int ipg_write_tmunu(ipg_ctx *c, const char *path, int it, int size_out, int eta_size_out, double L_out,
                    double deta_out);
/* the same from a host T [10][N*N] (ipg_get_tmunu's layout) */
int ipg_write_tmunu_file(const char *path, const double *T, int N, double L, int it, double dtau, int size_out,
                         int eta_size_out, double L_out, double deta_out);
/* Jazma-Hydro-t<tau>-<id>.dat (MyEigen.cpp:1047-1293; (writeOutputs % 4) / 2
 * == 1): g2mu2A g2mu2B interpolated onto the hydro grid and scaled to
 * integrate to Etot (ipg_write_hydro's; the reference passes 0 when the hydro
 * file is not written), in the hydro file's columns with u^tau = 1 */
int ipg_write_jazma(ipg_ctx *c, const char *path, int it, int size_out, int eta_size_out, double L_out,
                    double deta_out, double Etot);
int ipg_write_jazma_file(const char *path, const double *g2mu2A, const double *g2mu2B, int N, double L, int it,
                         double dtau, int size_out, int eta_size_out, double L_out, double deta_out, double Etot);
/* the same from a host Landau frame land [15][N*N] (ipg_get_landau's layout) */
int ipg_write_hydro_file(const char *path, const double *land, int N, double L, int it, double dtau, int size_out,
                         int eta_size_out, double L_out, double deta_out, double *Etot);
/* the maxX, maxY columns (Evolution.cpp:1492-1640) for eps, u^tau [N*N]
 * (site ix*N + iy) and the centroid (avx, avy) of the first pass */
int ipg_eccentricity_extents(const double *eps, const double *utau, int N, double L, double avx, double avy,
                             double cutoff, double *maxX, double *maxY);

/* ---- Coulomb gauge and gluon spectra (Evolution::multiplicity) -------- */
typedef struct ipg_multiplicity {
    double dNdy, dEdy;           /* sums over the modes: "dN/dy 1", "dE/dy 1" [GeV] */
    double dNdy_kt, dEdy_kt;     /* from the k_T histogram: "dN/dy 2", "dE/dy 2"    */
    double dNdy_cut3, dEdy_cut3; /* histogram above k_T = 3 GeV                     */
    double dNdy_cut6, dEdy_cut6; /* above 6 GeV                                     */
    double mean_pt;              /* gluon <p_T> = dEdy / dNdy [GeV]                 */
    double kt_bin;               /* bin width 2.83/100 in lattice units             */
    double n_kt[100], E_kt[100]; /* per-bin spectra n[ik], E[ik] after averaging     */
    int counts[100];             /* modes per bin                                   */
    int gauge_iterations;        /* gauge-fixing steps applied                      */
    double gauge_residual;       /* <tr(chi^+ chi)/Nc> when it stopped             */
    double tau_fm;
} ipg_multiplicity;
/* GaugeFix::FFTChi (GaugeFix.cpp:17-165): Fourier-accelerated steepest
 * descent to transverse Coulomb gauge on the current fields (periodic
 * neighbours), at most max_iterations steps, stopping when the residual is
 * below 1e-9 or rises again below 1e-6 (checked against its value every
 * tenth step); the links, E, phi and pi are transformed in place. */
int ipg_gauge_fix(ipg_ctx *c, int max_iterations, int *iterations, double *residual);
/* Evolution::multiplicity (Evolution.cpp:2284-3188) at step it (1 or itmax
 * in the reference): gauge fix (4000 steps at most), then the gluon number
 * and energy per unit rapidity from the transverse E and from pi,
 * n(k) = 2/(omega N^2) [g^2/((it - 1/2) dtau) Re tr E^i(k)E^i(-k) + (it - 1/2)
 * dtau Re tr pi(k)pi(-k)], and their 100-bin k_T spectra (runningCoupling 0,
 * usePseudoRapidity 0). */
int ipg_gluon_multiplicity(ipg_ctx *c, int it, ipg_multiplicity *out);
/* the NpartdNdy-t<tau>-<id>.dat line (Evolution.cpp:3153-3177): Npart
 * dN/dy Tpp b dE/dy seed N/A N/A N/A and the k_T > 3 and > 6 GeV cuts, then
 * the last column's coupling expression, as an ostream prints them */
int ipg_write_multiplicity(ipg_ctx *c, const char *path, const ipg_multiplicity *m);

/* field ids for ipg_get_field / ipg_set_field */
#define IPG_F_UX 0
#define IPG_F_UY 1
#define IPG_F_E1 2
#define IPG_F_E2 3
#define IPG_F_PHI 4
#define IPG_F_PI 5
#define IPG_F_VA 6      /* Wilson lines after setV (until fields are built) */
#define IPG_F_VB 7
#define IPG_F_G2MU2A 8  /* scalar fields: [S] doubles */
#define IPG_F_G2MU2B 9
#define IPG_F_TPA 10
#define IPG_F_TPB 11
/* current state (after the last completed step) as [S][18] or [S] */
int ipg_get_field(ipg_ctx *c, int which, double *out);
int ipg_set_field(ipg_ctx *c, int which, const double *in);
int ipg_get_tmunu(ipg_ctx *c, double *out);
/* nucleons after the +-b/2 shift: per nucleon x, y, z, phi, collided,
 * proton; A1 rows of nucleus A then A2 rows of nucleus B */
int ipg_get_nucleons(ipg_ctx *c, double *out, int *A1, int *A2);
int ipg_set_nucleons(ipg_ctx *c, const double *in, int A1, int A2);
/* per-link root-finder iteration counts [2S] (x links then y links) */
int ipg_get_newton_iterations(ipg_ctx *c, int *out);

/* colour-charge normals: inject z[2][Ny][S][8] (host copy; NULL clears), or
 * read back the device counter-map normals of slice (nucleus, k) as [S][8] */
int ipg_set_rho_normals(ipg_ctx *c, const double *z);
int ipg_get_rho_normals(ipg_ctx *c, int nucleus, int k, double *out);

/* GPU time of the last ipg_evolve in ms: out[0] total of the leapfrog
 * launches, out[1] the T^{mu nu} launch, out[2] number of timed step
 * launches.  Timing is recorded only after ipg_enable_timing(c, 1). */
int ipg_enable_timing(ipg_ctx *c, int on);
int ipg_get_timing(ipg_ctx *c, double *out);

/* ---- many events in one launch sequence (batched form) ----------------- */
/* nev events, each with its own parameters and seed; lattice sizes and
 * collision systems may differ.  Every kernel launch covers all events. */
ipg_batch *ipg_batch_create(const ipg_params *p, int nev, const uint64_t *seeds);
void ipg_batch_destroy(ipg_batch *b);
int ipg_batch_load_qs_table(ipg_batch *b, const char *path);
/* full Init::init + Evolution::run for all events; events that fail the
 * success test are re-drawn from their generator until they succeed or
 * max_retries is reached.  use_graph: capture the evolution launches in a
 * HIP graph. */
int ipg_batch_run(ipg_batch *b, int use_graph, int max_retries);
int ipg_batch_get_observables(ipg_batch *b, int ev, ipg_observables *o);

/* Reverse mode over a batch (no counterpart in the reference): Init::init
 * for every event in one launch sequence (failures re-drawn up to
 * max_retries, as ipg_batch_run does), then for each event the loss at
 * tau_max (IPG_LOSS_DEDETA, IPG_LOSS_ECC2, or IPG_LOSS_TMUNU with ten
 * weights) and its gradient with respect to rho, through the evolution,
 * pi, the root finder, the links and setV as ipg_adjoint_rho.  The events
 * go one after the other through one reverse-mode workspace, whose
 * checkpoint interval is chosen for the device memory left after the
 * per-event gradients (see ipg_set_checkpoint_interval; the batch's own
 * interval is set with ipg_batch_set_checkpoint_interval).  Each event's
 * gradient [2][Ny][S][8] stays in device memory (ipg_batch_get_rho_grad
 * copies it out).  values[nev] receive the losses (NaN for events that
 * never passed the success test). */
int ipg_batch_adjoint(ipg_batch *b, int loss, const double *weights, int max_retries, double *values);
int ipg_batch_get_rho_grad(ipg_batch *b, int ev, double *grad);
int ipg_batch_set_checkpoint_interval(ipg_batch *b, int K);
/* the largest device memory in use (all of the device, bytes) seen during
 * the last ipg_batch_adjoint, and the checkpoint interval it used */
int ipg_batch_adjoint_info(ipg_batch *b, unsigned long long *peak_bytes, int *K);
/* wall time of the last reverse-mode call's phases, ms (summed over the
 * events of a batch): [0] workspace allocation, [1] forward pass with its
 * checkpoints and the loss, [2] reverse leapfrog, [3] reverse initial
 * conditions (pi, root finder, links, Wilson lines), [4] the batch's initial
 * conditions (0 for a context).  No counterpart in the reference. */
int ipg_adjoint_phases(ipg_ctx *c, double *ms);
int ipg_batch_adjoint_phases(ipg_batch *b, double *ms);

#ifdef __cplusplus
}
#endif
#endif
