/* orc_glauber.c -- oracle nucleons, T_p, Q_s^2 and g^2mu^2.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Follows, in draw order:
 *   Init.cpp:2320-2345   b = sqrt((bmax^2-bmin^2) xb + bmin^2) (or linear in xb)
 *   Init.cpp:65-439      sampleTA: p at the origin, the deuteron (90-152),
 *                        Woods-Saxon nuclei, then rotate_nucleus_3D on A and B
 *   Glauber.cpp:1188-1243 SampleTARejection (deuteron p-n separation) with the
 *                        Hulthen thickness of Glauber.cpp:513-576, 847-1100
 *   Init.cpp:2721-3037   generate_nucleus_configuration{,_with_woods_saxon,
 *                        _with_deformed_woods_saxon{,2,_force_dmin}}
 *   Init.cpp:2818-2831   sample_r_from_woods_saxon
 *   Init.cpp:3084-3103   recenter_nucleus;  3154-3178 rotate_nucleus_3D
 *   Init.cpp:868-901     reaction plane angle and +-b/2 shift, anisotropy phi
 *   Init.cpp:3189-3205   sampleQsNormalization (smearQs)
 *   Init.cpp:1117-1216   wounding (Gaussian or hard sphere), N_coll, N_part
 *   Init.cpp:1009-1115   T_p sums
 *   Init.cpp:686-749     getNuclearQs2
 *   Init.cpp:1219-1421   Q_s^2 -> g^2mu^2 with the fluctuating-x fixed point
 *   Init.cpp:1423-1618   statistics and success test
 * The fixed point carries `localrapidity` (and QsA/QsB) from one site to the
 * next within an OpenMP thread; the oracle uses the single-thread order
 * (ix-major, iy-minor), the reference's result with OMP_NUM_THREADS=1.
 */
#include <math.h>
#include <stdlib.h>

#include "ipg_oracle.h"

static const double HBARC = 0.1973269718; /* Phys_consts.h:6 */

/* ---- deuteron ------------------------------------------------------------ */
/* the Hulthen density (1/r^2)(e^{-a r} - e^{-w r})^2, normalised to 2
 * nucleons, integrated over z in xi = e^{-a z} as the reference writes the
 * integrand (Glauber.cpp:874-903); s in fm */
typedef struct {
    double a, w, rho, sigma, s;
} hulthen_t;

static double hulthen_f(const hulthen_t *h, double xi) {
    if (xi == 0.0) xi = 1.0e-10;
    if (xi == 1.0) xi = 1.0 - 1.0e-10;
    double sa = h->s * h->a;
    double z = -log(xi);
    double r = sqrt(sa * sa + z * z);
    double v = 2.0 * h->a * h->rho * h->sigma;
    double g = (1.0 / r) * (exp(-r) - exp(-(h->w / h->a) * r));
    v *= g * g;
    return v;
}

/* Glauber::qnc7 (Glauber.cpp:946-1100) */
static double qnc7(const hulthen_t *h, double tol, double down, double dx, const double *f_of, double pre_sum,
                   double area, int *count) {
    static const double w[7] = {41.0 / 140.0, 54.0 / 35.0, 27.0 / 140.0, 68.0 / 35.0,
                                27.0 / 140,   54.0 / 35.0, 41.0 / 140.0};
    double fl[7], fr[7], left_sum = 0.0, right_sum = 0.0, ans;
    dx /= 2.0;
    fl[1] = hulthen_f(h, down + dx);
    fl[3] = hulthen_f(h, down + 3.0 * dx);
    fl[5] = hulthen_f(h, down + 5.0 * dx);
    fl[0] = f_of[0];
    fl[2] = f_of[1];
    fl[4] = f_of[2];
    fl[6] = f_of[3];
    *count += 3;
    for (int i = 0; i < 7; i++) left_sum += w[i] * fl[i];
    left_sum *= dx;
    fr[1] = hulthen_f(h, down + 7.0 * dx);
    fr[3] = hulthen_f(h, down + 9.0 * dx);
    fr[5] = hulthen_f(h, down + 11.0 * dx);
    fr[0] = f_of[3];
    fr[2] = f_of[4];
    fr[4] = f_of[5];
    fr[6] = f_of[6];
    *count += 3;
    for (int i = 0; i < 7; i++) right_sum += w[i] * fr[i];
    right_sum *= dx;
    ans = left_sum + right_sum;
    area += -fabs(pre_sum) + fabs(left_sum) + fabs(right_sum);
    if (fabs(ans - pre_sum) > tol * fabs(area) && (*count < 10000)) {
        tol /= 1.414213562;
        left_sum = qnc7(h, tol, down, dx, fl, left_sum, area, count);
        right_sum = qnc7(h, tol, down + dx * 6., dx, fr, right_sum, area, count);
        ans = left_sum + right_sum;
    }
    return ans;
}

/* InterNuPInSP / InterNuTInST: the table of NuInS over [0, 2 SCutOff] with
 * InterMax = 100 intervals (initGlauber, Glauber.cpp:1135-1180), cubic
 * through four nodes (FourPtInterpolate, LinearFindXorg) */
enum { DT_IMAX = 100 };
/* the last table built (the sampler calls this per trial); not thread-safe,
 * like the sampler's generator */
static struct {
    int valid;
    double a, w, R0, sigma;
    double vx[DT_IMAX + 1], vy[DT_IMAX + 1];
} dtab;

double orc_deuteron_thickness(double s, double a, double wr, double R0, double sigma_fm2) {
    enum { IMAX = DT_IMAX };
    double up = 2.0 * 12.0, dxt = (up - 0.0) / IMAX;
    if (!(dtab.valid && dtab.a == a && dtab.w == wr && dtab.R0 == R0 && dtab.sigma == sigma_fm2)) {
        double f0n = 2.0 * (a - wr) * (a - wr) * M_PI / a / wr / (a + wr);
        double f = (f0n * R0) / R0; /* AnumHulthen() / rho_WS with rho_WS = R_WS */
        hulthen_t h = {a, wr, 2.0 / f, sigma_fm2, 0.};
        for (int i = 0; i <= IMAX; i++) {
            dtab.vx[i] = dxt * i;
            h.s = dtab.vx[i];
            double g1[7], dx = (1.0 - 0.0) / 6.0;
            for (int k = 0; k < 7; k++) g1[k] = hulthen_f(&h, 0.0 + k * dx);
            int count = 7;
            dtab.vy[i] = qnc7(&h, 1.0e-6, 0.0, dx, g1, 0.0, 0.0, &count);
        }
        dtab.a = a;
        dtab.w = wr;
        dtab.R0 = R0;
        dtab.sigma = sigma_fm2;
        dtab.valid = 1;
    }
    const double *vx = dtab.vx, *vy = dtab.vy;
    if (s > up) return 0.0;
    double nx = IMAX * (s - vx[0]) / (vx[IMAX] - vx[0]);
    int x_org = (int)nx - 1;
    if (x_org <= 0) x_org = 0;
    else if (x_org >= IMAX - 3) x_org = IMAX - 3;
    double hh = (vx[IMAX] - vx[0]) / IMAX;
    double f0 = vy[x_org], f1 = vy[x_org + 1], f2 = vy[x_org + 2], f3 = vy[x_org + 3];
    double ca = (-f0 + 3.0 * f1 - 3.0 * f2 + f3) / (6.0 * hh * hh * hh);
    double cb = (2.0 * f0 - 5.0 * f1 + 4.0 * f2 - f3) / (2.0 * hh * hh);
    double cc = (-11.0 * f0 + 18.0 * f1 - 9.0 * f2 + 2.0 * f3) / (6.0 * hh);
    double y = ca * pow(s - vx[x_org], 3.);
    y += cb * pow(s - vx[x_org], 2.);
    y += cc * (s - vx[x_org]);
    y += f0;
    return y < 0.0 ? 0.0 : y;
}

/* Glauber::SampleTARejection (Glauber.cpp:1188-1243) */
static void sample_deuteron(orc_mt *rng, double a, double R0, double sigma_fm2, double *x, double *y) {
    const double wr = 1.18; /* Glauber.cpp:72, w_WS of "d" */
    double A = 1.2 * sigma_fm2 / 4.21325504715;
    double phi, r, tmp;
    do {
        phi = 2. * M_PI * orc_mt_real1(rng);
        r = 6.32456 * sqrt(-log((-0.00454545 * (-220. * A + A * 220. * (1. - exp(-0.025 * 15. * 15.)) *
                                                                 orc_mt_real1(rng))) / A));
        tmp = orc_mt_real1(rng);
    } while (tmp > r * orc_deuteron_thickness(r, a, wr, R0, sigma_fm2) / (A * r * 11. * exp(-r * r / 40.)));
    *x = r * cos(phi);
    *y = r * sin(phi);
}

/* stable sort of idx by key (ascending), insertion sort */
static void stable_sort_idx(int n, const double *key, int *idx) {
    for (int i = 1; i < n; i++) {
        int v = idx[i];
        int j = i - 1;
        while (j >= 0 && key[v] < key[idx[j]]) {
            idx[j + 1] = idx[j];
            j--;
        }
        idx[j + 1] = v;
    }
}

static double fermi(double r, double R, double a) {
    return 1.0 / (1.0 + exp((r - R) / a));
}

static double sample_r_ws(orc_mt *rng, double a, double R) {
    double rmax = R + 10.0 * a;
    double r;
    do {
        r = rmax * pow(orc_mt_real3(rng), 1.0 / 3.0);
    } while (orc_mt_real3(rng) > fermi(r, R, a));
    return r;
}

static double ylm(int l, double ct) {
    if (l == 2) return (3.0 * ct * ct - 1.0) * 0.31539156525252005;
    if (l == 3) return (5.0 * ct * ct * ct - 3.0 * ct) * 0.3731763325901154;
    return (35.0 * ct * ct * ct * ct - 30.0 * ct * ct + 3.0) * 0.10578554691520431;
}
static double y22(double ct, double phi) {
    return (1.0 - ct * ct) * cos(2.0 * phi) * 0.5462742152960397;
}

static void recenter(int A, double *x, double *y, double *z) {
    double mx = 0, my = 0, mz = 0;
    for (int i = 0; i < A; i++) {
        mx += x[i];
        my += y[i];
        mz += z[i];
    }
    mx /= (double)A;
    my /= (double)A;
    mz /= (double)A;
    for (int i = 0; i < A; i++) {
        x[i] -= mx;
        y[i] -= my;
        z[i] -= mz;
    }
}

/* place sorted radii on the sphere with the d_min rejection (<=100 tries) */
static void place_with_dmin(orc_mt *rng, int A, const double *r, const double *theta_fixed,
                            double d_min, double *x, double *y, double *z) {
    const double dsq_min = d_min * d_min;
    for (int i = 0; i < A; i++) {
        double ri = r[i], xi, yi, zi;
        int iter = 0, reject;
        do {
            iter++;
            reject = 0;
            double phi = 2. * M_PI * orc_mt_real3(rng);
            double th = theta_fixed ? theta_fixed[i] : acos(1. - 2. * orc_mt_real3(rng));
            xi = ri * sin(th) * cos(phi);
            yi = ri * sin(th) * sin(phi);
            zi = ri * cos(th);
            for (int j = i - 1; j >= 0; j--) {
                if ((ri - r[j]) * (ri - r[j]) > dsq_min) break;
                double d = (xi - x[j]) * (xi - x[j]) + (yi - y[j]) * (yi - y[j]) +
                           (zi - z[j]) * (zi - z[j]);
                if (d < dsq_min) {
                    reject = 1;
                    break;
                }
            }
        } while (reject == 1 && iter < 100);
        x[i] = xi;
        y[i] = yi;
        z[i] = zi;
    }
}

static void gen_nucleus(orc_mt *rng, int A, int Z, double a_ws, double R_ws, double b2,
                        double b3, double b4, double gam, int force_dmin, double d_min,
                        double dR, double da, orc_nucleon *out) {
    double *x = calloc(A, sizeof(double)), *y = calloc(A, sizeof(double)),
           *z = calloc(A, sizeof(double));
    int *prot = calloc(A, sizeof(int));
    int spherical = fabs(b2) < 1e-15 && fabs(b4) < 1e-15 && fabs(b3) < 1e-15 &&
                    fabs(gam) < 1e-15;
    if (spherical || (!force_dmin && fabs(gam) <= 1e-15)) {
        /* Init.cpp:2749-2816 (spherical) and 2833-2904 (deformed) */
        double *r = calloc(A, sizeof(double)), *ct = calloc(A, sizeof(double));
        double *rs = calloc(A, sizeof(double)), *th = NULL;
        int *idx = calloc(A, sizeof(int));
        for (int i = 0; i < A; i++) {
            double aa = i < Z ? a_ws : a_ws + da;
            double RR = i < Z ? R_ws : R_ws + dR;
            if (spherical) {
                r[i] = sample_r_ws(rng, aa, RR);
            } else {
                double rmax = RR + 10. * aa, Rth;
                do {
                    r[i] = rmax * pow(orc_mt_real3(rng), 1.0 / 3.0);
                    ct[i] = 1.0 - 2.0 * orc_mt_real3(rng);
                    Rth = RR * (1.0 + b2 * ylm(2, ct[i]) + b3 * ylm(3, ct[i]) +
                                b4 * ylm(4, ct[i]));
                } while (orc_mt_real3(rng) > fermi(r[i], Rth, aa));
            }
            idx[i] = i;
        }
        stable_sort_idx(A, r, idx);
        for (int i = 0; i < A; i++) rs[i] = r[idx[i]];
        if (!spherical) {
            th = calloc(A, sizeof(double));
            for (int i = 0; i < A; i++) th[i] = acos(ct[idx[i]]);
        }
        place_with_dmin(rng, A, rs, th, d_min, x, y, z);
        for (int i = 0; i < A; i++) prot[i] = spherical ? 1 : (idx[i] < Z);
        free(r);
        free(ct);
        free(rs);
        free(th);
        free(idx);
    } else {
        /* Init.cpp:2906-2982 (force_dmin) and 2984-3037 (gamma != 0) */
        double rmax = R_ws + dR + 10. * (a_ws + da);
        const double dsq_min = d_min * d_min;
        for (int i = 0; i < A; i++) {
            double Ri = i >= Z ? R_ws + dR : R_ws, ai = i >= Z ? a_ws + da : a_ws;
            int resample;
            double xi, yi, zi;
            do {
                double r, ct, phi, Rth;
                do {
                    r = rmax * pow(orc_mt_real3(rng), 1.0 / 3.0);
                    ct = 1.0 - 2.0 * orc_mt_real3(rng);
                    phi = 2. * M_PI * orc_mt_real3(rng);
                    Rth = Ri * (1.0 + b2 * (cos(gam) * ylm(2, ct) + sin(gam) * y22(ct, phi)) +
                                b3 * ylm(3, ct) + b4 * ylm(4, ct));
                } while (orc_mt_real3(rng) > fermi(r, Rth, ai));
                double st = sqrt(1. - ct * ct);
                xi = r * st * cos(phi);
                yi = r * st * sin(phi);
                zi = r * ct;
                resample = 0;
                if (force_dmin) {
                    for (int j = i - 1; j >= 0; j--) {
                        double d = (xi - x[j]) * (xi - x[j]) + (yi - y[j]) * (yi - y[j]) +
                                   (zi - z[j]) * (zi - z[j]);
                        if (d < dsq_min) {
                            resample = 1;
                            break;
                        }
                    }
                }
            } while (resample);
            x[i] = xi;
            y[i] = yi;
            z[i] = zi;
            prot[i] = i < Z;
        }
    }
    recenter(A, x, y, z);
    for (int i = 0; i < A; i++) {
        out[i].x = x[i];
        out[i].y = y[i];
        out[i].z = z[i];
        out[i].phi = atan2(y[i], x[i]);
        out[i].collided = 0;
        out[i].proton = prot[i];
    }
    free(x);
    free(y);
    free(z);
    free(prot);
}

static void rotate3d(orc_mt *rng, int A, orc_nucleon *n) {
    double al = 2 * M_PI * orc_mt_real3(rng);
    double be = acos(1. - 2. * orc_mt_real3(rng));
    double ga = 2 * M_PI * orc_mt_real3(rng);
    double c1 = cos(al), s1 = sin(al), c2 = cos(be), s2 = sin(be), c3 = cos(ga), s3 = sin(ga);
    for (int i = 0; i < A; i++) {
        double x = n[i].x, y = n[i].y, z = n[i].z;
        double xn = c2 * x - c3 * s2 * y + s2 * s3 * z;
        double yn = c1 * s2 * x + (c1 * c2 * c3 - s1 * s3) * y + (-c3 * s1 - c1 * c2 * s3) * z;
        double zn = s1 * s2 * x + (c1 * s3 + c2 * c3 * s1) * y + (c1 * c3 - c2 * s1 * s3) * z;
        n[i].x = xn;
        n[i].y = yn;
        n[i].z = zn;
    }
}

int orc_sample_event(orc_mt *rng, const orc_glauber_cfg *c, orc_nucleon *nA,
                     orc_nucleon *nB, orc_event_info *info) {
    if (c->A1 == 3 || c->A2 == 3) return -1; /* He3: unsupported here */
    double xb = orc_mt_real1(rng);
    double b = c->linear_b ? sqrt((c->bmax * c->bmax - c->bmin * c->bmin) * xb + c->bmin * c->bmin)
                           : (c->bmax - c->bmin) * xb + c->bmin;
    if (c->A1 == 1) {
        nA[0] = (orc_nucleon){0, 0, 0, 0, 0, 1};
    } else if (c->A1 == 2) {
        double dx, dy;
        sample_deuteron(rng, c->a1, c->R1, 0.1 * c->sigma_nn_mb, &dx, &dy);
        nA[0] = (orc_nucleon){dx / 2., dy / 2., 0., 0, 0, 1};
        nA[1] = (orc_nucleon){-(dx / 2.), -(dy / 2.), -0., 0, 0, 1};
    } else {
        gen_nucleus(rng, c->A1, c->Z1, c->a1, c->R1, c->beta2_1, c->beta3_1, c->beta4_1,
                    c->gamma_1, c->force_dmin, c->d_min, c->dR1, c->da1, nA);
    }
    if (c->A2 == 1) {
        nB[0] = (orc_nucleon){0, 0, 0, 0, 0, 1};
    } else if (c->A2 == 2) {
        double dx, dy;
        sample_deuteron(rng, c->a2, c->R2, 0.1 * c->sigma_nn_mb, &dx, &dy);
        nB[0] = (orc_nucleon){dx / 2., dy / 2., 0., 0, 0, 1};
        nB[1] = (orc_nucleon){-(dx / 2.), -(dy / 2.), -0., 0, 0, 0};
    } else {
        gen_nucleus(rng, c->A2, c->Z2, c->a2, c->R2, c->beta2_2, c->beta3_2, c->beta4_2,
                    c->gamma_2, c->force_dmin, c->d_min, c->dR2, c->da2, nB);
    }
    rotate3d(rng, c->A1, nA);
    rotate3d(rng, c->A2, nB);

    double phiRP = 0.;
    if (c->rotate_rp) phiRP = 2 * M_PI * orc_mt_real2(rng);
    for (int i = 0; i < c->A1; i++) {
        nA[i].x -= b / 2. * cos(phiRP);
        nA[i].y -= b / 2. * sin(phiRP);
    }
    for (int i = 0; i < c->A2; i++) {
        nB[i].x += b / 2. * cos(phiRP);
        nB[i].y += b / 2. * sin(phiRP);
    }
    if (c->xi != 0.) {
        for (int i = 0; i < c->A1; i++) nA[i].phi = 2 * M_PI * orc_mt_real2(rng);
        for (int i = 0; i < c->A2; i++) nB[i].phi = 2 * M_PI * orc_mt_real2(rng);
    } else {
        for (int i = 0; i < c->A1; i++) nA[i].phi = 0.;
        for (int i = 0; i < c->A2; i++) nB[i].phi = 0.;
    }

    /* wounding */
    double d2 = c->sigma_nn_mb / (M_PI * 10.);
    int Ncoll = 0;
    for (int i = 0; i < c->A1; i++) {
        for (int j = 0; j < c->A2; j++) {
            double dx = nB[j].x - nA[i].x, dy = nB[j].y - nA[i].y;
            double dij = dx * dx + dy * dy;
            int hit;
            if (c->gaussian_wounding) {
                double G = 0.92;
                double p = G * exp(-G * dij / d2);
                hit = orc_mt_real1(rng) < p;
            } else {
                hit = dij < d2;
            }
            if (hit) {
                Ncoll++;
                nB[j].collided = 1;
                nA[i].collided = 1;
            }
        }
    }
    if (c->A1 == 1 && c->A2 == 1) {
        nA[0].collided = 1;
        nB[0].collided = 1;
    }
    int Npart = 0;
    for (int i = 0; i < c->A1; i++) Npart += nA[i].collided == 1;
    for (int i = 0; i < c->A2; i++) Npart += nB[i].collided == 1;
    info->b = b;
    info->phiRP = phiRP;
    info->Ncoll = Ncoll;
    info->Npart = Npart;
    return 0;
}

void orc_tp_sum(int N, double L, double BG, double xi, const orc_nucleon *nuc, int A,
                double *Tp) {
    double a = L / N;
#pragma omp parallel for
    for (int ix = 0; ix < N; ix++) {
        double x = -L / 2. + a * ix;
        for (int iy = 0; iy < N; iy++) {
            double y = -L / 2. + a * iy;
            double acc = 0.;
            for (int i = 0; i < A; i++) {
                double xm = nuc[i].x, ym = nuc[i].y, phi = nuc[i].phi;
                double bp2 = (xm - x) * (xm - x) + (ym - y) * (ym - y) +
                             xi * pow((xm - x) * cos(phi) + (ym - y) * sin(phi), 2.);
                bp2 /= HBARC * HBARC;
                double T = sqrt(1 + xi) * exp(-bp2 / (2. * BG)) / (2. * M_PI * BG) * 1.0;
                acc = acc + T / 1.0;
            }
            Tp[ix * N + iy] = acc;
        }
    }
}

#define ORC_NT 200
#define ORC_NY 44
double orc_nuclear_qs2(const double *Tl, const double *Q, double T, double y) {
    const double dy = 0.25;
    int posy = (int)floor(y / dy + 0.0000001);
    if (y > ORC_NY * dy) return NAN; /* the reference exits here */
    double fracy = 0.;
    int tot = ORC_NT * ORC_NY;
#define QAT(it, iy) Q[((it) * ORC_NY + (iy)) < tot ? ((it) * ORC_NY + (iy)) : tot - 1]
    if (T > Tl[ORC_NT - 1]) {
        fracy = (y - (double)posy * dy) / dy;
        return fracy * QAT(ORC_NT - 1, posy + 1) + (1. - fracy) * QAT(ORC_NT - 1, posy);
    }
    if (T < Tl[0]) return 0.;
    int check = 0;
    double value = 0.;
    for (int iT = 0; iT < ORC_NT - 1; iT++) {
        if (T >= Tl[iT] && T < Tl[iT + 1]) {
            double fracT = (T - Tl[iT]) / (Tl[iT + 1] - Tl[iT]);
            fracy = (y - (double)posy * dy) / dy;
            double dn = fracT * QAT(iT + 1, posy) + (1. - fracT) * QAT(iT, posy);
            double up = fracT * QAT(iT + 1, posy + 1) + (1. - fracT) * QAT(iT, posy + 1);
            value = fracy * up + (1. - fracy) * dn;
            check++;
        }
    }
    if (check != 1) {
        if (T == Tl[ORC_NT - 1]) fracy = (y - (double)posy * dy) / dy;
        value = fracy * QAT(ORC_NT - 1, posy + 1) + (1. - fracy) * QAT(ORC_NT - 1, posy);
    }
#undef QAT
    return value;
}

static double distance_of(double BG, double Tp, double rmax) {
    double l = log(2 * M_PI * BG * Tp);
    if (l < 0.) {
        if (isinf(l)) return rmax + 1.;
        return sqrt(-2. * BG * l) * HBARC;
    }
    return 0.;
}

void orc_g2mu2(const orc_qs_cfg *c, const double *Tl, const double *Q, const double *TpA,
               const double *TpB, double *gA, double *gB) {
    const int N = c->N;
    const double a = c->L / N;
    const double yIn = c->rapidity;
    const double expo = 5.6;
    double ly = c->rapidity; /* carried from site to site (single thread) */
    double QsA = 1, QsB = 1;
    for (int ix = 0; ix < N; ix++) {
        for (int iy = 0; iy < N; iy++) {
            int s = ix * N + iy;
            gA[s] = 0.;
            gB[s] = 0.;
            double dA = distance_of(c->BG, TpA[s], c->rmax);
            double dB = distance_of(c->BG, TpB[s], c->rmax);
            int check = 0;
            if (dA < c->rmax) check = 1;
            if (dB < c->rmax && check == 1) check = 2;
            if (check != 2) continue;
            if (c->useFluctuatingx == 1) {
                double ydev = 10000;
                for (int guard = 0; fabs(ydev) > 0.001 && guard < 100000; guard++) {
                    if (ly >= 0) {
                        QsA = sqrt(orc_nuclear_qs2(Tl, Q, TpA[s], fabs(ly)));
                    } else {
                        double xv = QsA * c->xFactor / c->roots * exp(yIn);
                        if (xv == 0)
                            QsA = 0.;
                        else
                            QsA = sqrt(orc_nuclear_qs2(Tl, Q, TpA[s], 0.)) *
                                  sqrt(pow((1 - xv) / (1 - 0.01), expo) * pow((0.01 / xv), 0.2));
                    }
                    if (QsA == 0) {
                        ydev = 0;
                        gA[s] = 0.;
                    } else {
                        gA[s] = QsA * QsA / c->QsmuRatio / c->QsmuRatio * a * a / HBARC / HBARC /
                                c->g / c->g;
                        ydev = ly - log(0.01 / (QsA * c->xFactor / c->roots * exp(yIn)));
                        ly = log(0.01 / (QsA * c->xFactor / c->roots * exp(yIn)));
                    }
                }
                if (gA[s] != gA[s]) gA[s] = 0.;
                ly = c->rapidity;
                ydev = 10000;
                for (int guard = 0; fabs(ydev) > 0.001 && guard < 100000; guard++) {
                    if (ly >= 0) {
                        QsB = sqrt(orc_nuclear_qs2(Tl, Q, TpB[s], fabs(ly)));
                    } else {
                        double xv = QsB * c->xFactor / c->roots * exp(-yIn);
                        if (xv == 0)
                            QsB = 0.;
                        else
                            QsB = sqrt(orc_nuclear_qs2(Tl, Q, TpB[s], 0.)) *
                                  sqrt(pow((1 - xv) / (1 - 0.01), expo) * pow((0.01 / xv), 0.2));
                    }
                    if (QsB == 0) {
                        ydev = 0;
                        gB[s] = 0.;
                    } else {
                        gB[s] = QsB * QsB / c->QsmuRatioB / c->QsmuRatioB * a * a / HBARC /
                                HBARC / c->g / c->g;
                        ydev = ly - log(0.01 / (QsB * c->xFactor / c->roots * exp(-yIn)));
                        ly = log(0.01 / (QsB * c->xFactor / c->roots * exp(-yIn)));
                    }
                }
                if (gB[s] != gB[s]) gB[s] = 0.;
            } else {
                gA[s] = orc_nuclear_qs2(Tl, Q, TpA[s], ly) / c->QsmuRatio / c->QsmuRatio * a *
                        a / HBARC / HBARC / c->g / c->g;
                gB[s] = orc_nuclear_qs2(Tl, Q, TpB[s], ly) / c->QsmuRatioB / c->QsmuRatioB *
                        a * a / HBARC / HBARC / c->g / c->g;
            }
        }
    }
}

void orc_stats(const orc_qs_cfg *c, double sigma_nn_mb, int Npart, double minQs2ST,
               const orc_nucleon *nA, int A1, const orc_nucleon *nB, int A2,
               const double *gA, const double *gB, const double *TpA, const double *TpB,
               double out[8]) {
    const int N = c->N;
    const double L = c->L, a = L / N, g = c->g;
    const double rA = c->QsmuRatio, rB = c->QsmuRatioB;
    double aQs = 0, aQs2 = 0, aQs2Avg = 0, aQs2min = 0, aQs2min2 = 0, Tpp = 0;
    int count = 0;
    const double rcut = sqrt(0.1 * sigma_nn_mb / M_PI);
    for (int ix = 0; ix < N; ix++) {
        for (int iy = 0; iy < N; iy++) {
            int pos = ix * N + iy, check = 0;
            double x = -L / 2. + a * ix, y = -L / 2. + a * iy;
            double g2A = (pos > 0 && pos < (N - 1) * N + N - 1) ? gA[pos] : 0;
            double g2B = (pos > 0 && pos < (N - 1) * N + N - 1) ? gB[pos] : 0;
            if (g2B >= g2A)
                aQs2min2 += g2A * rA * rA / a / a * HBARC * HBARC * g * g;
            else
                aQs2min2 += g2B * rB * rB / a / a * HBARC * HBARC * g * g;
            for (int i = 0; i < A1; i++) {
                double r = sqrt((x - nA[i].x) * (x - nA[i].x) + (y - nA[i].y) * (y - nA[i].y));
                if (r < rcut && nA[i].collided == 1) check = 1;
            }
            for (int i = 0; i < A2; i++) {
                double r = sqrt((x - nB[i].x) * (x - nB[i].x) + (y - nB[i].y) * (y - nB[i].y));
                if (r < rcut && nB[i].collided == 1 && check == 1) check = 2;
            }
            if (check == 2) {
                if (g2B > g2A) {
                    aQs += sqrt(g2B * rB * rB / a / a * HBARC * HBARC * g * g);
                    aQs2 += g2B * rB * rB / a / a * HBARC * HBARC * g * g;
                    aQs2min += g2A * rA * rA / a / a * HBARC * HBARC * g * g;
                } else {
                    aQs += sqrt(g2A * rA * rA / a / a * HBARC * HBARC * g * g);
                    aQs2 += g2A * rA * rA / a / a * HBARC * HBARC * g * g;
                    aQs2min += g2B * rB * rB / a / a * HBARC * HBARC * g * g;
                }
                aQs2Avg += (g2B * rB * rB + g2A * rA * rA) / 2. / a / a * HBARC * HBARC * g * g;
                count++;
            }
            Tpp += TpB[pos] * TpA[pos] * a * a / HBARC / HBARC / HBARC / HBARC;
        }
    }
    aQs2 /= (double)count;
    aQs2Avg /= (double)count;
    aQs2min /= (double)count;
    double Qs = sqrt(aQs2), QsAvg = sqrt(aQs2Avg), Qsmin = sqrt(aQs2min);
    double alphas = g * g / 4. / M_PI; /* fixed coupling */
    double q2st = aQs2min2 * a * a / HBARC / HBARC;
    int success = Qs > 0 && QsAvg > 0 && aQs2 > 0 && Qsmin > 0 && aQs2Avg > 0 && alphas > 0 &&
                  Npart >= 2 && q2st > minQs2ST;
    out[0] = Qs;
    out[1] = QsAvg;
    out[2] = Qsmin;
    out[3] = Tpp;
    out[4] = q2st;
    out[5] = count;
    out[6] = alphas;
    out[7] = success;
}
