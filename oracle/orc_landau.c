/* orc_landau.c -- oracle Landau frame and spatial eccentricities.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   orc_eig4          the real nonsymmetric 4x4 eigenproblem that the
 *                     reference hands to GSL's gsl_eigen_nonsymmv
 *                     (MyEigen.cpp:78-95).  GSL is not in this image; the
 *                     method is restated from its published description:
 *                     Householder reduction to Hessenberg form, Francis
 *                     double-shift QR to real Schur form (Golub & Van Loan,
 *                     Matrix Computations, Alg. 7.4.2 and 7.5.1), 2x2 blocks
 *                     in standard form (the rotation of LAPACK's dlanv2),
 *                     eigenvectors by back substitution in the quasi-
 *                     triangular factor (LAPACK dtrevc's convention for a
 *                     complex pair), each scaled to unit Euclidean norm.
 *   orc_landau        MyEigen::flowVelocity4D (MyEigen.cpp:25-321), in its
 *                     site order, with the variables that the reference
 *                     carries from one site to the next (eps, u^mu: they
 *                     enter pi^{mu nu} at a site without a timelike
 *                     eigenvector, MyEigen.cpp:267-309)
 *   orc_eccentricity  Evolution::eccentricity (Evolution.cpp:1280-1810)
 *                     with gfactor = 1 (runningCoupling = 0)
 *
 * The reference's unqualified abs() on doubles resolves to the double
 * overload: the GSL headers pull in <math.h>, whose C++ form exports
 * std::abs into the global namespace.
 */
#include <float.h>
#include <math.h>
#include <string.h>

#include "ipg_oracle.h"

static const double HBARC = 0.1973269718;

static double sgnf(double a, double b) { return b >= 0. ? fabs(a) : -fabs(a); }

/* ---- 4x4 real Schur form ------------------------------------------------- */

/* P = I - 2 v v^T / (v^T v) on rows/cols r0..r0+n-1 of H (both sides) and
 * on the columns of Q */
static void reflect(double H[4][4], double Q[4][4], int r0, int n,
                    const double *v) {
    double vtv = 0.;
    for (int i = 0; i < n; i++) vtv += v[i] * v[i];
    if (vtv == 0.) return;
    const double f = 2. / vtv;
    for (int j = 0; j < 4; j++) {
        double s = 0.;
        for (int i = 0; i < n; i++) s += v[i] * H[r0 + i][j];
        s *= f;
        for (int i = 0; i < n; i++) H[r0 + i][j] -= s * v[i];
    }
    for (int i = 0; i < 4; i++) {
        double s = 0., t = 0.;
        for (int k = 0; k < n; k++) {
            s += H[i][r0 + k] * v[k];
            t += Q[i][r0 + k] * v[k];
        }
        s *= f;
        t *= f;
        for (int k = 0; k < n; k++) {
            H[i][r0 + k] -= s * v[k];
            Q[i][r0 + k] -= t * v[k];
        }
    }
}

/* Householder vector v with (I - 2vv^T/v^Tv) x = (beta, 0, ...); returns 0
 * when x is already of that form */
static int house(const double *x, int n, double *v) {
    double tail = 0.;
    for (int i = 1; i < n; i++) tail += x[i] * x[i];
    if (tail == 0.) return 0;
    const double alpha = sqrt(x[0] * x[0] + tail);
    const double beta = -sgnf(alpha, x[0]);
    v[0] = x[0] - beta;
    for (int i = 1; i < n; i++) v[i] = x[i];
    return 1;
}

/* rotation of rows/cols (i, i+1): H <- R^T H R, Q <- Q R with
 * R = [cs -sn; sn cs] */
static void rot(double H[4][4], double Q[4][4], int i, double cs, double sn) {
    for (int j = 0; j < 4; j++) {
        const double x = H[i][j], y = H[i + 1][j];
        H[i][j] = cs * x + sn * y;
        H[i + 1][j] = cs * y - sn * x;
    }
    for (int r = 0; r < 4; r++) {
        double x = H[r][i], y = H[r][i + 1];
        H[r][i] = cs * x + sn * y;
        H[r][i + 1] = cs * y - sn * x;
        x = Q[r][i];
        y = Q[r][i + 1];
        Q[r][i] = cs * x + sn * y;
        Q[r][i + 1] = cs * y - sn * x;
    }
}

/* standard form of the 2x2 block [a b; c d] = R [a' b'; c' d'] R^T: either
 * upper triangular (real pair) or a' = d', b'c' < 0 (complex pair) */
static void std2(double *a, double *b, double *c, double *d, double *cs,
                 double *sn) {
    const double eps = DBL_EPSILON;
    if (*c == 0.) {
        *cs = 1.;
        *sn = 0.;
    } else if (*b == 0.) {
        *cs = 0.;
        *sn = 1.;
        const double t = *d;
        *d = *a;
        *a = t;
        *b = -*c;
        *c = 0.;
    } else if (*a - *d == 0. && sgnf(1., *b) != sgnf(1., *c)) {
        *cs = 1.;
        *sn = 0.;
    } else {
        double temp = *a - *d, p = 0.5 * temp;
        const double bcmax = fmax(fabs(*b), fabs(*c));
        const double bcmis = fmin(fabs(*b), fabs(*c)) * sgnf(1., *b) * sgnf(1., *c);
        const double scale = fmax(fabs(p), bcmax);
        double z = p / scale * p + bcmax / scale * bcmis;
        if (z >= 4. * eps) {
            z = p + sgnf(sqrt(scale) * sqrt(z), p);
            *a = *d + z;
            *d = *d - bcmax / z * bcmis;
            const double tau = sqrt(*c * *c + z * z);
            *cs = z / tau;
            *sn = *c / tau;
            *b = *b - *c;
            *c = 0.;
        } else {
            const double sigma = *b + *c;
            const double tau = sqrt(sigma * sigma + temp * temp);
            *cs = sqrt(0.5 * (1. + fabs(sigma) / tau));
            *sn = -(p / (tau * *cs)) * sgnf(1., sigma);
            const double aa = *a * *cs + *b * *sn, bb = -*a * *sn + *b * *cs;
            const double cc = *c * *cs + *d * *sn, dd = -*c * *sn + *d * *cs;
            *a = aa * *cs + cc * *sn;
            *b = bb * *cs + dd * *sn;
            *c = -aa * *sn + cc * *cs;
            *d = -bb * *sn + dd * *cs;
            temp = 0.5 * (*a + *d);
            *a = temp;
            *d = temp;
            if (*c != 0.) {
                if (*b != 0.) {
                    if (sgnf(1., *b) == sgnf(1., *c)) {
                        const double sab = sqrt(fabs(*b)), sac = sqrt(fabs(*c));
                        p = sgnf(sab * sac, *c);
                        const double t2 = 1. / sqrt(fabs(*b + *c));
                        *a = temp + p;
                        *d = temp - p;
                        *b = *b - *c;
                        *c = 0.;
                        const double cs1 = sab * t2, sn1 = sac * t2;
                        temp = *cs * cs1 - *sn * sn1;
                        *sn = *cs * sn1 + *sn * cs1;
                        *cs = temp;
                    }
                } else {
                    *b = -*c;
                    *c = 0.;
                    temp = *cs;
                    *cs = -*sn;
                    *sn = temp;
                }
            }
        }
    }
}

/* complex helpers for the back substitution */
typedef struct {
    double r, i;
} cz;
static cz czdiv(cz a, cz b) {
    cz q;
    if (fabs(b.r) >= fabs(b.i)) {
        const double t = b.i / b.r, den = b.r + b.i * t;
        q.r = (a.r + a.i * t) / den;
        q.i = (a.i - a.r * t) / den;
    } else {
        const double t = b.r / b.i, den = b.r * t + b.i;
        q.r = (a.r * t + a.i) / den;
        q.i = (a.i * t - a.r) / den;
    }
    return q;
}
static cz czmul(cz a, cz b) {
    cz q = {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
    return q;
}
static cz czsub(cz a, cz b) {
    cz q = {a.r - b.r, a.i - b.i};
    return q;
}
static double cz1(cz a) { return fabs(a.r) + fabs(a.i); }
/* pivot no smaller than smin (dtrevc perturbs near-singular pivots) */
static cz czpiv(cz p, double smin) {
    if (cz1(p) < smin) {
        p.r = smin;
        p.i = 0.;
    }
    return p;
}

int orc_eig4(const double *A, double *wr, double *wi, double *vr, double *vi) {
    double H[4][4], Q[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            H[i][j] = A[4 * i + j];
            Q[i][j] = i == j ? 1. : 0.;
        }
    /* Hessenberg reduction */
    for (int k = 0; k < 2; k++) {
        double x[3], v[3];
        const int n = 3 - k;
        for (int i = 0; i < n; i++) x[i] = H[k + 1 + i][k];
        if (house(x, n, v)) {
            reflect(H, Q, k + 1, n, v);
            for (int i = k + 2; i < 4; i++) H[i][k] = 0.;
        }
    }
    double nrm = 0.;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) nrm = fmax(nrm, fabs(H[i][j]));
    /* Francis double-shift QR */
    int m = 3, its = 0, status = 0;
    while (m >= 0) {
        int l = 0;
        for (int k = m; k >= 1; k--) {
            double s = fabs(H[k - 1][k - 1]) + fabs(H[k][k]);
            if (s == 0.) s = nrm;
            if (fabs(H[k][k - 1]) <= DBL_EPSILON * s) {
                l = k;
                break;
            }
        }
        if (l > 0) H[l][l - 1] = 0.;
        if (l == m) {
            wr[m] = H[m][m];
            wi[m] = 0.;
            m -= 1;
            its = 0;
        } else if (l == m - 1) {
            double cs, sn;
            double a = H[m - 1][m - 1], b = H[m - 1][m], c = H[m][m - 1], d = H[m][m];
            std2(&a, &b, &c, &d, &cs, &sn);
            rot(H, Q, m - 1, cs, sn);
            H[m - 1][m - 1] = a;
            H[m - 1][m] = b;
            H[m][m - 1] = c;
            H[m][m] = d;
            wr[m - 1] = a;
            wr[m] = d;
            if (c == 0.) {
                wi[m - 1] = wi[m] = 0.;
            } else {
                wi[m - 1] = sqrt(fabs(b)) * sqrt(fabs(c));
                wi[m] = -wi[m - 1];
            }
            m -= 2;
            its = 0;
        } else {
            if (++its > 60) {
                status = 1;
                break;
            }
            double s, t;
            if (its == 10 || its == 20 || its == 40) {
                /* exceptional shift (EISPACK hqr) */
                const double w = fabs(H[m][m - 1]) + fabs(H[m - 1][m - 2]);
                s = 1.5 * w;
                t = w * w;
            } else {
                s = H[m - 1][m - 1] + H[m][m];
                t = H[m - 1][m - 1] * H[m][m] - H[m - 1][m] * H[m][m - 1];
            }
            double x[3];
            x[0] = H[l][l] * H[l][l] + H[l][l + 1] * H[l + 1][l] - s * H[l][l] + t;
            x[1] = H[l + 1][l] * (H[l][l] + H[l + 1][l + 1] - s);
            x[2] = H[l + 1][l] * H[l + 2][l + 1];
            for (int k = l; k <= m - 2; k++) {
                double v[3];
                if (house(x, 3, v)) {
                    reflect(H, Q, k, 3, v);
                    if (k > l) {
                        H[k + 1][k - 1] = 0.;
                        H[k + 2][k - 1] = 0.;
                    }
                }
                x[0] = H[k + 1][k];
                x[1] = H[k + 2][k];
                x[2] = k < m - 2 ? H[k + 3][k] : 0.;
            }
            double v[2];
            if (house(x, 2, v)) {
                reflect(H, Q, m - 1, 2, v);
                H[m][m - 2] = 0.;
            }
        }
    }
    if (status) return status;
    /* eigenvectors of the quasi-triangular factor, then back to A's basis */
    const double ulp = DBL_EPSILON, smlnum = DBL_MIN * (4. / ulp);
    for (int ki = 3; ki >= 0; ki--) {
        if (ki < 3 && H[ki + 1][ki] != 0.) continue; /* upper row of a pair */
        const int pair = ki > 0 && H[ki][ki - 1] != 0.;
        cz lam = {wr[ki], pair ? fabs(wi[ki]) : 0.};
        const double smin = fmax(ulp * (fabs(lam.r) + fabs(lam.i)), smlnum);
        cz x[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        int top;
        if (!pair) {
            x[ki].r = 1.;
            for (int k = 0; k < ki; k++) x[k].r = -H[k][ki];
            top = ki - 1;
        } else {
            if (fabs(H[ki - 1][ki]) >= fabs(H[ki][ki - 1])) {
                x[ki - 1].r = 1.;
                x[ki].i = lam.i / H[ki - 1][ki];
            } else {
                x[ki - 1].r = -lam.i / H[ki][ki - 1];
                x[ki].i = 1.;
            }
            for (int k = 0; k < ki - 1; k++) {
                x[k].r = -x[ki - 1].r * H[k][ki - 1];
                x[k].i = -x[ki].i * H[k][ki];
            }
            top = ki - 2;
        }
        for (int j = top; j >= 0;) {
            if (j > 0 && H[j][j - 1] != 0.) {
                /* 2x2 diagonal block rows j-1, j: (T - lam) y = x */
                cz c00 = {H[j - 1][j - 1] - lam.r, -lam.i}, c01 = {H[j - 1][j], 0.};
                cz c10 = {H[j][j - 1], 0.}, c11 = {H[j][j] - lam.r, -lam.i};
                cz b0 = x[j - 1], b1 = x[j], y0, y1;
                if (cz1(c00) >= cz1(c10)) {
                    cz p = czpiv(c00, smin);
                    cz f = czdiv(c10, p);
                    cz u = czpiv(czsub(c11, czmul(f, c01)), smin);
                    y1 = czdiv(czsub(b1, czmul(f, b0)), u);
                    y0 = czdiv(czsub(b0, czmul(c01, y1)), p);
                } else {
                    cz p = czpiv(c10, smin);
                    cz f = czdiv(c00, p);
                    cz u = czpiv(czsub(c01, czmul(f, c11)), smin);
                    y1 = czdiv(czsub(b0, czmul(f, b1)), u);
                    y0 = czdiv(czsub(b1, czmul(c11, y1)), p);
                }
                x[j - 1] = y0;
                x[j] = y1;
                for (int k = 0; k < j - 1; k++) {
                    x[k].r -= H[k][j - 1] * y0.r + H[k][j] * y1.r;
                    x[k].i -= H[k][j - 1] * y0.i + H[k][j] * y1.i;
                }
                j -= 2;
            } else {
                cz d = {H[j][j] - lam.r, -lam.i};
                x[j] = czdiv(x[j], czpiv(d, smin));
                for (int k = 0; k < j; k++) {
                    x[k].r -= H[k][j] * x[j].r;
                    x[k].i -= H[k][j] * x[j].i;
                }
                j -= 1;
            }
        }
        double zr[4], zi[4], n2 = 0.;
        for (int r = 0; r < 4; r++) {
            zr[r] = zi[r] = 0.;
            for (int k = 0; k <= ki; k++) {
                zr[r] += Q[r][k] * x[k].r;
                zi[r] += Q[r][k] * x[k].i;
            }
            n2 += zr[r] * zr[r] + zi[r] * zi[r];
        }
        const double inv = n2 > 0. ? 1. / sqrt(n2) : 0.;
        if (!pair) {
            for (int r = 0; r < 4; r++) {
                vr[4 * r + ki] = zr[r] * inv;
                vi[4 * r + ki] = 0.;
            }
        } else {
            /* column ki-1: eigenvalue wr + i|wi|; column ki its conjugate */
            const int up = wi[ki - 1] > 0. ? ki - 1 : ki, lo = up == ki ? ki - 1 : ki;
            for (int r = 0; r < 4; r++) {
                vr[4 * r + up] = zr[r] * inv;
                vi[4 * r + up] = zi[r] * inv;
                vr[4 * r + lo] = zr[r] * inv;
                vi[4 * r + lo] = -zi[r] * inv;
            }
        }
    }
    return 0;
}

/* ---- Landau frame ---------------------------------------------------------- */

void orc_landau(int N, double L, double dtau, int it, const double *T,
                double *out, double avg[4], int *nfail) {
    const size_t S = (size_t)N * N;
    const double a = L / N;
    double ux = 0., uy = 0., ueta = 0., utau = 0., eps = 0.;
    double aux = 0., auy = 0., aueta = 0., aeps = 0.;
    int fails = 0;
    for (int si = 0; si < N; si++) {
        for (int sj = 0; sj < N; sj++) {
            const size_t pos = (size_t)si * N + sj;
            const double Ttt = T[0 * S + pos], Txx = T[1 * S + pos], Tyy = T[2 * S + pos];
            const double Tee = T[3 * S + pos], Ttx = T[4 * S + pos], Tty = T[5 * S + pos];
            const double Tte = T[6 * S + pos], Txy = T[7 * S + pos], Txe = T[8 * S + pos];
            const double Tye = T[9 * S + pos];
            const double tt = (it * dtau * a) * (it * dtau * a);
            const double data[16] = {Ttt, -Ttx, -Tty, -tt * Tte, Ttx, -Txx, -Txy, -tt * Txe,
                                     Tty, -Txy, -Tyy, -tt * Tye, Tte, -Txe, -Tye, -tt * Tee};
            double wr[4], wi[4], vr[16], vi[16];
            const int rc = orc_eig4(data, wr, wi, vr, vi);
            double c_eps = Ttt, c_ut = 1., c_ux = 0., c_uy = 0., c_ue = 0.;
            if (rc) {
                fails++;
            } else {
                const double tau2 = a * it * dtau * a * it * dtau;
                for (int i = 0; i < 4; i++) {
                    /* square = z0^2 - z1^2 - z2^2 - tau^2 z3^2, complex */
                    double sqr = 0., sqi = 0., eur = 0.;
                    for (int j = 0; j < 4; j++) {
                        const double zr = vr[4 * j + i], zi = vi[4 * j + i];
                        const double pr = zr * zr - zi * zi, pim = 2. * zr * zi;
                        eur += pr;
                        if (j == 0) {
                            sqr += pr;
                            sqi += pim;
                        } else if (j < 3) {
                            sqr -= pr;
                            sqi -= pim;
                        } else {
                            const double ar = tau2 * zr, ai = tau2 * zi;
                            sqr -= ar * zr - ai * zi;
                            sqi -= ar * zi + ai * zr;
                        }
                    }
                    const double factor = sqrt(fabs(eur / sqr));
                    if (sqr > 0.) {
                        eps = wr[i];
                        if (fabs(wi[i]) > 0.001 && si > 0 && sj > 0 && si < N - 5 && sj < N - 5)
                            eps = Ttt;
                    }
                    double s2 = 0.;
                    for (int j = 0; j < 4; j++) {
                        const double zr = vr[4 * j + i] * factor, zi = vi[4 * j + i] * factor;
                        const double pr = zr * zr - zi * zi;
                        if (j == 0)
                            s2 += pr;
                        else if (j < 3)
                            s2 -= pr;
                        else
                            s2 -= (tau2 * zr) * zr - (tau2 * zi) * zi;
                    }
                    if (s2 > 0.) {
                        int flip = 0;
                        for (int j = 0; j < 4; j++) {
                            double zr = vr[4 * j + i] * factor, zi = vi[4 * j + i] * factor;
                            if (j == 0 && zr < 0.) {
                                flip = 1;
                                zr = -zr;
                                zi = -zi;
                            }
                            if (j > 0 && flip) {
                                zr = -zr;
                                zi = -zi;
                            }
                            if (j == 0) utau = eps > 0.1 ? zr : 1.;
                            if (j == 1) ux = eps > 0.1 ? zr : 0.;
                            if (j == 2) uy = eps > 0.1 ? zr : 0.;
                            if (j == 3) ueta = eps > 0.1 ? zr : 0.;
                            if (fabs(zi) > 0.001 && eps > 0.001 && si > 10 && sj > 10 && si < N - 10 &&
                                sj < N - 10) {
                                utau = 1.;
                                ux = 0.;
                                uy = 0.;
                                ueta = 0.;
                                eps = Ttt;
                            }
                        }
                        c_ut = utau;
                        c_ux = ux;
                        c_uy = uy;
                        c_ue = ueta;
                        c_eps = eps;
                        aux += ux * ux * eps;
                        auy += uy * uy * eps;
                        aueta += ueta * ueta * eps * it * dtau * a * it * dtau * a;
                        aeps += eps;
                    }
                }
            }
            out[0 * S + pos] = c_eps;
            out[1 * S + pos] = c_ut;
            out[2 * S + pos] = c_ux;
            out[3 * S + pos] = c_uy;
            out[4 * S + pos] = c_ue;
            double *pi = out + 5 * S;
            if (utau == 1 && ux == 0 && uy == 0 && ueta == 0) {
                for (int k = 0; k < 10; k++) pi[k * S + pos] = 0.;
            } else {
                pi[0 * S + pos] = Ttt - 4. / 3. * eps * utau * utau + eps / 3.;
                pi[1 * S + pos] = Txx - 4. / 3. * eps * ux * ux - eps / 3.;
                pi[2 * S + pos] = Tyy - 4. / 3. * eps * uy * uy - eps / 3.;
                pi[3 * S + pos] = Tee - 4. / 3. * eps * ueta * ueta - eps / 3. / it / dtau / a / it / dtau / a;
                pi[4 * S + pos] = Ttx - 4. / 3. * eps * utau * ux;
                pi[5 * S + pos] = Tty - 4. / 3. * eps * utau * uy;
                pi[6 * S + pos] = Tte - 4. / 3. * eps * utau * ueta;
                pi[7 * S + pos] = Txy - 4. / 3. * eps * ux * uy;
                pi[8 * S + pos] = Txe - 4. / 3. * eps * ux * ueta;
                pi[9 * S + pos] = Tye - 4. / 3. * eps * uy * ueta;
            }
        }
    }
    avg[0] = aux;
    avg[1] = auy;
    avg[2] = aueta;
    avg[3] = aeps;
    if (nfail) *nfail = fails;
}

/* ---- eccentricities ------------------------------------------------------ */

static double ecc_weight(double e, double ut, double cutoff) {
    return e < cutoff ? 0. : e * ut;
}

void orc_eccentricity(int N, double L, double QsmuRatio, double QsmuRatioB,
                      double cutoff, const double *eps, const double *utau,
                      const double *g2A, const double *g2B, double *res) {
    const double a = L / N;
    double avx = 0., avy = 0., toteps = 0., area = 0., avgeden = 0., avgqs = 0.;
    int sum = 0;
    for (int ix = 0; ix < N; ix++) {
        const double x = -L / 2. + a * ix;
        for (int iy = 0; iy < N; iy++) {
            const double y = -L / 2. + a * iy;
            const size_t pos = (size_t)ix * N + iy;
            double w;
            if (eps[pos] < cutoff) {
                w = 0.;
            } else {
                w = eps[pos] * utau[pos];
                area += a * a;
                sum += 1;
                avgeden += eps[pos] * HBARC;
                avgqs += g2A[pos] * QsmuRatio * QsmuRatio * g2B[pos] * QsmuRatioB * QsmuRatioB / a / a / a / a;
            }
            avx += x * w;
            avy += y * w;
            toteps += w;
        }
    }
    avx /= toteps;
    avy /= toteps;
    avgeden /= (double)sum;
    avgqs /= (double)sum;
    double c[7] = {0}, sn[7] = {0}, rn[7] = {0};
    for (int ix = 2; ix < N - 2; ix++) {
        const double x = -L / 2. + a * ix - avx;
        for (int iy = 2; iy < N - 2; iy++) {
            const size_t pos = (size_t)ix * N + iy;
            const double y = -L / 2. + a * iy - avy;
            double phiA;
            if (x >= 0) {
                phiA = atan(y / x);
                if (x == 0) {
                    if (y >= 0)
                        phiA = M_PI / 2.;
                    else
                        phiA = 3. * M_PI / 2.;
                }
            } else {
                phiA = atan(y / x) + M_PI;
            }
            const double w = ecc_weight(eps[pos], utau[pos], cutoff);
            const double r = sqrt(x * x + y * y);
            /* n = 1 is weighted with r^3, n >= 2 with r^n
             * (Evolution.cpp:1614-1633) */
            rn[1] += r * r * r * w;
            rn[2] += r * r * w;
            rn[3] += r * r * r * w;
            rn[4] += r * r * r * r * w;
            rn[5] += r * r * r * r * r * w;
            rn[6] += r * r * r * r * r * r * w;
            c[1] += r * r * r * cos(phiA) * w;
            sn[1] += r * r * r * sin(phiA) * w;
            c[2] += r * r * cos(2. * phiA) * w;
            sn[2] += r * r * sin(2. * phiA) * w;
            c[3] += r * r * r * cos(3. * phiA) * w;
            sn[3] += r * r * r * sin(3. * phiA) * w;
            c[4] += r * r * r * r * cos(4. * phiA) * w;
            sn[4] += r * r * r * r * sin(4. * phiA) * w;
            c[5] += r * r * r * r * r * cos(5. * phiA) * w;
            sn[5] += r * r * r * r * r * sin(5. * phiA) * w;
            c[6] += r * r * r * r * r * r * cos(6. * phiA) * w;
            sn[6] += r * r * r * r * r * r * sin(6. * phiA) * w;
        }
    }
    for (int n = 1; n <= 6; n++) {
        res[n - 1] = sqrt(c[n] * c[n] + sn[n] * sn[n]) / rn[n];
        res[6 + n - 1] = (atan(sn[n] / c[n]) + M_PI) / n;
    }
    double avxSq = 0., avySq = 0.;
    for (int ix = 0; ix < N; ix++) {
        const double x = -L / 2. + a * ix - avx;
        for (int iy = 0; iy < N; iy++) {
            const size_t pos = (size_t)ix * N + iy;
            const double w = ecc_weight(eps[pos], utau[pos], cutoff);
            const double y = -L / 2. + a * iy - avy;
            avxSq += x * x * w;
            avySq += y * y * w;
        }
    }
    avxSq /= toteps;
    avySq /= toteps;
    res[12] = area;
    res[13] = 1. / sqrt(1. / avxSq + 1. / avySq);
    res[14] = avgeden;
    res[15] = avgqs * HBARC;
    res[16] = sqrt(rn[2] / toteps);
    res[17] = avx;
    res[18] = avy;
}
