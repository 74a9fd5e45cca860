// landau.hpp -- Landau frame of T^{mu nu} at one site (MyEigen.cpp:25-321).
//
// The reference hands the mixed tensor T^mu_nu to GSL's real nonsymmetric
// eigensolver and takes the timelike eigenvector as u^mu.  Here the 4x4
// problem is solved in registers: Householder reduction to Hessenberg form,
// Francis double-shift QR to real Schur form, 2x2 blocks in standard form,
// eigenvectors by back substitution in the quasi-triangular factor, unit
// Euclidean norm.  Every array index is a compile-time constant once the
// loops are unrolled (the active window is a run-time predicate), so nothing
// lives in scratch memory.  Only +, -, *, /, sqrt and comparisons are used, and
// the translation unit is built without contraction: the result is the same
// bit pattern on the host and the device.
//
// The functions are host+device so that the CPU build of this same code can
// be compared with the oracle.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#define IPGL_HD __host__ __device__ inline

namespace ipgl {

struct Eig4 {
    double wr[4], wi[4];
    double vr[4][4], vi[4][4];  // eigenvector k in column k
};

IPGL_HD double sgnf(double a, double b) { return b >= 0. ? fabs(a) : -fabs(a); }

// P = I - 2 v v^T / v^T v on rows/columns R0..R0+NN-1 (both sides of H) and
// on the columns of Q
template <int R0, int NN>
IPGL_HD void reflect(double (&H)[4][4], double (&Q)[4][4], const double (&v)[3]) {
    double vtv = 0.;
#pragma unroll
    for (int i = 0; i < NN; i++) vtv += v[i] * v[i];
    if (vtv == 0.) return;
    const double f = 2. / vtv;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        double s = 0.;
#pragma unroll
        for (int i = 0; i < NN; i++) s += v[i] * H[R0 + i][j];
        s *= f;
#pragma unroll
        for (int i = 0; i < NN; i++) H[R0 + i][j] -= s * v[i];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double s = 0., t = 0.;
#pragma unroll
        for (int k = 0; k < NN; k++) {
            s += H[i][R0 + k] * v[k];
            t += Q[i][R0 + k] * v[k];
        }
        s *= f;
        t *= f;
#pragma unroll
        for (int k = 0; k < NN; k++) {
            H[i][R0 + k] -= s * v[k];
            Q[i][R0 + k] -= t * v[k];
        }
    }
}

// Householder vector of x (length NN): 0 if x is already (x0, 0, ...)
template <int NN>
IPGL_HD bool house(const double (&x)[3], double (&v)[3]) {
    double tail = 0.;
#pragma unroll
    for (int i = 1; i < NN; i++) tail += x[i] * x[i];
    if (tail == 0.) return false;
    const double alpha = sqrt(x[0] * x[0] + tail);
    const double beta = -sgnf(alpha, x[0]);
    v[0] = x[0] - beta;
#pragma unroll
    for (int i = 1; i < NN; i++) v[i] = x[i];
    return true;
}

// H <- R^T H R, Q <- Q R on rows/columns (I, I+1), R = [cs -sn; sn cs]
template <int I>
IPGL_HD void rot(double (&H)[4][4], double (&Q)[4][4], double cs, double sn) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const double x = H[I][j], y = H[I + 1][j];
        H[I][j] = cs * x + sn * y;
        H[I + 1][j] = cs * y - sn * x;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        double x = H[r][I], y = H[r][I + 1];
        H[r][I] = cs * x + sn * y;
        H[r][I + 1] = cs * y - sn * x;
        x = Q[r][I];
        y = Q[r][I + 1];
        Q[r][I] = cs * x + sn * y;
        Q[r][I + 1] = cs * y - sn * x;
    }
}

// standard form of the 2x2 block [a b; c d] = R [a' b'; c' d'] R^T: upper
// triangular for a real pair, a' = d' and b'c' < 0 for a complex pair
IPGL_HD void std2(double& a, double& b, double& c, double& d, double& cs, double& sn) {
    const double eps = DBL_EPSILON;
    if (c == 0.) {
        cs = 1.;
        sn = 0.;
    } else if (b == 0.) {
        cs = 0.;
        sn = 1.;
        const double t = d;
        d = a;
        a = t;
        b = -c;
        c = 0.;
    } else if (a - d == 0. && sgnf(1., b) != sgnf(1., c)) {
        cs = 1.;
        sn = 0.;
    } else {
        double temp = a - d, p = 0.5 * temp;
        const double bcmax = fmax(fabs(b), fabs(c));
        const double bcmis = fmin(fabs(b), fabs(c)) * sgnf(1., b) * sgnf(1., c);
        const double scale = fmax(fabs(p), bcmax);
        double z = p / scale * p + bcmax / scale * bcmis;
        if (z >= 4. * eps) {
            z = p + sgnf(sqrt(scale) * sqrt(z), p);
            a = d + z;
            d = d - bcmax / z * bcmis;
            const double tau = sqrt(c * c + z * z);
            cs = z / tau;
            sn = c / tau;
            b = b - c;
            c = 0.;
        } else {
            const double sigma = b + c;
            const double tau = sqrt(sigma * sigma + temp * temp);
            cs = sqrt(0.5 * (1. + fabs(sigma) / tau));
            sn = -(p / (tau * cs)) * sgnf(1., sigma);
            const double aa = a * cs + b * sn, bb = -a * sn + b * cs;
            const double cc = c * cs + d * sn, dd = -c * sn + d * cs;
            a = aa * cs + cc * sn;
            b = bb * cs + dd * sn;
            c = -aa * sn + cc * cs;
            d = -bb * sn + dd * cs;
            temp = 0.5 * (a + d);
            a = temp;
            d = temp;
            if (c != 0.) {
                if (b != 0.) {
                    if (sgnf(1., b) == sgnf(1., c)) {
                        const double sab = sqrt(fabs(b)), sac = sqrt(fabs(c));
                        p = sgnf(sab * sac, c);
                        const double t2 = 1. / sqrt(fabs(b + c));
                        a = temp + p;
                        d = temp - p;
                        b = b - c;
                        c = 0.;
                        const double cs1 = sab * t2, sn1 = sac * t2;
                        temp = cs * cs1 - sn * sn1;
                        sn = cs * sn1 + sn * cs1;
                        cs = temp;
                    }
                } else {
                    b = -c;
                    c = 0.;
                    temp = cs;
                    cs = -sn;
                    sn = temp;
                }
            }
        }
    }
}

// standardise the block at rows (M-1, M) and record its eigenvalues
template <int M>
IPGL_HD void block2(double (&H)[4][4], double (&Q)[4][4], Eig4& r) {
    double cs, sn;
    double a = H[M - 1][M - 1], b = H[M - 1][M], c = H[M][M - 1], d = H[M][M];
    std2(a, b, c, d, cs, sn);
    rot<M - 1>(H, Q, cs, sn);
    H[M - 1][M - 1] = a;
    H[M - 1][M] = b;
    H[M][M - 1] = c;
    H[M][M] = d;
    r.wr[M - 1] = a;
    r.wr[M] = d;
    if (c == 0.) {
        r.wi[M - 1] = r.wi[M] = 0.;
    } else {
        r.wi[M - 1] = sqrt(fabs(b)) * sqrt(fabs(c));
        r.wi[M] = -r.wi[M - 1];
    }
}

// complex helpers of the back substitution
struct cz {
    double r, i;
};
IPGL_HD cz czdiv(cz a, cz b) {
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
IPGL_HD cz czmul(cz a, cz b) { return cz{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
IPGL_HD cz czsub(cz a, cz b) { return cz{a.r - b.r, a.i - b.i}; }
IPGL_HD double cz1(cz a) { return fabs(a.r) + fabs(a.i); }
IPGL_HD cz czpiv(cz p, double smin) { return cz1(p) < smin ? cz{smin, 0.} : p; }

// eigenvector of the eigenvalue at KI (the lower row of a complex pair)
template <int KI>
IPGL_HD void eigvec(const double (&H)[4][4], const double (&Q)[4][4], Eig4& r) {
    constexpr int KU = KI < 3 ? KI + 1 : 3;
    if (KI < 3 && H[KU][KI] != 0.) return;  // upper row: written with its pair
    constexpr double ulp = DBL_EPSILON;
    const double smlnum = DBL_MIN * (4. / ulp);
    const bool pair = KI > 0 && H[KI][KI > 0 ? KI - 1 : 0] != 0.;
    const cz lam = {r.wr[KI], pair ? fabs(r.wi[KI]) : 0.};
    const double smin = fmax(ulp * (fabs(lam.r) + fabs(lam.i)), smlnum);
    cz x[4] = {{0., 0.}, {0., 0.}, {0., 0.}, {0., 0.}};
    constexpr int K1 = KI > 0 ? KI - 1 : 0;
    if (!pair) {
        x[KI].r = 1.;
#pragma unroll
        for (int k = 0; k < KI; k++) x[k].r = -H[k][KI];
    } else {
        if (fabs(H[K1][KI]) >= fabs(H[KI][K1])) {
            x[K1].r = 1.;
            x[KI].i = lam.i / H[K1][KI];
        } else {
            x[K1].r = -lam.i / H[KI][K1];
            x[KI].i = 1.;
        }
#pragma unroll
        for (int k = 0; k < K1; k++) {
            x[k].r = -x[K1].r * H[k][K1];
            x[k].i = -x[KI].i * H[k][KI];
        }
    }
    // upward solve of (T - lam) y = x over the rows above the start row
    bool skip = pair;  // a pair's upper row is already set
#pragma unroll
    for (int j = KI - 1; j >= 0; j--) {
        if (skip) {
            skip = false;
            continue;
        }
        if (j > 0 && H[j][j > 0 ? j - 1 : 0] != 0.) {
            const int j1 = j > 0 ? j - 1 : 0;
            const cz c00 = {H[j1][j1] - lam.r, -lam.i}, c01 = {H[j1][j], 0.};
            const cz c10 = {H[j][j1], 0.}, c11 = {H[j][j] - lam.r, -lam.i};
            const cz b0 = x[j1], b1 = x[j];
            cz y0, y1;
            if (cz1(c00) >= cz1(c10)) {
                const cz p = czpiv(c00, smin);
                const cz f = czdiv(c10, p);
                const cz u = czpiv(czsub(c11, czmul(f, c01)), smin);
                y1 = czdiv(czsub(b1, czmul(f, b0)), u);
                y0 = czdiv(czsub(b0, czmul(c01, y1)), p);
            } else {
                const cz p = czpiv(c10, smin);
                const cz f = czdiv(c00, p);
                const cz u = czpiv(czsub(c01, czmul(f, c11)), smin);
                y1 = czdiv(czsub(b0, czmul(f, b1)), u);
                y0 = czdiv(czsub(b1, czmul(c11, y1)), p);
            }
            x[j1] = y0;
            x[j] = y1;
#pragma unroll
            for (int k = 0; k < j1; k++) {
                x[k].r -= H[k][j1] * y0.r + H[k][j] * y1.r;
                x[k].i -= H[k][j1] * y0.i + H[k][j] * y1.i;
            }
            skip = true;
        } else {
            const cz d = {H[j][j] - lam.r, -lam.i};
            x[j] = czdiv(x[j], czpiv(d, smin));
#pragma unroll
            for (int k = 0; k < j; k++) {
                x[k].r -= H[k][j] * x[j].r;
                x[k].i -= H[k][j] * x[j].i;
            }
        }
    }
    double zr[4], zi[4], n2 = 0.;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        zr[q] = zi[q] = 0.;
#pragma unroll
        for (int k = 0; k <= KI; k++) {
            zr[q] += Q[q][k] * x[k].r;
            zi[q] += Q[q][k] * x[k].i;
        }
        n2 += zr[q] * zr[q] + zi[q] * zi[q];
    }
    const double inv = n2 > 0. ? 1. / sqrt(n2) : 0.;
    if (!pair) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            r.vr[q][KI] = zr[q] * inv;
            r.vi[q][KI] = 0.;
        }
    } else {
        // wr + i|wi| goes to the column whose wi is positive (the upper row;
        // the lower one only if sqrt|b| sqrt|c| underflowed)
        const bool upper = r.wi[K1] > 0.;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double a = zr[q] * inv, b = zi[q] * inv;
            r.vr[q][K1] = a;
            r.vi[q][K1] = upper ? b : -b;
            r.vr[q][KI] = a;
            r.vi[q][KI] = upper ? -b : b;
        }
    }
}

// returns false when the QR iteration does not converge
IPGL_HD bool eig4(const double (&A)[4][4], Eig4& r) {
    double H[4][4], Q[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            H[i][j] = A[i][j];
            Q[i][j] = i == j ? 1. : 0.;
        }
    {
        double x[3] = {H[1][0], H[2][0], H[3][0]}, v[3] = {0., 0., 0.};
        if (house<3>(x, v)) {
            reflect<1, 3>(H, Q, v);
            H[2][0] = 0.;
            H[3][0] = 0.;
        }
        double x2[3] = {H[2][1], H[3][1], 0.}, v2[3] = {0., 0., 0.};
        if (house<2>(x2, v2)) {
            reflect<2, 2>(H, Q, v2);
            H[3][1] = 0.;
        }
    }
    double nrm = 0.;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) nrm = fmax(nrm, fabs(H[i][j]));
    int m = 3, its = 0;
    while (m >= 0) {
        int l = 0;
#pragma unroll
        for (int k = 1; k <= 3; k++) {
            double s = fabs(H[k - 1][k - 1]) + fabs(H[k][k]);
            if (s == 0.) s = nrm;
            if (k <= m && fabs(H[k][k - 1]) <= DBL_EPSILON * s) l = k;
        }
#pragma unroll
        for (int k = 1; k <= 3; k++)
            if (k == l) H[k][k - 1] = 0.;
        if (l == m) {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (k == m) {
                    r.wr[k] = H[k][k];
                    r.wi[k] = 0.;
                }
            m -= 1;
            its = 0;
        } else if (l == m - 1) {
            if (m == 1) block2<1>(H, Q, r);
            if (m == 2) block2<2>(H, Q, r);
            if (m == 3) block2<3>(H, Q, r);
            m -= 2;
            its = 0;
        } else {
            if (++its > 60) return false;
            // the window [l, m] has l <= 1 and m >= l + 2
            double hm1m1 = 0., hmm = 0., hm1m = 0., hmm1 = 0., hm1m2 = 0.;
#pragma unroll
            for (int k = 2; k <= 3; k++)
                if (k == m) {
                    hm1m1 = H[k - 1][k - 1];
                    hmm = H[k][k];
                    hm1m = H[k - 1][k];
                    hmm1 = H[k][k - 1];
                    hm1m2 = H[k - 1][k - 2];
                }
            double s, t;
            if (its == 10 || its == 20 || its == 40) {
                const double w = fabs(hmm1) + fabs(hm1m2);
                s = 1.5 * w;
                t = w * w;
            } else {
                s = hm1m1 + hmm;
                t = hm1m1 * hmm - hm1m * hmm1;
            }
            double x[3];
            if (l == 0) {
                x[0] = H[0][0] * H[0][0] + H[0][1] * H[1][0] - s * H[0][0] + t;
                x[1] = H[1][0] * (H[0][0] + H[1][1] - s);
                x[2] = H[1][0] * H[2][1];
            } else {
                x[0] = H[1][1] * H[1][1] + H[1][2] * H[2][1] - s * H[1][1] + t;
                x[1] = H[2][1] * (H[1][1] + H[2][2] - s);
                x[2] = H[2][1] * H[3][2];
            }
            // k = 0 (only when l = 0)
            if (l == 0) {
                double v[3] = {0., 0., 0.};
                if (house<3>(x, v)) reflect<0, 3>(H, Q, v);
                x[0] = H[1][0];
                x[1] = H[2][0];
                x[2] = m > 2 ? H[3][0] : 0.;
            }
            // k = 1 (when m = 3)
            if (m == 3) {
                double v[3] = {0., 0., 0.};
                if (house<3>(x, v)) {
                    reflect<1, 3>(H, Q, v);
                    if (l == 0) {
                        H[2][0] = 0.;
                        H[3][0] = 0.;
                    }
                }
                x[0] = H[2][1];
                x[1] = H[3][1];
                x[2] = 0.;
            }
            {
                double v[3] = {0., 0., 0.};
                if (house<2>(x, v)) {
                    if (m == 2) {
                        reflect<1, 2>(H, Q, v);
                        H[2][0] = 0.;
                    } else {
                        reflect<2, 2>(H, Q, v);
                        H[3][1] = 0.;
                    }
                }
            }
        }
    }
    eigvec<3>(H, Q, r);
    eigvec<2>(H, Q, r);
    eigvec<1>(H, Q, r);
    eigvec<0>(H, Q, r);
    return true;
}

// ---- Landau matching (MyEigen.cpp:49-264) -------------------------------------

// T order: tautau, xx, yy, etaeta, taux, tauy, taueta, xy, xeta, yeta
struct LandauSite {
    double eps, u[4];   // cell values (defaults T^tautau, (1,0,0,0))
    double eps_var;     // the reference's running eps after this site
    int flags;          // bit0: eps_var set here, bit1: u^mu set here, bit2: QR failed
    double acc[4];      // sum u_x^2 eps, u_y^2 eps, (tau u^eta)^2 eps, eps
};

IPGL_HD void landau_site(const double (&T)[10], int N, double a, double dtau, int it, int si, int sj,
                        LandauSite& o) {
    const double Ttt = T[0], Txx = T[1], Tyy = T[2], Tee = T[3], Ttx = T[4], Tty = T[5];
    const double Tte = T[6], Txy = T[7], Txe = T[8], Tye = T[9];
    const double tt = (it * dtau * a) * (it * dtau * a);
    const double A[4][4] = {{Ttt, -Ttx, -Tty, -tt * Tte},
                            {Ttx, -Txx, -Txy, -tt * Txe},
                            {Tty, -Txy, -Tyy, -tt * Tye},
                            {Tte, -Txe, -Tye, -tt * Tee}};
    o.eps = Ttt;
    o.u[0] = 1.;
    o.u[1] = o.u[2] = o.u[3] = 0.;
    o.eps_var = 0.;
    o.flags = 0;
    o.acc[0] = o.acc[1] = o.acc[2] = o.acc[3] = 0.;
    Eig4 r;
    if (!eig4(A, r)) {
        o.flags = 4;
        return;
    }
    const double tau2 = a * it * dtau * a * it * dtau;
    double eps = 0., ut = 0., ux = 0., uy = 0., ue = 0.;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double sqr = 0., eur = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double zr = r.vr[j][i], zi = r.vi[j][i];
            const double pr = zr * zr - zi * zi;
            eur += pr;
            if (j == 0)
                sqr += pr;
            else if (j < 3)
                sqr -= pr;
            else
                sqr -= (tau2 * zr) * zr - (tau2 * zi) * zi;
        }
        const double factor = sqrt(fabs(eur / sqr));
        if (sqr > 0.) {
            eps = r.wr[i];
            if (fabs(r.wi[i]) > 0.001 && si > 0 && sj > 0 && si < N - 5 && sj < N - 5) eps = Ttt;
            o.flags |= 1;
        }
        double s2 = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double zr = r.vr[j][i] * factor, zi = r.vi[j][i] * factor;
            const double pr = zr * zr - zi * zi;
            if (j == 0)
                s2 += pr;
            else if (j < 3)
                s2 -= pr;
            else
                s2 -= (tau2 * zr) * zr - (tau2 * zi) * zi;
        }
        if (s2 > 0.) {
            bool flip = false;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                double zr = r.vr[j][i] * factor, zi = r.vi[j][i] * factor;
                if (j == 0 && zr < 0.) {
                    flip = true;
                    zr = -zr;
                    zi = -zi;
                }
                if (j > 0 && flip) {
                    zr = -zr;
                    zi = -zi;
                }
                if (j == 0) ut = eps > 0.1 ? zr : 1.;
                if (j == 1) ux = eps > 0.1 ? zr : 0.;
                if (j == 2) uy = eps > 0.1 ? zr : 0.;
                if (j == 3) ue = eps > 0.1 ? zr : 0.;
                if (fabs(zi) > 0.001 && eps > 0.001 && si > 10 && sj > 10 && si < N - 10 && sj < N - 10) {
                    ut = 1.;
                    ux = 0.;
                    uy = 0.;
                    ue = 0.;
                    eps = Ttt;
                }
            }
            o.u[0] = ut;
            o.u[1] = ux;
            o.u[2] = uy;
            o.u[3] = ue;
            o.eps = eps;
            o.flags |= 2;
            o.acc[0] += ux * ux * eps;
            o.acc[1] += uy * uy * eps;
            o.acc[2] += ue * ue * eps * it * dtau * a * it * dtau * a;
            o.acc[3] += eps;
        }
    }
    o.eps_var = eps;
}

// ---- reverse mode of the cell values (eps, u^tau) --------------------------------
// x = M^T-solve of the 5x5 bordered system (partial pivoting); b overwritten
IPGL_HD void solve5(double (&M)[5][5], double (&b)[5]) {
    for (int c = 0; c < 5; c++) {
        int pr = c;
        for (int r = c + 1; r < 5; r++)
            if (fabs(M[r][c]) > fabs(M[pr][c])) pr = r;
        if (pr != c) {
            for (int k = 0; k < 5; k++) {
                const double t = M[c][k];
                M[c][k] = M[pr][k];
                M[pr][k] = t;
            }
            const double t = b[c];
            b[c] = b[pr];
            b[pr] = t;
        }
        const double d = M[c][c];
        if (d == 0.) continue;
        for (int r = c + 1; r < 5; r++) {
            const double f = M[r][c] / d;
            for (int k = c; k < 5; k++) M[r][k] -= f * M[c][k];
            b[r] -= f * b[c];
        }
    }
    for (int c = 4; c >= 0; c--) {
        double v = b[c];
        for (int k = c + 1; k < 5; k++) v -= M[c][k] * b[k];
        b[c] = M[c][c] != 0. ? v / M[c][c] : 0.;
    }
}

// Given eps_b = dL/d eps and ut_b = dL/d u^tau of the cell values of one site
// (LandauSite::eps, u[0]), adds dL/dT[k] to Tb (T's order).  The branches of
// landau_site are replayed: a value set to a constant (the rest frame, u^tau
// = 1) carries no derivative, eps = T^tautau (the fallbacks and sites without
// a timelike vector) carries dL/dT^tautau, and a real eigenpair (lambda, u)
// of A = T g, g = diag(1, -1, -1, -tau^2), u^T g u = 1, gives
//   d lambda = (g u)^T dA u
//   du       from [[A - lambda, u], [(g u)^T, 0]] [du; s] = [d lambda u - dA u; 0],
// so that with M^T [l; m] = [ut_b e_tau; 0] and rho = (eps_b + l.u) g u - l,
// dL = rho^T dA u.  Where the reference keeps eps = Re(lambda) of a complex
// eigenvalue (near the lattice edges) its derivative is taken; u^tau from the
// real part of a complex eigenvector is differentiated by central differences.
IPGL_HD void landau_site_adj(const double (&T)[10], int N, double a, double dtau, int it, int si, int sj,
                            double eps_b, double ut_b, double (&Tb)[10]) {
    const double Ttt = T[0], Txx = T[1], Tyy = T[2], Tee = T[3], Ttx = T[4], Tty = T[5];
    const double Tte = T[6], Txy = T[7], Txe = T[8], Tye = T[9];
    const double tt = (it * dtau * a) * (it * dtau * a);
    const double A[4][4] = {{Ttt, -Ttx, -Tty, -tt * Tte},
                            {Ttx, -Txx, -Txy, -tt * Txe},
                            {Tty, -Txy, -Tyy, -tt * Tye},
                            {Tte, -Txe, -Tye, -tt * Tee}};
    Eig4 r;
    if (!eig4(A, r)) {
        Tb[0] += eps_b;
        return;
    }
    const double tau2 = a * it * dtau * a * it * dtau;
    // replay: sources of the running eps and of the cell values
    // (-1: constant, 0..3: eigenpair, 4: T^tautau)
    double eps = 0.;
    int e_src = -1, cell_e = 4, cell_u = -1;
    double cell_flip = 1.;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double sqr = 0., eur = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double zr = r.vr[j][i], zi = r.vi[j][i];
            const double pr = zr * zr - zi * zi;
            eur += pr;
            if (j == 0)
                sqr += pr;
            else if (j < 3)
                sqr -= pr;
            else
                sqr -= (tau2 * zr) * zr - (tau2 * zi) * zi;
        }
        const double factor = sqrt(fabs(eur / sqr));
        if (sqr > 0.) {
            eps = r.wr[i];
            e_src = i;
            if (fabs(r.wi[i]) > 0.001 && si > 0 && sj > 0 && si < N - 5 && sj < N - 5) {
                eps = Ttt;
                e_src = 4;
            }
        }
        double s2 = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double zr = r.vr[j][i] * factor, zi = r.vi[j][i] * factor;
            const double pr = zr * zr - zi * zi;
            if (j == 0)
                s2 += pr;
            else if (j < 3)
                s2 -= pr;
            else
                s2 -= (tau2 * zr) * zr - (tau2 * zi) * zi;
        }
        if (s2 > 0.) {
            const double flip = r.vr[0][i] * factor < 0. ? -1. : 1.;
            int u_src = eps > 0.1 ? i : -1;
            bool fb = false;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const double zi = r.vi[j][i] * factor;
                if (fabs(zi) > 0.001 && eps > 0.001 && si > 10 && sj > 10 && si < N - 10 && sj < N - 10) fb = true;
            }
            if (fb) {
                u_src = -1;
                eps = Ttt;
                e_src = 4;
            }
            cell_e = e_src;
            cell_u = u_src;
            cell_flip = flip;
        }
    }
    // eps from T^tautau
    if (cell_e == 4) Tb[0] += eps_b;
    const double g[4] = {1., -1., -1., -tt};
    // the normalised real eigenvector of pair k (u^tau > 0 when flipped as the
    // cell's), and its lambda
    auto vec = [&](int k, double fl, double (&u)[4]) {
        double sqr = 0.;
        for (int j = 0; j < 4; j++) sqr += g[j] * r.vr[j][k] * r.vr[j][k];
        const double f = fl / sqrt(fabs(sqr));
        for (int j = 0; j < 4; j++) u[j] = r.vr[j][k] * f;
    };
    auto add = [&](const double (&rho)[4], const double (&u)[4]) {
        // Tb[k] += rho^T (dA / dT[k]) u, A[m][n] = T^{mn} g_n
        const int ia[10] = {0, 1, 2, 3, 0, 0, 0, 1, 1, 2};
        const int ib[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
        for (int k = 0; k < 10; k++) {
            const int al = ia[k], be = ib[k];
            double v = rho[al] * g[be] * u[be];
            if (al != be) v += rho[be] * g[al] * u[al];
            Tb[k] += v;
        }
    };
    // eps = Re(lambda) of a complex eigenvalue (kept at the lattice edges):
    // d lambda = (g v)^T dA v / (v^T g v) with the complex eigenvector v
    if (cell_e >= 0 && cell_e < 4 && r.wi[cell_e] != 0.) {
        double dr = 0., di = 0.;
        for (int j = 0; j < 4; j++) {
            const double vr = r.vr[j][cell_e], vi = r.vi[j][cell_e];
            dr += g[j] * (vr * vr - vi * vi);
            di += g[j] * 2. * vr * vi;
        }
        const double dd = dr * dr + di * di;
        if (dd > 0.) {
            // rho = eps_b g v / (v^T g v)
            double pr[4], pi_[4];
            for (int j = 0; j < 4; j++) {
                const double gr = g[j] * r.vr[j][cell_e], gi = g[j] * r.vi[j][cell_e];
                pr[j] = eps_b * (gr * dr + gi * di) / dd;
                pi_[j] = eps_b * (gi * dr - gr * di) / dd;
            }
            const int ia[10] = {0, 1, 2, 3, 0, 0, 0, 1, 1, 2};
            const int ib[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
            for (int k = 0; k < 10; k++) {
                const int al = ia[k], be = ib[k];
                const double vbr = g[be] * r.vr[be][cell_e], vbi = g[be] * r.vi[be][cell_e];
                double v = pr[al] * vbr - pi_[al] * vbi;
                if (al != be) {
                    const double var = g[al] * r.vr[al][cell_e], vai = g[al] * r.vi[al][cell_e];
                    v += pr[be] * var - pi_[be] * vai;
                }
                Tb[k] += v;
            }
        }
    }
    // u^mu from the real part of a complex eigenvector (the reference's
    // u outside the central region): its phase follows the solver's
    // back-substitution convention, so u^tau is differentiated by central
    // differences of landau_site in each of the ten components
    if (cell_u >= 0 && r.wi[cell_u] != 0. && ut_b != 0.) {
        // fourth-order stencil (8 (f(h) - f(-h)) - (f(2h) - f(-2h))) / 12 h
        for (int k = 0; k < 10; k++) {
            const double hk = 1e-5 * (fabs(T[k]) + 1e-4 * fabs(T[0])) + 1e-300;
            double f[4];
            const double off[4] = {hk, -hk, 2. * hk, -2. * hk};
            for (int m = 0; m < 4; m++) {
                double tq[10];
                for (int q = 0; q < 10; q++) tq[q] = T[q];
                tq[k] += off[m];
                LandauSite o;
                landau_site(tq, N, a, dtau, it, si, sj, o);
                f[m] = o.u[0];
            }
            Tb[k] += ut_b * (8. * (f[0] - f[1]) - (f[2] - f[3])) / (12. * hk);
        }
    }
    const bool e_pair = cell_e >= 0 && cell_e < 4 && r.wi[cell_e] == 0.;
    const bool u_pair = cell_u >= 0 && r.wi[cell_u] == 0. && ut_b != 0.;
    if (u_pair) {
        double u[4];
        vec(cell_u, cell_flip, u);
        const double lam = r.wr[cell_u];
        double gu[4];
        for (int j = 0; j < 4; j++) gu[j] = g[j] * u[j];
        // M^T = [[A^T - lam, g u], [u^T, 0]]
        double M[5][5], b[5] = {ut_b, 0., 0., 0., 0.};
        for (int i = 0; i < 4; i++) {
            for (int j = 0; j < 4; j++) M[i][j] = A[j][i] - (i == j ? lam : 0.);
            M[i][4] = gu[i];
            M[4][i] = u[i];
        }
        M[4][4] = 0.;
        solve5(M, b);
        double lu = 0.;
        for (int j = 0; j < 4; j++) lu += b[j] * u[j];
        const double eb = (e_pair && cell_e == cell_u) ? eps_b : 0.;
        double rho[4];
        for (int j = 0; j < 4; j++) rho[j] = (eb + lu) * gu[j] - b[j];
        add(rho, u);
        if (e_pair && cell_e != cell_u) {
            double v[4], gv[4];
            vec(cell_e, 1., v);
            for (int j = 0; j < 4; j++) gv[j] = g[j] * v[j] * eps_b;
            add(gv, v);
        }
    } else if (e_pair) {
        double v[4], gv[4];
        vec(cell_e, 1., v);
        for (int j = 0; j < 4; j++) gv[j] = g[j] * v[j] * eps_b;
        add(gv, v);
    }
}

// ---- eps_2 in reverse (Evolution.cpp:1280-1660) ------------------------------
// eps_2 = |C + i Sn| / R over the sites 2 <= ix, iy < N-2, relative to the
// centroid (xb, yb) = sum_all w (x, y) / W of the weights w = eps u^tau:
//   C = sum w (dx^2 - dy^2), Sn = sum 2 w dx dy, R = sum w (dx^2 + dy^2)
// (the reference's r^2 cos 2phi, r^2 sin 2phi, r^2).  Mx, My = sum w dx, w dy
// over the same sites.  Returns dL/dw at a site (dx, dy from the centroid)
// for dL/d eps_2 = e2b, through the sums and the centroid.
IPGL_HD double ecc2_wbar(double C, double Sn, double R, double Mx, double My, double W, double dx, double dy,
                        bool interior, double e2b) {
    const double m = sqrt(C * C + Sn * Sn);
    if (m == 0. || R == 0. || W == 0.) return 0.;
    const double e2 = m / R;
    const double aC = e2b * C / (m * R), aS = e2b * Sn / (m * R), aR = -e2b * e2 / R;
    const double xbb = -2. * Mx * (aC + aR) - 2. * My * aS;
    const double ybb = 2. * My * aC - 2. * Mx * aS - 2. * My * aR;
    double v = (xbb * dx + ybb * dy) / W;
    if (interior) v += aC * (dx * dx - dy * dy) + aS * 2. * dx * dy + aR * (dx * dx + dy * dy);
    return v;
}

// pi^{mu nu} from T and the running (eps, u^mu) (MyEigen.cpp:267-309)
IPGL_HD void landau_pi(const double (&T)[10], double eps, double ut, double ux, double uy, double ue, int it,
                      double dtau, double a, double (&pi)[10]) {
    if (ut == 1 && ux == 0 && uy == 0 && ue == 0) {
#pragma unroll
        for (int k = 0; k < 10; k++) pi[k] = 0.;
        return;
    }
    pi[0] = T[0] - 4. / 3. * eps * ut * ut + eps / 3.;
    pi[1] = T[1] - 4. / 3. * eps * ux * ux - eps / 3.;
    pi[2] = T[2] - 4. / 3. * eps * uy * uy - eps / 3.;
    pi[3] = T[3] - 4. / 3. * eps * ue * ue - eps / 3. / it / dtau / a / it / dtau / a;
    pi[4] = T[4] - 4. / 3. * eps * ut * ux;
    pi[5] = T[5] - 4. / 3. * eps * ut * uy;
    pi[6] = T[6] - 4. / 3. * eps * ut * ue;
    pi[7] = T[7] - 4. / 3. * eps * ux * uy;
    pi[8] = T[8] - 4. / 3. * eps * ux * ue;
    pi[9] = T[9] - 4. / 3. * eps * uy * ue;
}

}  // namespace ipgl
