// gaugefix.hpp -- per-site pieces of the Fourier-accelerated Coulomb gauge
// fixing (GaugeFix::FFTChi, GaugeFix.cpp:17-165) on host and device.
//
//   expm_pade6   Matrix::expm (Matrix.cpp:431-560): diagonal Pade
//                approximant of order 6 with scaling and squaring, the
//                reference's rounded coefficients, the irreducible fraction
//                evaluated by Horner in H^2 and inverted by the adjugate
//   reunitarize  Matrix::reu (Matrix.h:67-90): rows 1 and 2 by Gram-Schmidt,
//                row 3 the conjugated cross product, each normalised
#pragma once
#include "su3.hpp"

namespace ipggf {

IPG_HD double cabs2(cd a) { return a.r * a.r + a.i * a.i; }
IPG_HD double cabs(cd a) { return sqrt(cabs2(a)); }

IPG_HD M3 add_diag(const M3& a, double s) {
    M3 c = a;
    c.e[0].r += s;
    c.e[4].r += s;
    c.e[8].r += s;
    return c;
}

// exp(A) for a 3x3 complex A; *ok = false when A holds NaN or infinity
IPG_HD M3 expm_pade6(const M3& A, bool* ok) {
    // diagonal Pade (6, 6) coefficients, rounded as in the reference
    const double c[7] = {1., 0.5, 0.1136363636, 0.01515151515, 0.001262626263, 6.313131313e-05,
                         1.503126503e-06};
    *ok = true;
    double norm = 0.;
    for (int i = 0; i < 3; i++) {
        double row = 0.;
        for (int j = 0; j < 3; j++) row += cabs(A.e[3 * i + j]);
        norm = fmax(norm, row);
    }
    if (norm == 0.0) {
        bool zero = true;
        for (int k = 0; k < 9; k++) zero = zero && A.e[k].r == 0.0 && A.e[k].i == 0.0;
        if (!zero) *ok = false;
        return m_one();
    }
    if (!(norm == norm) || norm > 1e300) {
        *ok = false;
        return m_one();
    }
    // scale by 2^-s so that the norm is below 1/2
    int s = 0;
    M3 U = A;
    if (norm > 0.5) {
        s = (int)(log(norm) / log(2.0) + 2.0);
        if (s < 0) s = 0;
        const double scale = 1.0 / pow(2.0, (double)s);
        U = scale * A;
    }
    const M3 H2 = U * U;
    // Horner in H2: Q from c6, P from c5, alternating down to c0
    M3 Q = add_diag(m_zero(), c[6]);
    M3 P = add_diag(m_zero(), c[5]);
    int odd = 1;
    for (int k = 5; k > 0; --k) {
        if (odd == 1)
            Q = add_diag(Q * H2, c[k - 1]);
        else
            P = add_diag(P * H2, c[k - 1]);
        odd = 1 - odd;
    }
    if (odd == 1)
        Q = Q * U;
    else
        P = P * U;
    Q = Q - P;
    // Q^-1 = adj(Q) / det(Q)
    const cd* q = Q.e;
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
    const cd det = q[0] * q[4] * q[8] + q[1] * q[5] * q[6] + q[2] * q[3] * q[7] - q[2] * q[4] * q[6] -
                   q[1] * q[3] * q[8] - q[5] * q[7] * q[0];
    const cd inv = cdiv(cd{1., 0.}, det);
    for (int k = 0; k < 9; k++) Qi.e[k] = Qi.e[k] * inv;
    M3 R = (2.0 * Qi) * P;
    if (odd == 1)
        R = -1. * add_diag(R, 1.);
    else
        R = add_diag(R, 1.);
    for (int i = 0; i < s; ++i) R = R * R;
    return R;
}

// rows of the matrix as vectors: normalise row 0, Gram-Schmidt row 1
// against it, row 2 = conj(row0 x row1) normalised
IPG_HD void normalize3(cd (&v)[3]) {
    const double n = sqrt(cabs2(v[0]) + cabs2(v[1]) + cabs2(v[2]));
    for (int k = 0; k < 3; k++) v[k] = cd{v[k].r / n, v[k].i / n};
}
IPG_HD M3 reunitarize(const M3& m) {
    cd e1[3] = {m.e[0], m.e[1], m.e[2]};
    cd a2[3] = {m.e[3], m.e[4], m.e[5]};
    normalize3(e1);
    // Spinor::GramSchmidt normalises its argument again
    cd u[3] = {e1[0], e1[1], e1[2]};
    normalize3(u);
    const cd p = cmul(u[0], a2[0]) + cmul(u[1], a2[1]) + cmul(u[2], a2[2]);  // <u, a2>
    cd e2[3];
    for (int k = 0; k < 3; k++) e2[k] = a2[k] - p * u[k];
    normalize3(e2);
    cd e3[3] = {cconj(e1[1] * e2[2] - e1[2] * e2[1]), cconj(e1[2] * e2[0] - e1[0] * e2[2]),
                cconj(e1[0] * e2[1] - e1[1] * e2[0])};
    normalize3(e3);
    M3 r;
    for (int k = 0; k < 3; k++) {
        r.e[k] = e1[k];
        r.e[3 + k] = e2[k];
        r.e[6 + k] = e3[k];
    }
    return r;
}

// the gauge-fixing rotation of one site: g = reu(exp(i chi^a t^a)), or 1
// when the exponential is not finite (GaugeFix.cpp:128-147)
IPG_HD M3 gauge_rotation(const double (&chi)[8]) {
    M3 H = m_zero();
    const double s3 = 1.0 / (2.0 * sqrt(3.0));
    // sum_a chi^a t^a (Group.cpp:68-149)
    H.e[1] = cd{0.5 * chi[0], -0.5 * chi[1]};
    H.e[3] = cd{0.5 * chi[0], 0.5 * chi[1]};
    H.e[0] = cd{0.5 * chi[2] + s3 * chi[7], 0.};
    H.e[4] = cd{-0.5 * chi[2] + s3 * chi[7], 0.};
    H.e[2] = cd{0.5 * chi[3], -0.5 * chi[4]};
    H.e[6] = cd{0.5 * chi[3], 0.5 * chi[4]};
    H.e[5] = cd{0.5 * chi[5], -0.5 * chi[6]};
    H.e[7] = cd{0.5 * chi[5], 0.5 * chi[6]};
    H.e[8] = cd{-2.0 * s3 * chi[7], 0.};
    bool ok = true;
    M3 g = expm_pade6(mul_i(1.0, H), &ok);
    if (!ok) return m_one();
    g = reunitarize(g);
    if (!(g.e[2].r == g.e[2].r) || !(g.e[2].i == g.e[2].i)) return m_one();
    return g;
}

}  // namespace ipggf
