/* orc_su3.c -- oracle SU(3) exponential.  TEST INFRASTRUCTURE ONLY.
 *
 * orc_expm_coeff restates Matrix::expmCoeff (Matrix.cpp:310-426): the
 * Cayley-Hamilton closed form of exp(i Q^a t^a) (Morningstar & Peardon,
 * Phys. Rev. D69 (2004) 054501) returning the 9 coefficients u0, u_a with
 * exp(i Q t) = u0 * 1 + sum_a u_a t^a, with NaN coefficients replaced by 0.
 * orc_u_from_exponent restates Init::getUfromExponent (Init.cpp:3483-3499):
 * the identity is returned when |Re u0| < 1e-15.
 * Pinned by: scipy.linalg.expm(i Q^a t^a), unitarity and det = 1
 * (tests/test_oracle_su3.py).
 */
#include <math.h>

#include "ipg_oracle.h"
#include "orc_internal.h"

void orc_generator(int a, M3 *t) {
    *t = m_zero();
    const double s3 = 1.0 / (2.0 * sqrt(3.0));
    switch (a) {
        case 0: t->e[1] = 0.5; t->e[3] = 0.5; break;
        case 1: t->e[1] = CMPLX(0, -0.5); t->e[3] = CMPLX(0, 0.5); break;
        case 2: t->e[0] = 0.5; t->e[4] = -0.5; break;
        case 3: t->e[2] = 0.5; t->e[6] = 0.5; break;
        case 4: t->e[2] = CMPLX(0, -0.5); t->e[6] = CMPLX(0, 0.5); break;
        case 5: t->e[5] = 0.5; t->e[7] = 0.5; break;
        case 6: t->e[5] = CMPLX(0, -0.5); t->e[7] = CMPLX(0, 0.5); break;
        default:
            t->e[0] = s3;
            t->e[4] = s3;
            t->e[8] = -1.0 / sqrt(3.0);
            break;
    }
}

void orc_expm_coeff(const double Q[8], double coeff[18]) {
    const double r3 = sqrt(3.0);
    /* cubic invariant c0 = det(Q.t) written out in the t^a basis */
    double c0 = r3 * (Q[0] * Q[0] * Q[7] + Q[1] * Q[1] * Q[7] + Q[2] * Q[2] * Q[7]);
    c0 -= Q[7] * Q[7] * Q[7] / r3;
    c0 -= (r3 / 2.0) * (Q[3] * Q[3] * Q[7] + Q[4] * Q[4] * Q[7] +
                        Q[5] * Q[5] * Q[7] + Q[6] * Q[6] * Q[7]);
    c0 += 3.0 * (Q[0] * Q[3] * Q[5] + Q[0] * Q[4] * Q[6] + Q[1] * Q[4] * Q[5] -
                 Q[1] * Q[3] * Q[6]);
    c0 += 1.5 * (Q[2] * Q[3] * Q[3] + Q[2] * Q[4] * Q[4] - Q[2] * Q[5] * Q[5] -
                 Q[2] * Q[6] * Q[6]);
    c0 /= 12.0;
    /* quadratic invariant c1 = tr(Q.t)^2 / 2 */
    double c1 = 0.0;
    for (int a = 0; a < 8; a++) c1 += Q[a] * Q[a];
    c1 *= 0.25;

    /* pow(c1/3, 1.5) written as x*sqrt(x): two correctly rounded operations,
     * so the degenerate-eigenvalue case |c0/c0max| ~ 1 (where acos turns NaN)
     * is decided identically on every platform */
    const double x3 = c1 / 3.0;
    double c0max = fmax(1e-15, 2.0 * (x3 * sqrt(x3)));
    double th3 = acos(c0 / c0max) / 3.0;
    double u = sqrt(c1 / 3.0) * cos(th3);
    double w = sqrt(c1) * sin(th3);
    double xi0 = sin(w) / w;
    double den = 9.0 * u * u - w * w;
    cplx iu = CMPLX(0.0, 1.0) * u;
    double cw = cos(w);
    cplx e2iu = cexp(2.0 * iu);
    cplx emiu = cexp(-iu);

    cplx f0 = (u * u - w * w) * e2iu +
              emiu * (8.0 * u * u * cw + 2.0 * iu * xi0 * (3.0 * u * u + w * w));
    f0 /= den;
    cplx f1 = 2.0 * u * e2iu -
              emiu * (2.0 * u * cw - CMPLX(0.0, 1.0) * (3.0 * u * u - w * w) * xi0);
    f1 /= den;
    cplx f2 = e2iu - emiu * (cw + 3.0 * iu * xi0);
    f2 /= den;

    cplx u0 = f0 + 2.0 / 3.0 * c1 * f2;
    f1 /= (0.5 * f2);
    cplx ua[8];
    for (int i = 0; i < 8; i++) ua[i] = f1 * Q[i];
    /* d-symbol contractions d_abc Q^b Q^c */
    ua[0] += (Q[3] * Q[5] + Q[4] * Q[6] + 2.0 / r3 * Q[0] * Q[7]);
    ua[1] += (2.0 * Q[1] * Q[7] / r3 - Q[3] * Q[6] + Q[4] * Q[5]);
    ua[2] += (2.0 * Q[2] * Q[7] / r3 + 0.5 * Q[3] * Q[3] + 0.5 * Q[4] * Q[4] -
              0.5 * Q[5] * Q[5] - 0.5 * Q[6] * Q[6]);
    ua[3] += (-1.0 / r3 * Q[3] * Q[7] + Q[0] * Q[5] - Q[1] * Q[6] + Q[2] * Q[3]);
    ua[4] += (-1.0 / r3 * Q[4] * Q[7] + Q[0] * Q[6] + Q[1] * Q[5] + Q[2] * Q[4]);
    ua[5] += (-1.0 / r3 * Q[5] * Q[7] + Q[0] * Q[3] + Q[1] * Q[4] - Q[2] * Q[5]);
    ua[6] += (-1.0 / r3 * Q[6] * Q[7] + Q[0] * Q[4] - Q[1] * Q[3] - Q[2] * Q[6]);
    ua[7] += (Q[0] * Q[0] + Q[1] * Q[1] + Q[2] * Q[2] - Q[7] * Q[7] -
              0.5 * Q[3] * Q[3] - 0.5 * Q[4] * Q[4] - 0.5 * Q[5] * Q[5] -
              0.5 * Q[6] * Q[6]) / r3;

    cplx out[9];
    out[0] = u0;
    for (int i = 0; i < 8; i++) out[i + 1] = ua[i] * 0.5 * f2;
    for (int i = 0; i < 9; i++) {
        if (isnan(creal(out[i])) || isnan(cimag(out[i]))) out[i] = 0.0;
        coeff[2 * i] = creal(out[i]);
        coeff[2 * i + 1] = cimag(out[i]);
    }
}

M3 orc_exp_iQt(const double Q[8]) {
    double c[18];
    orc_expm_coeff(Q, c);
    if (fabs(c[0]) < 1e-15) return m_one();
    cplx u[9];
    for (int i = 0; i < 9; i++) u[i] = CMPLX(c[2 * i], c[2 * i + 1]);
    /* u0*1 + sum_a u_{a+1} t^a, entry by entry (zero terms dropped) */
    const double s3 = 1.0 / (2.0 * sqrt(3.0));
    const double t8 = -1.0 / sqrt(3.0);
    M3 U;
    U.e[0] = (u[0] + u[3] * 0.5) + u[8] * s3;
    U.e[1] = u[1] * 0.5 + u[2] * CMPLX(0, -0.5);
    U.e[2] = u[4] * 0.5 + u[5] * CMPLX(0, -0.5);
    U.e[3] = u[1] * 0.5 + u[2] * CMPLX(0, 0.5);
    U.e[4] = (u[0] + u[3] * -0.5) + u[8] * s3;
    U.e[5] = u[6] * 0.5 + u[7] * CMPLX(0, -0.5);
    U.e[6] = u[4] * 0.5 + u[5] * CMPLX(0, 0.5);
    U.e[7] = u[6] * 0.5 + u[7] * CMPLX(0, 0.5);
    U.e[8] = u[0] + u[8] * t8;
    return U;
}

void orc_u_from_exponent(const double Q[8], double U[18]) {
    m_store(U, orc_exp_iQt(Q));
}

void orc_matmul(const double *A, const double *B, double *C) {
    m_store(C, m_mul(m_load(A), m_load(B)));
}
