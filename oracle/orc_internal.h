/* orc_internal.h -- 3x3 complex helpers shared by the oracle sources.
 * TEST INFRASTRUCTURE ONLY.
 *
 * The arithmetic follows Matrix.cpp:100-306 (operator*, conjg, prodABconj,
 * prodAconjB) and Matrix.h:135-178 (+=, -=, square) term by term and in the
 * same summation order, so that the oracle reproduces the reference's
 * rounding wherever the reference's own expression order is known.
 */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include <complex.h>
#include <string.h>

typedef double complex cplx;

/* coefficient factors for mutation tests (orc_set_mutation); all 1.0 unless
 * a test plants a deliberate mistake, and x * 1.0 == x, so they change
 * nothing otherwise */
extern double orc_mut[32];

typedef struct {
    cplx e[9];
} M3;

static inline M3 m_load(const double *p) {
    M3 m;
    for (int i = 0; i < 9; i++) m.e[i] = CMPLX(p[2 * i], p[2 * i + 1]);
    return m;
}
static inline void m_store(double *p, M3 m) {
    for (int i = 0; i < 9; i++) {
        p[2 * i] = creal(m.e[i]);
        p[2 * i + 1] = cimag(m.e[i]);
    }
}
static inline M3 m_zero(void) {
    M3 m;
    for (int i = 0; i < 9; i++) m.e[i] = 0.0;
    return m;
}
static inline M3 m_one(void) {
    M3 m = m_zero();
    m.e[0] = m.e[4] = m.e[8] = 1.0;
    return m;
}
/* a*b (Matrix.cpp:100-138) */
static inline M3 m_mul(M3 a, M3 b) {
    M3 c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            c.e[3 * i + j] = a.e[3 * i] * b.e[j] + a.e[3 * i + 1] * b.e[3 + j] +
                             a.e[3 * i + 2] * b.e[6 + j];
    return c;
}
/* a*b^dagger (Matrix.cpp:209-248) */
static inline M3 m_mul_bd(M3 a, M3 b) {
    M3 c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            c.e[3 * i + j] = a.e[3 * i] * conj(b.e[3 * j]) +
                             a.e[3 * i + 1] * conj(b.e[3 * j + 1]) +
                             a.e[3 * i + 2] * conj(b.e[3 * j + 2]);
    return c;
}
/* a^dagger*b (Matrix.cpp:250-289) */
static inline M3 m_mul_ad(M3 a, M3 b) {
    M3 c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            c.e[3 * i + j] = conj(a.e[i]) * b.e[j] + conj(a.e[3 + i]) * b.e[3 + j] +
                             conj(a.e[6 + i]) * b.e[6 + j];
    return c;
}
/* hermitian conjugate (Matrix.cpp:186-207) */
static inline M3 m_dag(M3 a) {
    M3 c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c.e[3 * j + i] = conj(a.e[3 * i + j]);
    return c;
}
static inline M3 m_add(M3 a, M3 b) {
    for (int i = 0; i < 9; i++) a.e[i] += b.e[i];
    return a;
}
static inline M3 m_sub(M3 a, M3 b) {
    for (int i = 0; i < 9; i++) a.e[i] -= b.e[i];
    return a;
}
static inline M3 m_scale(double s, M3 a) {
    for (int i = 0; i < 9; i++) a.e[i] = a.e[i] * s;
    return a;
}
static inline M3 m_cscale(cplx s, M3 a) {
    for (int i = 0; i < 9; i++) a.e[i] = a.e[i] * s;
    return a;
}
static inline cplx m_tr(M3 a) { return a.e[0] + a.e[4] + a.e[8]; }
/* tr(a*b) summed as in Matrix::traceOfProdcutOfMatrix (Matrix.cpp:594-604) */
static inline cplx m_trprod(M3 a, M3 b) {
    cplx t = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t += a.e[3 * i + j] * b.e[3 * j + i];
    return t;
}
/* a - tr(a)/3 * 1 as written at Evolution.cpp:257-258 */
static inline M3 m_traceless(M3 a) {
    cplx t = m_tr(a);
    M3 o = m_one();
    o = m_cscale(t / 3.0, o);
    return m_sub(a, o);
}
/* 0.5 * sum |e|^2 (Matrix.h:172-178) */
static inline double m_square(M3 a) {
    double t = 0.0;
    for (int i = 0; i < 9; i++) t += creal(a.e[i]) * creal(a.e[i]) + cimag(a.e[i]) * cimag(a.e[i]);
    return 0.5 * t;
}
static inline M3 m_mul3(M3 a, M3 b, M3 c) { return m_mul(m_mul(a, b), c); }
static inline M3 m_mul4(M3 a, M3 b, M3 c, M3 d) {
    return m_mul(m_mul(m_mul(a, b), c), d);
}

/* SU(3) generators t^a = lambda^a / 2 (Group.cpp:68-149) */
void orc_generator(int a, M3 *t);
M3 orc_exp_iQt(const double Q[8]);

#endif
