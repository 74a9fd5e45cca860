// Host build of ipglasma_amd/csrc/adjoint.hpp: the per-site forward and
// reverse functions over flat arrays (matrices as 18 doubles, row-major,
// interleaved re/im), for finite-difference checks.
#include "../../ipglasma_amd/csrc/adjoint.hpp"

using namespace ipgad;

static M3 ld(const double* p) {
    M3 m;
    for (int k = 0; k < 9; k++) m.e[k] = cd{p[2 * k], p[2 * k + 1]};
    return m;
}
static void st(double* p, const M3& m) {
    for (int k = 0; k < 9; k++) {
        p[2 * k] = m.e[k].r;
        p[2 * k + 1] = m.e[k].i;
    }
}
template <class T>
static void unpack(const double* x, T& q, int n) {
    M3* m = reinterpret_cast<M3*>(&q);
    for (int k = 0; k < n; k++) m[k] = ld(x + 18 * k);
}
template <class T>
static void pack(double* x, const T& q, int n) {
    const M3* m = reinterpret_cast<const M3*>(&q);
    for (int k = 0; k < n; k++) st(x + 18 * k, m[k]);
}

extern "C" {
// pi: 10 input matrices, 1 output
void adjh_pi(const double* x, double cp, const double* yb, double* y, double* xb) {
    PiIn q;
    unpack(x, q, 10);
    st(y, pi_fwd(q, cp));
    PiBar b;
    for (int k = 0; k < 10; k++) reinterpret_cast<M3*>(&b)[k] = m_zero();
    pi_adj(q, cp, ld(yb), b);
    pack(xb, b, 10);
}
// E: 15 inputs, 2 outputs
void adjh_e(const double* x, double tau, double dtau, double g, const double* yb, double* y, double* xb) {
    EIn q;
    unpack(x, q, 15);
    M3 a, c;
    e_fwd(q, tau, dtau, g, a, c);
    st(y, a);
    st(y + 18, c);
    EBar b;
    for (int k = 0; k < 15; k++) reinterpret_cast<M3*>(&b)[k] = m_zero();
    e_adj(q, tau, dtau, g, ld(yb), ld(yb + 18), b);
    pack(xb, b, 15);
}
// U update: inputs U, E; output U'
void adjh_u(const double* x, double kappa, const double* yb, double* y, double* xb) {
    const M3 U = ld(x), E = ld(x + 18);
    st(y, u_fwd(U, E, kappa));
    M3 Ub = m_zero(), Eb = m_zero();
    u_adj(U, E, kappa, ld(yb), Ub, Eb);
    st(xb, Ub);
    st(xb + 18, Eb);
}
// T^tautau: 16 inputs, scalar output
double adjh_tt(const double* x, int plaq, double tau, double g, double yb, double* xb) {
    TIn q;
    unpack(x, q, 16);
    q.plaq = plaq;
    const double y = tt_fwd(q, tau, g);
    TBar b;
    for (int k = 0; k < 16; k++) reinterpret_cast<M3*>(&b)[k] = m_zero();
    tt_adj(q, tau, g, yb, b);
    pack(xb, b, 16);
    return y;
}
// root finder: inputs U1, U2 and the solution U; Ub -> U1b, U2b
void adjh_newton(const double* u1, const double* u2, const double* u, const double* ub, double* u1b,
                 double* u2b) {
    M3 b1 = m_zero(), b2 = m_zero();
    newton_adj(ld(u1), ld(u2), ld(u), ld(ub), b1, b2);
    st(u1b, b1);
    st(u2b, b2);
}
// one direction of the pi assembly: inputs D, U, Dm, Um; output acc
void adjh_acc(const double* x, const double* ab, double* y, double* xb) {
    const M3 D = ld(x), U = ld(x + 18), Dm = ld(x + 36), Um = ld(x + 54);
    st(y, acc_dir(D, U, Dm, Um));
    M3 b[4] = {m_zero(), m_zero(), m_zero(), m_zero()};
    acc_dir_adj(D, U, Dm, Um, ld(ab), b[0], b[1], b[2], b[3]);
    for (int k = 0; k < 4; k++) st(xb + 18 * k, b[k]);
}
// the exponential from the reverse's coefficients (expq_value), against exp_iQt
void adjh_expq_value(const double* q, double* e) {
    double Q[8];
    for (int a = 0; a < 8; a++) Q[a] = q[a];
    ExpQ x;
    expq_prep(Q, x);
    st(e, expq_value(x));
}
// the forward slices' exponential (expq_forward)
void adjh_expq_forward(const double* q, double* e) {
    double Q[8];
    for (int a = 0; a < 8; a++) Q[a] = q[a];
    st(e, expq_forward(Q));
}
// the setV slice step V <- exp(i Q^a t^a) V (expq_apply) and det Q (gen_det)
void adjh_expq_apply(const double* q, const double* v, double* out) {
    double Q[8];
    for (int a = 0; a < 8; a++) Q[a] = q[a];
    M3 V = ld(v);
    expq_apply(Q, V);
    st(out, V);
}
double adjh_gen_det(const double* q) {
    double Q[8];
    for (int a = 0; a < 8; a++) Q[a] = q[a];
    return gen_det(Q);
}
void adjh_expq(const double* q, const double* eb, double* e, double* qb) {
    double Q[8], Qb[8];
    for (int a = 0; a < 8; a++) Q[a] = q[a];
    st(e, exp_iQt(Q));
    expq_adj(Q, ld(eb), Qb);
    for (int a = 0; a < 8; a++) qb[a] = Qb[a];
}
}
