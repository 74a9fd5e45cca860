// kernels_adjoint.hip -- reverse mode of dE/deta through the leapfrog.
//
// dE/deta = tau a^2 hbarc sum_s T^tautau(s) at it = itmax is a function of the
// state after step itmax-1, which the leapfrog (Evolution.cpp:366-431) builds
// from the tau = 0 fields.  The backward pass walks the steps in reverse: for
// step it it needs the state before and after the step, which the host
// recomputes segment by segment from checkpoints (every K = ceil(sqrt(T))
// steps).
//
// No atomics in the step: every site writes the cotangents of its own fields
// ("self") and, per neighbour offset, the cotangent it sends to that
// neighbour ("outbox"); the next reverse step gathers the outboxes of the
// sites whose clamped neighbour it is.  One step runs as three site-parallel
// passes (local updates; the three plaquettes of E'; the commutators of E'
// and the pi' stencil) so that few matrices are live at a time.
//
// A state slot is 72 planes of S doubles: U_x, U_y (18 each, full complex),
// E1, E2, phi, pi (9 each, Hermitian storage).  A compact cotangent set is
// 108 planes: U_x, U_y, E1, E2, phi, pi, all full complex (18 each).  A
// pending set is kPend planes:
//   self     U_x, U_y (18 each), E1, E2, phi, pi (9 each, Hermitian)
//   outbox   U_x to (0,+1) (0,-1) (-1,+1) (-1,0)   4 x 18
//            U_y to (+1,0) (+1,-1) (0,-1) (-1,0)   4 x 18
//            phi to (+1,0) (0,+1) (-1,0) (0,-1)    4 x 9 (Hermitian)
// The cotangents of E, phi and pi are used only through their Hermitian
// part, so self and outbox keep that part.
#include "adjoint.hpp"
#include "ipg_device.hpp"
#include "ipg_internal.hpp"

namespace {

using namespace ipgad;

struct Slot {
    const double* Ux;
    const double* Uy;
    const double* E1;
    const double* E2;
    const double* phi;
    const double* pi;
};
__device__ __forceinline__ Slot slot_of(const double* base, size_t S) {
    return Slot{base, base + 18 * S, base + 36 * S, base + 45 * S, base + 54 * S, base + 63 * S};
}
// cotangent set: field f (0 Ux, 1 Uy, 2 E1, 3 E2, 4 phi, 5 pi)
__device__ __forceinline__ double* cot(double* base, size_t S, int f) { return base + (size_t)f * 18 * S; }

// pending set planes
constexpr int kPU1 = 0, kPU2 = 18, kPE1 = 36, kPE2 = 45, kPPh = 54, kPPi = 63;
constexpr int kPOx = 72, kPOy = 144, kPOph = 216;
static_assert(kPOph + 36 == IPG_PEND_PLANES, "pending set layout");

__device__ __forceinline__ void atomic_add_m3(double* f, size_t S, size_t s, const M3& m) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
        atomicAdd(f + (2 * k) * S + s, m.e[k].r);
        atomicAdd(f + (2 * k + 1) * S + s, m.e[k].i);
    }
}
__device__ __forceinline__ void add_m3(double* f, size_t S, size_t s, const M3& m) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
        f[(2 * k) * S + s] += m.e[k].r;
        f[(2 * k + 1) * S + s] += m.e[k].i;
    }
}

// adds hproj(m) to a Hermitian-storage plane group
__device__ __forceinline__ void atomic_add_h(double* f, size_t S, size_t s, const M3& m) {
    const M3 h = hproj(m);
    atomicAdd(f + s, h.e[0].r);
    atomicAdd(f + S + s, h.e[4].r);
    atomicAdd(f + 2 * S + s, h.e[8].r);
    atomicAdd(f + 3 * S + s, h.e[1].r);
    atomicAdd(f + 4 * S + s, h.e[1].i);
    atomicAdd(f + 5 * S + s, h.e[2].r);
    atomicAdd(f + 6 * S + s, h.e[2].i);
    atomicAdd(f + 7 * S + s, h.e[5].r);
    atomicAdd(f + 8 * S + s, h.e[5].i);
}
__device__ __forceinline__ void add_h(double* f, size_t S, size_t s, const M3& h) {
    sth<true>(f, S, s, ldh<true>(f, S, s) + h);
}

// The sites u along one axis whose clamped neighbour clamp(u + d) is t: one
// in the interior, two or none at the edges.
struct Src {
    int a, b, n;
};
__device__ __forceinline__ Src srcs(int t, int d, int N) {
    if (d == 0) return Src{t, t, 1};
    if (d > 0) return t == N - 1 ? Src{t - 1, t, 2} : (t >= 1 ? Src{t - 1, t - 1, 1} : Src{0, 0, 0});
    return t == 0 ? Src{1, 0, 2} : (t <= N - 2 ? Src{t + 1, t + 1, 1} : Src{0, 0, 0});
}
// what site (i, j) receives through one outbox offset (di, dj)
template <bool HERM>
__device__ __forceinline__ M3 gather(const double* f, size_t S, int N, int i, int j, int di, int dj) {
    const Src a = srcs(i, di, N), b = srcs(j, dj, N);
    M3 r = m_zero();
    for (int p = 0; p < a.n; p++) {
        const int si = p ? a.b : a.a;
        for (int q = 0; q < b.n; q++) {
            const int sj = q ? b.b : b.a;
            const size_t u = (size_t)si * N + sj;
            r = r + (HERM ? ldh<true>(f, S, u) : ldm<true>(f, S, u));
        }
    }
    return r;
}
// total cotangents of U_x, U_y and phi at (i, j) = s in a pending set
__device__ __forceinline__ M3 cin_ux(const double* P, size_t S, int N, int i, int j, size_t s) {
    M3 r = ldm<true>(P + kPU1 * S, S, s);
    r = r + gather<false>(P + (kPOx + 0) * S, S, N, i, j, 0, 1);
    r = r + gather<false>(P + (kPOx + 18) * S, S, N, i, j, 0, -1);
    r = r + gather<false>(P + (kPOx + 36) * S, S, N, i, j, -1, 1);
    r = r + gather<false>(P + (kPOx + 54) * S, S, N, i, j, -1, 0);
    return r;
}
__device__ __forceinline__ M3 cin_uy(const double* P, size_t S, int N, int i, int j, size_t s) {
    M3 r = ldm<true>(P + kPU2 * S, S, s);
    r = r + gather<false>(P + (kPOy + 0) * S, S, N, i, j, 1, 0);
    r = r + gather<false>(P + (kPOy + 18) * S, S, N, i, j, 1, -1);
    r = r + gather<false>(P + (kPOy + 36) * S, S, N, i, j, 0, -1);
    r = r + gather<false>(P + (kPOy + 54) * S, S, N, i, j, -1, 0);
    return r;
}
__device__ __forceinline__ M3 cin_ph(const double* P, size_t S, int N, int i, int j, size_t s) {
    M3 r = ldh<true>(P + kPPh * S, S, s);
    r = r + gather<true>(P + (kPOph + 0) * S, S, N, i, j, 1, 0);
    r = r + gather<true>(P + (kPOph + 9) * S, S, N, i, j, 0, 1);
    r = r + gather<true>(P + (kPOph + 18) * S, S, N, i, j, -1, 0);
    r = r + gather<true>(P + (kPOph + 27) * S, S, N, i, j, 0, -1);
    return r;
}

// the site of this thread in a launch of pad8(ceil(S / IPG_BLOCK)) blocks:
// the XCD-aware block order of the forward kernels (xcd_remap), so that the
// blocks of neighbouring lattice rows run on one XCD and share its L2
__device__ __forceinline__ long long xcd_site() {
    return (long long)xcd_remap((int)blockIdx.x, (int)gridDim.x) * IPG_BLOCK + threadIdx.x;
}

struct Lat {
    int N;
    long long S;
    __device__ __forceinline__ size_t at(int i, int j) const {
        return (size_t)clampi(i, N) * N + clampi(j, N);
    }
};

// ---- state copies between the live buffers and a slot -----------------------------
// A state slot is 72 planes of S doubles: U_x, U_y (18 each), E1, E2, phi,
// pi (9 each).  One thread per double; S is a multiple of the block size
// (N >= 16), so each block stays inside one plane.
__device__ __forceinline__ double* state_plane(const EvDesc& ev, int parity, int q, int* k) {
    if (q < 18) return *k = q, ev.U[parity][0];
    if (q < 36) return *k = q - 18, ev.U[parity][1];
    if (q < 45) return *k = q - 36, ev.E1;
    if (q < 54) return *k = q - 45, ev.E2;
    if (q < 63) return *k = q - 54, ev.phi[parity];
    return *k = q - 63, ev.pi;
}
__global__ __launch_bounds__(IPG_BLOCK) void k_state_save(const EvDesc* __restrict__ evs, int parity,
                                                         double* __restrict__ slot) {
    const EvDesc& ev = evs[0];
    const long long S = ev.S;
    const long long t = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (t >= 72 * S) return;
    const int q = (int)(t / S);
    const long long s = t - (long long)q * S;
    int k = 0;
    const double* src = state_plane(ev, parity, q, &k);
    slot[t] = src[(size_t)k * S + s];
}
__global__ __launch_bounds__(IPG_BLOCK) void k_state_load(const EvDesc* __restrict__ evs, int parity,
                                                         const double* __restrict__ slot) {
    const EvDesc& ev = evs[0];
    const long long S = ev.S;
    const long long t = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (t >= 72 * S) return;
    const int q = (int)(t / S);
    const long long s = t - (long long)q * S;
    int k = 0;
    double* dst = state_plane(ev, parity, q, &k);
    dst[(size_t)k * S + s] = slot[t];
}

__global__ __launch_bounds__(IPG_BLOCK) void k_zero(double* __restrict__ p, long long n) {
    const long long t = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (t < n) p[t] = 0.;
}

// ---- loss: dL/dT^tautau(s) = tau_fm hbarc / a^2 for every site --------------------
// into the own-field planes of a zeroed pending set
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_loss(const EvDesc* __restrict__ evs, const double* __restrict__ xs,
                                                       int it, double yb, double* __restrict__ cout) {
    const EvDesc& ev = evs[0];
    const Lat L{ev.N, ev.S};
    const long long s = xcd_site();
    if (s >= L.S) return;
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const Slot X = slot_of(xs, S);
    const size_t pX = L.at(i + 1, j), pY = L.at(i, j + 1), pXY = L.at(i + 1, j + 1);
    TIn q;
    q.e1 = ldh<true>(X.E1, S, s);
    q.e1py = ldh<true>(X.E1, S, pY);
    q.e2 = ldh<true>(X.E2, S, s);
    q.e2px = ldh<true>(X.E2, S, pX);
    q.p = ldh<true>(X.pi, S, s);
    q.ppx = ldh<true>(X.pi, S, pX);
    q.ppy = ldh<true>(X.pi, S, pY);
    q.ppxy = ldh<true>(X.pi, S, pXY);
    q.ux = ldm<true>(X.Ux, S, s);
    q.uy = ldm<true>(X.Uy, S, s);
    q.uy_px = ldm<true>(X.Uy, S, pX);
    q.ux_py = ldm<true>(X.Ux, S, pY);
    q.ph = ldh<true>(X.phi, S, s);
    q.phpx = ldh<true>(X.phi, S, pX);
    q.phpy = ldh<true>(X.phi, S, pY);
    q.phpxy = ldh<true>(X.phi, S, pXY);
    q.plaq = i != N - 1;
    TBar b{};  // zeros (named members: registers, not an indexed array)
    tt_adj(q, it * ev.dtau, ev.g, yb, b);
    double* cU1 = cout + kPU1 * S;
    double* cU2 = cout + kPU2 * S;
    double* cE1 = cout + kPE1 * S;
    double* cE2 = cout + kPE2 * S;
    double* cPh = cout + kPPh * S;
    double* cPi = cout + kPPi * S;
    atomic_add_h(cE1, S, s, b.e1);
    atomic_add_h(cE1, S, pY, b.e1py);
    atomic_add_h(cE2, S, s, b.e2);
    atomic_add_h(cE2, S, pX, b.e2px);
    atomic_add_h(cPi, S, s, b.p);
    atomic_add_h(cPi, S, pX, b.ppx);
    atomic_add_h(cPi, S, pY, b.ppy);
    atomic_add_h(cPi, S, pXY, b.ppxy);
    atomic_add_m3(cU1, S, s, b.ux);
    atomic_add_m3(cU2, S, s, b.uy);
    atomic_add_m3(cU2, S, pX, b.uy_px);
    atomic_add_m3(cU1, S, pY, b.ux_py);
    atomic_add_h(cPh, S, s, b.ph);
    atomic_add_h(cPh, S, pX, b.phpx);
    atomic_add_h(cPh, S, pY, b.phpy);
    atomic_add_h(cPh, S, pXY, b.phpxy);
}

// ---- T^{mu nu} at itmax in reverse (Evolution.cpp:659-1246) -----------------------
// tb: cotangents of the ten stored components (T's order, after the a^-4,
// a^-5, a^-6 scalings) at every site; the cotangents of the fields of the
// state xs go into the own-field planes of a zeroed pending set (atomics:
// each site's terms reach up to two sites away).  Three passes as the
// forward's: diagonal components with T^{tau eta} and T^{xy}; the clovers of
// T^{tau x}, T^{tau y}, T^{x eta}, T^{y eta}; their pi and E pi parts.
// f = Re tr(A B): Ab += c B^+, Bb += c A^+;  f = Im tr(A B): Ab += c i B^+, Bb += c i A^+.

struct TScat {
    double* P;
    size_t S;
    __device__ void u1(size_t q, const M3& m) const { atomic_add_m3(P + kPU1 * S, S, q, m); }
    __device__ void u2(size_t q, const M3& m) const { atomic_add_m3(P + kPU2 * S, S, q, m); }
    __device__ void e1(size_t q, const M3& m) const { atomic_add_h(P + kPE1 * S, S, q, m); }
    __device__ void e2(size_t q, const M3& m) const { atomic_add_h(P + kPE2 * S, S, q, m); }
    __device__ void ph(size_t q, const M3& m) const { atomic_add_h(P + kPPh * S, S, q, m); }
    __device__ void pi(size_t q, const M3& m) const { atomic_add_h(P + kPPi * S, S, q, m); }
};

// the phi gradient G = U phi(n) U^+ - phi(m) in reverse: U at uq (field f:
// 1 = U_x, 2 = U_y), given Gb
__device__ __forceinline__ void grad_adj(const TScat& sc, const Slot& X, size_t S, int f, size_t uq, size_t n,
                                         size_t m, const M3& Gb) {
    const M3 u = ldm<true>(f == 1 ? X.Ux : X.Uy, S, uq);
    M3 ub = m_zero(), pb = m_zero();
    adj_conj(u, ldh<true>(X.phi, S, n), Gb, ub, pb);
    if (f == 1)
        sc.u1(uq, ub);
    else
        sc.u2(uq, ub);
    sc.ph(n, pb);
    sc.ph(m, -1. * Gb);
}
__device__ __forceinline__ M3 grad_fwd(const Slot& X, size_t S, int f, size_t uq, size_t n, size_t m) {
    const M3 u = ldm<true>(f == 1 ? X.Ux : X.Uy, S, uq);
    return (u * ldh<true>(X.phi, S, n)) * dag(u) - ldh<true>(X.phi, S, m);
}

__global__ __launch_bounds__(IPG_BLOCK) void k_adj_tmunu_a(const EvDesc* __restrict__ evs, const double* __restrict__ xs,
                                                          int it, const double* __restrict__ tb,
                                                          double* __restrict__ pout) {
    const EvDesc& ev = evs[0];
    const Lat L{ev.N, ev.S};
    const long long s = xcd_site();
    if (s >= L.S) return;
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const Slot X = slot_of(xs, S);
    const TScat sc{pout, S};
    const double tau = it * ev.dtau, g = ev.g, a = ev.a;
    const double a2 = a * a, a4 = a2 * a2, a5 = a4 * a, a6 = a4 * a2;
    const size_t pX = L.at(i + 1, j), pY = L.at(i, j + 1), pXY = L.at(i + 1, j + 1);
    const double b0 = tb[s] / a4, b1 = tb[S + s] / a4, b2 = tb[2 * S + s] / a4, b3 = tb[3 * S + s] / a6;
    const double gt = g * g / tau / tau, it2 = 1. / tau / tau, h = 0.5 / tau / tau;
    const double t1b = gt / 2. * (b0 - b1 + b2 + it2 * b3), t2b = gt / 2. * (b0 + b1 - b2 + it2 * b3);
    const double lb = b0 + b1 + b2 - it2 * b3;  // the pi sum and the plaquette term
    const double Ab = h * (b0 + b1 - b2 + it2 * b3), Cb = h * (b0 - b1 + b2 + it2 * b3);
    const double reb = tb[6 * S + s] * g / (tau * tau * tau) / a5;  // T^{tau eta}
    const double bb = tb[7 * S + s] * 0.5 / (tau * tau) / a4;       // T^{xy}
    // pi: sum Re tr(pi^2) / 4
    sc.pi(s, (lb / 2.) * ldh<true>(X.pi, S, s));
    sc.pi(pX, (lb / 2.) * ldh<true>(X.pi, S, pX));
    sc.pi(pY, (lb / 2.) * ldh<true>(X.pi, S, pY));
    sc.pi(pXY, (lb / 2.) * ldh<true>(X.pi, S, pXY));
    // plaquette: 2/g^2 (3 - Re tr(ux uy(+x) ux(+y)^+ uy^+)), 1 on the last row
    if (i != N - 1) {
        M3 B1, B2, B3, B4;
        adj_prod4(ldm<true>(X.Ux, S, s), ldm<true>(X.Uy, S, pX), dag(ldm<true>(X.Ux, S, pY)), dag(ldm<true>(X.Uy, S, s)),
                  (-2. / (g * g) * lb) * m_one(), B1, B2, B3, B4);
        sc.u1(s, B1);
        sc.u2(pX, B2);
        sc.u1(pY, dag(B3));
        sc.u2(s, dag(B4));
    }
    const M3 e1 = ldh<true>(X.E1, S, s), e1p = ldh<true>(X.E1, S, pY), e2 = ldh<true>(X.E2, S, s), e2p = ldh<true>(X.E2, S, pX);
    M3 e1b = (2. * t1b) * e1, e1pb = (2. * t1b) * e1p, e2b = (2. * t2b) * e2, e2pb = (2. * t2b) * e2p;
    const M3 gx = grad_fwd(X, S, 1, s, pX, s), gy = grad_fwd(X, S, 2, s, pY, s);
    const M3 gxy1 = grad_fwd(X, S, 1, pY, pXY, pY), gxy2 = grad_fwd(X, S, 2, pX, pXY, pX);
    M3 gxb = (2. * Ab) * gx, gxy1b = (2. * Ab) * gxy1, gyb = (2. * Cb) * gy, gxy2b = (2. * Cb) * gxy2;
    // T^{tau eta}: Re tr(E1 gX) + Re tr(E1(+y) gXY1) + Re tr(E2 gY) + Re tr(E2(+x) gXY2)
    e1b = e1b + reb * dag(gx);
    gxb = gxb + reb * dag(e1);
    e1pb = e1pb + reb * dag(gxy1);
    gxy1b = gxy1b + reb * dag(e1p);
    e2b = e2b + reb * dag(gy);
    gyb = gyb + reb * dag(e2);
    e2pb = e2pb + reb * dag(gxy2);
    gxy2b = gxy2b + reb * dag(e2p);
    // T^{xy}: 4 Re tr(c P Q) + Re tr(gx gy) + Re tr(yl gy) + Re tr(gx xl) + Re tr(yl xl)
    {
        const M3 ux = ldm<true>(X.Ux, S, s), uy = ldm<true>(X.Uy, S, s);
        M3 uxb = m_zero(), uyb = m_zero();
        const double c = -0.25 * g * g;
        const M3 Pm = e1 + (uy * e1p) * dag(uy), Qm = e2 + (ux * e2p) * dag(ux);
        const M3 Pb = (4. * bb * c) * dag(Qm), Qb = (4. * bb * c) * dag(Pm);
        e1b = e1b + Pb;
        adj_conj(uy, e1p, Pb, uyb, e1pb);
        e2b = e2b + Qb;
        adj_conj(ux, e2p, Qb, uxb, e2pb);
        const M3 yl = (uy * gxy1) * dag(uy), xl = (ux * gxy2) * dag(ux);
        M3 ylb = bb * (dag(gy) + dag(xl));
        M3 xlb = bb * (dag(gx) + dag(yl));
        gxb = gxb + bb * (dag(gy) + dag(xl));
        gyb = gyb + bb * (dag(gx) + dag(yl));
        adj_conj(uy, gxy1, ylb, uyb, gxy1b);
        adj_conj(ux, gxy2, xlb, uxb, gxy2b);
        sc.u1(s, uxb);
        sc.u2(s, uyb);
    }
    sc.e1(s, e1b);
    sc.e1(pY, e1pb);
    sc.e2(s, e2b);
    sc.e2(pX, e2pb);
    grad_adj(sc, X, S, 1, s, pX, s, gxb);
    grad_adj(sc, X, S, 2, s, pY, s, gyb);
    grad_adj(sc, X, S, 1, pY, pXY, pY, gxy1b);
    grad_adj(sc, X, S, 2, pX, pXY, pX, gxy2b);
}

// one clover product m4(f1 q1 [d1], ...) in reverse: factor k is field fk
// (1 = U_x, 2 = U_y) at site qk, daggered when dk
struct Fac {
    int f;
    size_t q;
    bool d;
};
__device__ __forceinline__ M3 fac_ld(const Slot& X, size_t S, const Fac& a) {
    const M3 m = ldm<true>(a.f == 1 ? X.Ux : X.Uy, S, a.q);
    return a.d ? dag(m) : m;
}
__device__ __forceinline__ void fac_st(const TScat& sc, const Fac& a, const M3& b) {
    const M3 m = a.d ? dag(b) : b;
    if (a.f == 1)
        sc.u1(a.q, m);
    else
        sc.u2(a.q, m);
}
__device__ __forceinline__ M3 prod4(const Slot& X, size_t S, const Fac (&f)[4]) {
    return ((fac_ld(X, S, f[0]) * fac_ld(X, S, f[1])) * fac_ld(X, S, f[2])) * fac_ld(X, S, f[3]);
}
__device__ __forceinline__ void prod4_adj(const TScat& sc, const Slot& X, size_t S, const Fac (&f)[4], const M3& Yb) {
    M3 B1, B2, B3, B4;
    adj_prod4(fac_ld(X, S, f[0]), fac_ld(X, S, f[1]), fac_ld(X, S, f[2]), fac_ld(X, S, f[3]), Yb, B1, B2, B3, B4);
    fac_st(sc, f[0], B1);
    fac_st(sc, f[1], B2);
    fac_st(sc, f[2], B3);
    fac_st(sc, f[3], B4);
}
// Each clover of the forward, K = TL(P1 - P2) + TL(P3 - P4), pairs a
// plaquette product with its conjugate (P2 = P1^+, P4 = P3^+), so
// K = ah(P1) + ah(P3) with ah(P) = TL(P - P^+).  Reverse of ah given Kb:
// t = TL(Kb), Pb = t - t^+ (the cotangent of P^+ is the conjugate of that
// of P).  The T^{tau x} and T^{tau y} clovers at s share the plaquette at s
// in the same cyclic order (with opposite signs), so the four clovers need
// seven products and seven product reverses.
__device__ __forceinline__ M3 ah(const M3& p) { return traceless(p - dag(p)); }
__device__ __forceinline__ M3 ah_bar(const M3& kb) {
    const M3 t = traceless(kb);
    return t - dag(t);
}

__global__ __launch_bounds__(IPG_BLOCK) void k_adj_tmunu_b(const EvDesc* __restrict__ evs, const double* __restrict__ xs,
                                                          int it, const double* __restrict__ tb,
                                                          double* __restrict__ pout) {
    const EvDesc& ev = evs[0];
    const Lat L{ev.N, ev.S};
    const long long s = xcd_site();
    if (s >= L.S) return;
    const size_t S = L.S;
    const int i = (int)(s / L.N), j = (int)(s - (long long)i * L.N);
    const Slot X = slot_of(xs, S);
    const TScat sc{pout, S};
    const double tau = it * ev.dtau, g = ev.g, a = ev.a;
    const double a4 = a * a * a * a, a5 = a4 * a;
    const size_t s0 = (size_t)s;
    const size_t pX = L.at(i + 1, j), pY = L.at(i, j + 1), pXY = L.at(i + 1, j + 1), mX = L.at(i - 1, j);
    const size_t mY = L.at(i, j - 1), mXpY = L.at(i - 1, j + 1), pXmY = L.at(i + 1, j - 1);
    const size_t p2X = L.at(i + 2, j), p2Y = L.at(i, j + 2);
    const double txb = tb[4 * S + s] / (4. * tau * a4), tyb = tb[5 * S + s] / (4. * tau * a4);
    const double xeb = -tb[8 * S + s] / (4. * g * tau * tau * a5), yeb = -tb[9 * S + s] / (4. * g * tau * tau * a5);
    // the plaquette at s from U_y(s): shared by the clovers X1 and Y1
    const Fac pa[4] = {{2, s0, false}, {1, pY, false}, {2, pX, true}, {1, s0, true}};
    const M3 ta = ah(prod4(X, S, pa));
    M3 tab;
    // X1 (tau x, x eta with E2(s), gY): K = ah(pa) + ah(pc)
    {
        const Fac pc[4] = {{1, mX, true}, {2, mX, false}, {1, mXpY, false}, {2, s0, true}};
        const M3 K = ta + ah(prod4(X, S, pc));
        const M3 e2 = ldh<true>(X.E2, S, s0), gy = grad_fwd(X, S, 2, s0, pY, s0);
        // ttx_im = Im tr(E2 K), xe_im = -Im tr(K gY)
        const M3 Kb = mul_i(txb, dag(e2)) + mul_i(-xeb, dag(gy));
        sc.e2(s0, mul_i(txb, dag(K)));
        grad_adj(sc, X, S, 2, s0, pY, s0, mul_i(-xeb, dag(K)));
        tab = ah_bar(Kb);
        prod4_adj(sc, X, S, pc, tab);
    }
    // Y1 (tau y, y eta with E1(s), gX): K = ah(pa^+) + ah(pc) = -ah(pa) + ah(pc)
    {
        const Fac pc[4] = {{2, mY, true}, {1, mY, false}, {2, pXmY, false}, {1, s0, true}};
        const M3 K = ah(prod4(X, S, pc)) - ta;
        const M3 e1 = ldh<true>(X.E1, S, s0), gx = grad_fwd(X, S, 1, s0, pX, s0);
        const M3 Kb = mul_i(tyb, dag(e1)) + mul_i(-yeb, dag(gx));
        sc.e1(s0, mul_i(tyb, dag(K)));
        grad_adj(sc, X, S, 1, s0, pX, s0, mul_i(-yeb, dag(K)));
        const M3 tb2 = ah_bar(Kb);
        prod4_adj(sc, X, S, pc, tb2);
        tab = tab - tb2;
    }
    prod4_adj(sc, X, S, pa, tab);
    // X2 (tau x, x eta with E2(+x), gXY2): K = ah(pe) + ah(pg)
    {
        const Fac pe[4] = {{2, pX, false}, {1, pXY, false}, {2, p2X, true}, {1, pX, true}};
        const Fac pg[4] = {{1, s0, true}, {2, s0, false}, {1, pY, false}, {2, pX, true}};
        const M3 K = ah(prod4(X, S, pe)) + ah(prod4(X, S, pg));
        const M3 e2 = ldh<true>(X.E2, S, pX), gg = grad_fwd(X, S, 2, pX, pXY, pX);
        const M3 Kb = mul_i(txb, dag(e2)) + mul_i(-xeb, dag(gg));
        sc.e2(pX, mul_i(txb, dag(K)));
        grad_adj(sc, X, S, 2, pX, pXY, pX, mul_i(-xeb, dag(K)));
        const M3 t = ah_bar(Kb);
        prod4_adj(sc, X, S, pe, t);
        prod4_adj(sc, X, S, pg, t);
    }
    // Y2 (tau y, y eta with E1(+y), gXY1): K = ah(qe) + ah(qg)
    {
        const Fac qe[4] = {{1, pY, false}, {2, pXY, false}, {1, p2Y, true}, {2, pY, true}};
        const Fac qg[4] = {{2, s0, true}, {1, s0, false}, {2, pX, false}, {1, pY, true}};
        const M3 K = ah(prod4(X, S, qe)) + ah(prod4(X, S, qg));
        const M3 e1 = ldh<true>(X.E1, S, pY), gg = grad_fwd(X, S, 1, pY, pXY, pY);
        const M3 Kb = mul_i(tyb, dag(e1)) + mul_i(-yeb, dag(gg));
        sc.e1(pY, mul_i(tyb, dag(K)));
        grad_adj(sc, X, S, 1, pY, pXY, pY, mul_i(-yeb, dag(K)));
        const M3 t = ah_bar(Kb);
        prod4_adj(sc, X, S, qe, t);
        prod4_adj(sc, X, S, qg, t);
    }
}

// Re tr(PI(w) (U phi(n) U^+ - V^+ phi(m) V)) in reverse with weight c; U at uq, V at vq.
// This helper and epi_adj stay out of line: inlined, their twelve calls in
// k_adj_tmunu_c keep ~1.3 KB per lane live and spill it to scratch.
__device__ __attribute__((noinline)) void pigrad_adj(const TScat& sc, const Slot& X, size_t S, int f, double c, size_t w,
                                           size_t uq, size_t n, size_t vq, size_t m) {
    const double* U = f == 1 ? X.Ux : X.Uy;
    const M3 u = ldm<true>(U, S, uq), v = ldm<true>(U, S, vq), pw = ldh<true>(X.pi, S, w);
    const M3 G = (u * ldh<true>(X.phi, S, n)) * dag(u) - (dag(v) * ldh<true>(X.phi, S, m)) * v;
    sc.pi(w, c * dag(G));
    const M3 Yb = c * dag(pw);
    M3 ub = m_zero(), vb = m_zero(), nb = m_zero(), mb = m_zero();
    adj_conj(u, ldh<true>(X.phi, S, n), Yb, ub, nb);
    adj_conjd(v, ldh<true>(X.phi, S, m), -1. * Yb, vb, mb);
    if (f == 1) {
        sc.u1(uq, ub);
        sc.u1(vq, vb);
    } else {
        sc.u2(uq, ub);
        sc.u2(vq, vb);
    }
    sc.ph(n, nb);
    sc.ph(m, mb);
}
// Re tr(E(w) (PI(w) + U PI(n) U^+)) in reverse with weight c; E field fe
__device__ __attribute__((noinline)) void epi_adj(const TScat& sc, const Slot& X, size_t S, int fe, double c, size_t w,
                                        size_t uq, size_t n) {
    const double* U = fe == 1 ? X.Ux : X.Uy;
    const M3 u = ldm<true>(U, S, uq), e = ldh<true>(fe == 1 ? X.E1 : X.E2, S, w);
    const M3 Z = ldh<true>(X.pi, S, w) + (u * ldh<true>(X.pi, S, n)) * dag(u);
    if (fe == 1)
        sc.e1(w, c * dag(Z));
    else
        sc.e2(w, c * dag(Z));
    const M3 Zb = c * dag(e);
    sc.pi(w, Zb);
    M3 ub = m_zero(), nb = m_zero();
    adj_conj(u, ldh<true>(X.pi, S, n), Zb, ub, nb);
    if (fe == 1)
        sc.u1(uq, ub);
    else
        sc.u2(uq, ub);
    sc.pi(n, nb);
}

__global__ __launch_bounds__(IPG_BLOCK) void k_adj_tmunu_c(const EvDesc* __restrict__ evs, const double* __restrict__ xs,
                                                          int it, const double* __restrict__ tb,
                                                          double* __restrict__ pout) {
    const EvDesc& ev = evs[0];
    const Lat L{ev.N, ev.S};
    const long long s = xcd_site();
    if (s >= L.S) return;
    const size_t S = L.S;
    const int i = (int)(s / L.N), j = (int)(s - (long long)i * L.N);
    const Slot X = slot_of(xs, S);
    const TScat sc{pout, S};
    const double tau = it * ev.dtau, g = ev.g, a = ev.a;
    const double a4 = a * a * a * a, a5 = a4 * a;
    const size_t s0 = (size_t)s;
    const size_t pX = L.at(i + 1, j), pY = L.at(i, j + 1), pXY = L.at(i + 1, j + 1), mX = L.at(i - 1, j);
    const size_t mY = L.at(i, j - 1), mXpY = L.at(i - 1, j + 1), pXmY = L.at(i + 1, j - 1);
    const size_t p2X = L.at(i + 2, j), p2Y = L.at(i, j + 2), p2XY = L.at(i + 2, j + 1), pX2Y = L.at(i + 1, j + 2);
    // ttx_re, tty_re enter T^{tau x}, T^{tau y} with -2/(8 tau) a^-4
    const double cx = -tb[4 * S + s] / (4. * tau * a4), cy = -tb[5 * S + s] / (4. * tau * a4);
    pigrad_adj(sc, X, S, 1, cx, s0, s0, pX, mX, mX);
    pigrad_adj(sc, X, S, 1, cx, pY, pY, pXY, mXpY, mXpY);
    pigrad_adj(sc, X, S, 1, cx, pX, pX, p2X, s0, s0);
    pigrad_adj(sc, X, S, 1, cx, pXY, pXY, p2XY, pY, pY);
    pigrad_adj(sc, X, S, 2, cy, s0, s0, pY, mY, mY);
    pigrad_adj(sc, X, S, 2, cy, pX, pX, pXY, pXmY, pXmY);
    pigrad_adj(sc, X, S, 2, cy, pY, pY, p2Y, s0, s0);
    pigrad_adj(sc, X, S, 2, cy, pXY, pXY, pX2Y, pX, pX);
    // r1 enters T^{x eta}, T^{y eta} with -2/tau^2 g/4 a^-5
    const double rx = -tb[8 * S + s] * g / (2. * tau * tau * a5), ry = -tb[9 * S + s] * g / (2. * tau * tau * a5);
    epi_adj(sc, X, S, 1, rx, s0, s0, pX);
    epi_adj(sc, X, S, 1, rx, pY, pY, pXY);
    epi_adj(sc, X, S, 2, ry, s0, s0, pY);
    epi_adj(sc, X, S, 2, ry, pX, pX, pXY);
}

// uniform weights of the ten components (the linear loss sum_s sum_k w_k T_k(s))
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_fill_tb(long long S, const double* __restrict__ w,
                                                          double* __restrict__ tb) {
    const long long t = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (t >= 10 * S) return;
    tb[t] = w[t / S];
}

// ---- one leapfrog step in reverse ------------------------------------------------------
// xb: state before the step, xa: state after; pin: pending cotangents of the
// state after; pout: pending cotangents of the state before (every plane is
// written: pass 1 the own-field planes, pass 2 the link outboxes, pass 3 the
// phi outboxes)

// pass 1, site-local: U' = M(E') U, phi' = phi + (tau + dtau/2) dtau pi',
// and the traceless part taken by E'.  Leaves in pout: the U cotangents from
// the link update, Z_i = TL(E_i'b) (the own E_i cotangent), phi'b and pi'b.
__device__ __forceinline__ void adj_local_site(const EvDesc& ev, long long s, const double* __restrict__ xb,
                                               const double* __restrict__ xa, int it, const double* __restrict__ pin,
                                               double* __restrict__ pout) {
    const size_t S = ev.S;
    const int N = ev.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const double dt = ev.dtau, tau = it * dt, g = ev.g;
    const Slot B = slot_of(xb, S), A = slot_of(xa, S);
    const double kappa = g * g * dt / (tau + dt / 2.);
    {
        M3 e1b = ldh<true>(pin + kPE1 * S, S, s), ub = m_zero();
        u_adj(ldm<true>(B.Ux, S, s), ldh<true>(A.E1, S, s), kappa, cin_ux(pin, S, N, i, j, s), ub, e1b);
        stm<true>(pout + kPU1 * S, S, s, ub);
        sth<true>(pout + kPE1 * S, S, s, traceless(hproj(e1b)));
    }
    {
        M3 e2b = ldh<true>(pin + kPE2 * S, S, s), ub = m_zero();
        u_adj(ldm<true>(B.Uy, S, s), ldh<true>(A.E2, S, s), kappa, cin_uy(pin, S, N, i, j, s), ub, e2b);
        stm<true>(pout + kPU2 * S, S, s, ub);
        sth<true>(pout + kPE2 * S, S, s, traceless(hproj(e2b)));
    }
    const M3 phb = cin_ph(pin, S, N, i, j, s);
    sth<true>(pout + kPPh * S, S, s, phb);
    sth<true>(pout + kPPi * S, S, s, ldh<true>(pin + kPPi * S, S, s) + ((tau + dt / 2.) * dt) * phb);
}

// pass 2: the plaquettes of E' (U12 = ux uy_px ux_py^+ uy^+,
// U1m2 = ux uy_pxmy^+ ux_my^+ uy_my, U2m1 = uy ux_mxpy^+ uy_mx^+ ux_mx)
__device__ __forceinline__ void adj_plaq_site(const EvDesc& ev, long long s, const double* __restrict__ xb, int it,
                                              double* __restrict__ pout) {
    const Lat L{ev.N, ev.S};
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const double dt = ev.dtau, tau = it * dt, g = ev.g;
    const Slot B = slot_of(xb, S);
    M3 U12b, U1m2b, U2m1b;
    plaq_bars(ldh<true>(pout + kPE1 * S, S, s), ldh<true>(pout + kPE2 * S, S, s), tau, dt, g, U12b, U1m2b, U2m1b);
    M3 B1, B2, B3, B4, uxs, uys;
    {
        const M3 ux = ldm<true>(B.Ux, S, s);
        adj_prod4(ux, ldm<true>(B.Uy, S, L.at(i + 1, j)), dag(ldm<true>(B.Ux, S, L.at(i, j + 1))), dag(ldm<true>(B.Uy, S, s)), U12b,
                  B1, B2, B3, B4);
        uxs = B1;
        stm<true>(pout + (kPOy + 0) * S, S, s, B2);
        stm<true>(pout + (kPOx + 0) * S, S, s, dag(B3));
        uys = dag(B4);
        adj_prod4(ux, dag(ldm<true>(B.Uy, S, L.at(i + 1, j - 1))), dag(ldm<true>(B.Ux, S, L.at(i, j - 1))),
                  ldm<true>(B.Uy, S, L.at(i, j - 1)), U1m2b, B1, B2, B3, B4);
    }
    uxs = uxs + B1;
    stm<true>(pout + (kPOy + 18) * S, S, s, dag(B2));
    stm<true>(pout + (kPOx + 18) * S, S, s, dag(B3));
    stm<true>(pout + (kPOy + 36) * S, S, s, B4);
    adj_prod4(ldm<true>(B.Uy, S, s), dag(ldm<true>(B.Ux, S, L.at(i - 1, j + 1))), dag(ldm<true>(B.Uy, S, L.at(i - 1, j))),
              ldm<true>(B.Ux, S, L.at(i - 1, j)), U2m1b, B1, B2, B3, B4);
    uys = uys + B1;
    stm<true>(pout + (kPOx + 36) * S, S, s, dag(B2));
    stm<true>(pout + (kPOy + 54) * S, S, s, dag(B3));
    stm<true>(pout + (kPOx + 54) * S, S, s, B4);
    add_m3(pout + kPU1 * S, S, s, uxs);
    add_m3(pout + kPU2 * S, S, s, uys);
}

// pass 3: the commutators of E' and the pi' stencil
//   pi' = pi + (dtau/tau) [Ux phi(x+x) Ux^+ + Ux(x-x)^+ phi(x-x) Ux(x-x) + (y) - 4 phi]
__device__ __forceinline__ void adj_phi_site(const EvDesc& ev, long long s, const double* __restrict__ xb, int it,
                                             double* __restrict__ pout) {
    const Lat L{ev.N, ev.S};
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const double dt = ev.dtau, tau = it * dt;
    const Slot B = slot_of(xb, S);
    const cd c2c{0., -dt / tau};  // conj(i dtau / tau)
    const M3 Bb = (dt / tau) * ldh<true>(pout + kPPi * S, S, s);
    const M3 ph = ldh<true>(B.phi, S, s);
    M3 phs = -4. * Bb;
    {
        const M3 u = ldm<true>(B.Ux, S, s), pn = ldh<true>(B.phi, S, L.at(i + 1, j));
        M3 ub = m_zero(), pnb = m_zero();
        comm_adj(u, pn, ph, cscale(c2c, ldh<true>(pout + kPE1 * S, S, s)), phs, ub, pnb);
        adj_conj(u, pn, Bb, ub, pnb);
        add_m3(pout + kPU1 * S, S, s, ub);
        sth<true>(pout + (kPOph + 0) * S, S, s, hproj(pnb));
    }
    {
        const M3 u = ldm<true>(B.Uy, S, s), pn = ldh<true>(B.phi, S, L.at(i, j + 1));
        M3 ub = m_zero(), pnb = m_zero();
        comm_adj(u, pn, ph, cscale(c2c, ldh<true>(pout + kPE2 * S, S, s)), phs, ub, pnb);
        adj_conj(u, pn, Bb, ub, pnb);
        add_m3(pout + kPU2 * S, S, s, ub);
        sth<true>(pout + (kPOph + 9) * S, S, s, hproj(pnb));
    }
    {
        const size_t sm = L.at(i - 1, j);
        M3 ub = m_zero(), pnb = m_zero();
        adj_conjd(ldm<true>(B.Ux, S, sm), ldh<true>(B.phi, S, sm), Bb, ub, pnb);
        add_m3(pout + (kPOx + 54) * S, S, s, ub);
        sth<true>(pout + (kPOph + 18) * S, S, s, hproj(pnb));
    }
    {
        const size_t sm = L.at(i, j - 1);
        M3 ub = m_zero(), pnb = m_zero();
        adj_conjd(ldm<true>(B.Uy, S, sm), ldh<true>(B.phi, S, sm), Bb, ub, pnb);
        add_m3(pout + (kPOy + 36) * S, S, s, ub);
        sth<true>(pout + (kPOph + 27) * S, S, s, hproj(pnb));
    }
    add_h(pout + kPPh * S, S, s, hproj(phs));
}

// The three passes as kernels of their own (few matrices live at a time).
// Each pass reads only its own site's planes of pout, so a site can also run
// them back to back (below).
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_local(const EvDesc* __restrict__ evs, const double* __restrict__ xb,
                                                        const double* __restrict__ xa, int it,
                                                        const double* __restrict__ pin, double* __restrict__ pout) {
    const long long s = xcd_site();
    if (s < evs[0].S) adj_local_site(evs[0], s, xb, xa, it, pin, pout);
}
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_plaq(const EvDesc* __restrict__ evs, const double* __restrict__ xb,
                                                       int it, double* __restrict__ pout) {
    const long long s = xcd_site();
    if (s < evs[0].S) adj_plaq_site(evs[0], s, xb, it, pout);
}
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_phi(const EvDesc* __restrict__ evs, const double* __restrict__ xb,
                                                      int it, double* __restrict__ pout) {
    const long long s = xcd_site();
    if (s < evs[0].S) adj_phi_site(evs[0], s, xb, it, pout);
}


// The passes back to back with their partial sums on chip: the link
// cotangents of the own site (pass 1 starts them, passes 2 and 3 add) and the
// two link outboxes that passes 2 and 3 both write (U_x to (-1,0), U_y to
// (0,-1)) live in LDS, 72 doubles per thread, until their last term; every
// other plane of pout is written once, when final.  Kept in pout instead,
// each of those planes crosses HBM three times (written, read back after the
// L2 has evicted it, written again).  Same sums in the same order as the
// three kernels.
constexpr int kAccSlots = 72;  // U_x, U_y, outbox U_y (0,-1), outbox U_x (-1,0)
__device__ __forceinline__ M3 acc_ld(const double* a) {
    M3 m;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        m.e[k].r = a[(2 * k) * IPG_BLOCK];
        m.e[k].i = a[(2 * k + 1) * IPG_BLOCK];
    }
    return m;
}
__device__ __forceinline__ void acc_st(double* a, const M3& m) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
        a[(2 * k) * IPG_BLOCK] = m.e[k].r;
        a[(2 * k + 1) * IPG_BLOCK] = m.e[k].i;
    }
}
// passes 2 and 3 of one site with the partial sums in LDS; the link
// cotangents of pass 1 come from LDS (P1_LDS) or from pout
template <bool P1_LDS>
__device__ __forceinline__ void adj_plaq_phi_lds(const EvDesc& ev, long long s, const double* __restrict__ xb, int it,
                                                 double* __restrict__ pout, double* aU1) {
    double* const aU2 = aU1 + 18 * IPG_BLOCK;
    double* const aOy36 = aU1 + 36 * IPG_BLOCK;
    double* const aOx54 = aU1 + 54 * IPG_BLOCK;
    const Lat L{ev.N, ev.S};
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const double dt = ev.dtau, tau = it * dt, g = ev.g;
    const Slot B = slot_of(xb, S);
    // pass 2: the plaquettes of E'
    {
        M3 U12b, U1m2b, U2m1b;
        plaq_bars(ldh<true>(pout + kPE1 * S, S, s), ldh<true>(pout + kPE2 * S, S, s), tau, dt, g, U12b, U1m2b, U2m1b);
        M3 B1, B2, B3, B4, uxs, uys;
        {
            const M3 ux = ldm<true>(B.Ux, S, s);
            adj_prod4(ux, ldm<true>(B.Uy, S, L.at(i + 1, j)), dag(ldm<true>(B.Ux, S, L.at(i, j + 1))), dag(ldm<true>(B.Uy, S, s)),
                      U12b, B1, B2, B3, B4);
            uxs = B1;
            stm<true>(pout + (kPOy + 0) * S, S, s, B2);
            stm<true>(pout + (kPOx + 0) * S, S, s, dag(B3));
            uys = dag(B4);
            adj_prod4(ux, dag(ldm<true>(B.Uy, S, L.at(i + 1, j - 1))), dag(ldm<true>(B.Ux, S, L.at(i, j - 1))),
                      ldm<true>(B.Uy, S, L.at(i, j - 1)), U1m2b, B1, B2, B3, B4);
        }
        uxs = uxs + B1;
        stm<true>(pout + (kPOy + 18) * S, S, s, dag(B2));
        stm<true>(pout + (kPOx + 18) * S, S, s, dag(B3));
        acc_st(aOy36, B4);
        adj_prod4(ldm<true>(B.Uy, S, s), dag(ldm<true>(B.Ux, S, L.at(i - 1, j + 1))), dag(ldm<true>(B.Uy, S, L.at(i - 1, j))),
                  ldm<true>(B.Ux, S, L.at(i - 1, j)), U2m1b, B1, B2, B3, B4);
        uys = uys + B1;
        stm<true>(pout + (kPOx + 36) * S, S, s, dag(B2));
        stm<true>(pout + (kPOy + 54) * S, S, s, dag(B3));
        acc_st(aOx54, B4);
        acc_st(aU1, (P1_LDS ? acc_ld(aU1) : ldm<true>(pout + kPU1 * S, S, s)) + uxs);
        acc_st(aU2, (P1_LDS ? acc_ld(aU2) : ldm<true>(pout + kPU2 * S, S, s)) + uys);
    }
    asm volatile("" ::: "memory");
    // pass 3: the commutators of E' and the pi' stencil
    {
        const cd c2c{0., -dt / tau};  // conj(i dtau / tau)
        const M3 Bb = (dt / tau) * ldh<true>(pout + kPPi * S, S, s);
        const M3 ph = ldh<true>(B.phi, S, s);
        M3 phs = -4. * Bb;
        {
            const M3 u = ldm<true>(B.Ux, S, s), pn = ldh<true>(B.phi, S, L.at(i + 1, j));
            M3 ub = m_zero(), pnb = m_zero();
            comm_adj(u, pn, ph, cscale(c2c, ldh<true>(pout + kPE1 * S, S, s)), phs, ub, pnb);
            adj_conj(u, pn, Bb, ub, pnb);
            stm<true>(pout + kPU1 * S, S, s, acc_ld(aU1) + ub);
            sth<true>(pout + (kPOph + 0) * S, S, s, hproj(pnb));
        }
        {
            const M3 u = ldm<true>(B.Uy, S, s), pn = ldh<true>(B.phi, S, L.at(i, j + 1));
            M3 ub = m_zero(), pnb = m_zero();
            comm_adj(u, pn, ph, cscale(c2c, ldh<true>(pout + kPE2 * S, S, s)), phs, ub, pnb);
            adj_conj(u, pn, Bb, ub, pnb);
            stm<true>(pout + kPU2 * S, S, s, acc_ld(aU2) + ub);
            sth<true>(pout + (kPOph + 9) * S, S, s, hproj(pnb));
        }
        {
            const size_t sm = L.at(i - 1, j);
            M3 ub = m_zero(), pnb = m_zero();
            adj_conjd(ldm<true>(B.Ux, S, sm), ldh<true>(B.phi, S, sm), Bb, ub, pnb);
            stm<true>(pout + (kPOx + 54) * S, S, s, acc_ld(aOx54) + ub);
            sth<true>(pout + (kPOph + 18) * S, S, s, hproj(pnb));
        }
        {
            const size_t sm = L.at(i, j - 1);
            M3 ub = m_zero(), pnb = m_zero();
            adj_conjd(ldm<true>(B.Uy, S, sm), ldh<true>(B.phi, S, sm), Bb, ub, pnb);
            stm<true>(pout + (kPOy + 36) * S, S, s, acc_ld(aOy36) + ub);
            sth<true>(pout + (kPOph + 27) * S, S, s, hproj(pnb));
        }
        add_h(pout + kPPh * S, S, s, hproj(phs));
    }
}
__global__ __launch_bounds__(IPG_BLOCK, 1) void k_adj_step_lds(const EvDesc* __restrict__ evs,
                                                              const double* __restrict__ xb,
                                                              const double* __restrict__ xa, int it,
                                                              const double* __restrict__ pin,
                                                              double* __restrict__ pout) {
    __shared__ double acc[kAccSlots * IPG_BLOCK];
    double* const aU1 = acc + threadIdx.x;
    double* const aU2 = aU1 + 18 * IPG_BLOCK;
    const EvDesc& ev = evs[0];
    const Lat L{ev.N, ev.S};
    const long long s = xcd_site();
    if (s >= L.S) return;
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const double dt = ev.dtau, tau = it * dt, g = ev.g;
    const Slot B = slot_of(xb, S), A = slot_of(xa, S);
    // pass 1: the link update, the traceless part of E', phi' and pi'
    {
        const double kappa = g * g * dt / (tau + dt / 2.);
        {
            M3 e1b = ldh<true>(pin + kPE1 * S, S, s), ub = m_zero();
            u_adj(ldm<true>(B.Ux, S, s), ldh<true>(A.E1, S, s), kappa, cin_ux(pin, S, N, i, j, s), ub, e1b);
            acc_st(aU1, ub);
            sth<true>(pout + kPE1 * S, S, s, traceless(hproj(e1b)));
        }
        {
            M3 e2b = ldh<true>(pin + kPE2 * S, S, s), ub = m_zero();
            u_adj(ldm<true>(B.Uy, S, s), ldh<true>(A.E2, S, s), kappa, cin_uy(pin, S, N, i, j, s), ub, e2b);
            acc_st(aU2, ub);
            sth<true>(pout + kPE2 * S, S, s, traceless(hproj(e2b)));
        }
        const M3 phb = cin_ph(pin, S, N, i, j, s);
        sth<true>(pout + kPPh * S, S, s, phb);
        sth<true>(pout + kPPi * S, S, s, ldh<true>(pin + kPPi * S, S, s) + ((tau + dt / 2.) * dt) * phb);
    }
    // the partial sums are read back from LDS, not kept in registers
    asm volatile("" ::: "memory");
    adj_plaq_phi_lds<true>(ev, s, xb, it, pout, aU1);
}
// passes 2 and 3 as one kernel after k_adj_local (the whole-step kernel
// above spills: pass 1's link exponential reverse and pass 3 do not fit the
// register file together)
__global__ __launch_bounds__(IPG_BLOCK, 1) void k_adj_plaq_phi(const EvDesc* __restrict__ evs,
                                                              const double* __restrict__ xb, int it,
                                                              double* __restrict__ pout) {
    __shared__ double acc[kAccSlots * IPG_BLOCK];
    const long long s = xcd_site();
    if (s < evs[0].S) adj_plaq_phi_lds<false>(evs[0], s, xb, it, pout, acc + threadIdx.x);
}
// ---- the tau = 0 kick (phi += dtau^2/2 pi, U = M(E) U) in reverse ----------------
// pin: pending cotangents after the kick; cout: compact cotangents of the
// tau = 0 fields
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_kick(const EvDesc* __restrict__ evs, const double* __restrict__ x0,
                                                       const double* __restrict__ pin, double* __restrict__ cout) {
    const EvDesc& ev = evs[0];
    const long long s = xcd_site();
    if (s >= ev.S) return;
    const size_t S = ev.S;
    const int N = ev.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const double dt = ev.dtau, tau = 0., g = ev.g;
    const Slot X = slot_of(x0, S);
    const double kappa = g * g * dt / (tau + dt / 2.);
    {
        M3 e1b = ldh<true>(pin + kPE1 * S, S, s), ub = m_zero();
        u_adj(ldm<true>(X.Ux, S, s), ldh<true>(X.E1, S, s), kappa, cin_ux(pin, S, N, i, j, s), ub, e1b);
        stm<true>(cot(cout, S, 0), S, s, ub);
        stm<true>(cot(cout, S, 2), S, s, hproj(e1b));
    }
    {
        M3 e2b = ldh<true>(pin + kPE2 * S, S, s), ub = m_zero();
        u_adj(ldm<true>(X.Uy, S, s), ldh<true>(X.E2, S, s), kappa, cin_uy(pin, S, N, i, j, s), ub, e2b);
        stm<true>(cot(cout, S, 1), S, s, ub);
        stm<true>(cot(cout, S, 3), S, s, hproj(e2b));
    }
    const M3 phb = cin_ph(pin, S, N, i, j, s);
    stm<true>(cot(cout, S, 4), S, s, phb);
    stm<true>(cot(cout, S, 5), S, s, ldh<true>(pin + kPPi * S, S, s) + ((tau + dt / 2.) * dt) * phb);
}

// ---- back from the tau = 0 fields to the Wilson lines ------------------------------
// links from the saved Wilson lines (k_links, clamped neighbours)
__device__ __forceinline__ M3 link_of(const double* V, const Lat& L, int i, int j, int dir) {
    const size_t s = (size_t)i * L.N + j;
    const size_t n = dir == 0 ? L.at(i + 1, j) : L.at(i, j + 1);
    return mul_bd(ldm<true>(V, L.S, s), ldm<true>(V, L.S, n));
}

// pi = -(i/4g) acc at tau = 0 (k_pi_init) in reverse: accb = conj(-i/4g) pib;
// adds to the cotangents of the tau = 0 links U (ulb, 2 x 18 planes) and of
// D = L1 - L2 (db, 2 x 18 planes)
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_pi_init(const EvDesc* __restrict__ evs,
                                                          const double* __restrict__ s0,
                                                          const double* __restrict__ vab,
                                                          const double* __restrict__ cs0, double* __restrict__ ulb,
                                                          double* __restrict__ db) {
    const EvDesc& ev = evs[0];
    const Lat L{ev.N, ev.S};
    const long long s = xcd_site();
    if (s >= L.S) return;
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const Slot X = slot_of(s0, S);
    const double* VA = vab;
    const double* VB = vab + 18 * S;
    const double c = -1. / (4. * ev.g);
    const M3 accb = cscale(cd{0., -c}, hproj(ldm<true>(cot(const_cast<double*>(cs0), S, 5), S, s)));
    for (int dir = 0; dir < 2; dir++) {
        const int im = dir == 0 ? i - 1 : i, jm = dir == 0 ? j : j - 1;
        const size_t sm = L.at(im, jm);
        const int imc = (int)(sm / N), jmc = (int)(sm - (size_t)imc * N);
        const M3 D = link_of(VA, L, i, j, dir) - link_of(VB, L, i, j, dir);
        const M3 Dm = link_of(VA, L, imc, jmc, dir) - link_of(VB, L, imc, jmc, dir);
        const double* Uf = dir == 0 ? X.Ux : X.Uy;
        M3 Db = m_zero(), Ub = m_zero(), Dmb = m_zero(), Umb = m_zero();
        acc_dir_adj(D, ldm<true>(Uf, S, s), Dm, ldm<true>(Uf, S, sm), accb, Db, Ub, Dmb, Umb);
        double* ul = ulb + (size_t)dir * 18 * S;
        double* dd = db + (size_t)dir * 18 * S;
        atomic_add_m3(ul, S, s, Ub);
        atomic_add_m3(ul, S, sm, Umb);
        atomic_add_m3(dd, S, s, Db);
        atomic_add_m3(dd, S, sm, Dmb);
    }
}

// root finder (implicit) and links in reverse: for each link, U1b, U2b from
// the total cotangent of its solution, plus Db (L1b += Db, L2b -= Db); then
// L = V(s) V(s+d)^+ adds to the Wilson-line cotangents (vb, 2 x 18 planes).
// One launch per direction DIR (fewer live registers than both in one).
template <int DIR>
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_links_newton(const EvDesc* __restrict__ evs,
                                                               const double* __restrict__ s0,
                                                               const double* __restrict__ vab,
                                                               const double* __restrict__ ulb,
                                                               const double* __restrict__ db,
                                                               double* __restrict__ vb) {
    const EvDesc& ev = evs[0];
    const Lat L{ev.N, ev.S};
    const long long s = xcd_site();
    if (s >= L.S) return;
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const Slot X = slot_of(s0, S);
    const double* VA = vab;
    const double* VB = vab + 18 * S;
    {
        constexpr int dir = DIR;
        const size_t n = dir == 0 ? L.at(i + 1, j) : L.at(i, j + 1);
        M3 U1b = m_zero(), U2b = m_zero();
        {
            const M3 U1 = mul_bd(ldm<true>(VA, S, s), ldm<true>(VA, S, n));
            const M3 U2 = mul_bd(ldm<true>(VB, S, s), ldm<true>(VB, S, n));
            newton_adj(U1, U2, ldm<true>(dir == 0 ? X.Ux : X.Uy, S, s), ldm<true>(ulb + (size_t)dir * 18 * S, S, s),
                       U1b, U2b);
        }
        const M3 Db = ldm<true>(db + (size_t)dir * 18 * S, S, s);
        U1b = U1b + Db;
        U2b = U2b - Db;
        // L = V(s) V(n)^+:  V(s) b += Lb V(n),  V(n) b += Lb^+ V(s); the
        // Wilson lines are loaded again here (cache hits) rather than kept
        // live across the solve
        atomic_add_m3(vb, S, s, U1b * ldm<true>(VA, S, n));
        atomic_add_m3(vb, S, n, mul_ad(U1b, ldm<true>(VA, S, s)));
        atomic_add_m3(vb + 18 * S, S, s, U2b * ldm<true>(VB, S, n));
        atomic_add_m3(vb + 18 * S, S, n, mul_ad(U2b, ldm<true>(VB, S, s)));
    }
}

}  // namespace

static unsigned nb_of(long long n) { return (unsigned)((n + IPG_BLOCK - 1) / IPG_BLOCK); }
// blocks of a site-parallel launch: a multiple of 8 for xcd_site
static unsigned nb8_of(long long S) { return (nb_of(S) + 7) / 8 * 8; }

void ipg_adj_pi_init(const EvDesc* d, long long S, const double* s0, const double* vab, const double* cs0,
                     double* ulb, double* db, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_pi_init, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, s0, vab, cs0, ulb, db);
}
void ipg_adj_links_newton(const EvDesc* d, long long S, const double* s0, const double* vab, const double* ulb,
                          const double* db, double* vb, hipStream_t st) {
    const dim3 g(nb8_of(S));
    hipLaunchKernelGGL(k_adj_links_newton<0>, g, dim3(IPG_BLOCK), 0, st, d, s0, vab, ulb, db, vb);
    hipLaunchKernelGGL(k_adj_links_newton<1>, g, dim3(IPG_BLOCK), 0, st, d, s0, vab, ulb, db, vb);
}


void ipg_adj_state_save(const EvDesc* d, long long S, int parity, double* slot, hipStream_t st) {
    hipLaunchKernelGGL(k_state_save, dim3(nb_of(72 * S)), dim3(IPG_BLOCK), 0, st, d, parity, slot);
}
void ipg_adj_state_load(const EvDesc* d, long long S, int parity, const double* slot, hipStream_t st) {
    hipLaunchKernelGGL(k_state_load, dim3(nb_of(72 * S)), dim3(IPG_BLOCK), 0, st, d, parity, slot);
}
void ipg_adj_zero(double* p, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3(nb_of(n)), dim3(IPG_BLOCK), 0, st, p, n);
}
void ipg_adj_loss(const EvDesc* d, long long S, const double* xs, int it, double yb, double* cout, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_loss, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, xs, it, yb, cout);
}
void ipg_adj_step(const EvDesc* d, long long S, const double* xb, const double* xa, int it, const double* cin,
                  double* cout, hipStream_t st, int variant) {
    if (variant == 1) {
        hipLaunchKernelGGL(k_adj_step_lds, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, xb, xa, it, cin, cout);
        return;
    }
    hipLaunchKernelGGL(k_adj_local, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, xb, xa, it, cin, cout);
    if (variant == 2) {
        hipLaunchKernelGGL(k_adj_plaq_phi, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, xb, it, cout);
        return;
    }
    hipLaunchKernelGGL(k_adj_plaq, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, xb, it, cout);
    hipLaunchKernelGGL(k_adj_phi, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, xb, it, cout);
}
void ipg_adj_kick(const EvDesc* d, long long S, const double* x0, const double* cin, double* cout, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_kick, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, x0, cin, cout);
}

void ipg_adj_tmunu(const EvDesc* d, long long S, const double* xs, int it, const double* tb, double* pout,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_adj_tmunu_a, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, xs, it, tb, pout);
    hipLaunchKernelGGL(k_adj_tmunu_b, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, xs, it, tb, pout);
    hipLaunchKernelGGL(k_adj_tmunu_c, dim3(nb8_of(S)), dim3(IPG_BLOCK), 0, st, d, xs, it, tb, pout);
}
void ipg_adj_fill_tb(long long S, const double* w, double* tb, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_fill_tb, dim3(nb_of(10 * S)), dim3(IPG_BLOCK), 0, st, S, w, tb);
}
