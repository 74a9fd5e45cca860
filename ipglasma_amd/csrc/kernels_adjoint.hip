// kernels_adjoint.hip -- reverse mode of dE/deta through the leapfrog.
//
// dE/deta = tau a^2 hbarc sum_s T^tautau(s) at it = itmax is a function of the
// state after step itmax-1, which the leapfrog (Evolution.cpp:366-431) builds
// from the tau = 0 fields.  The backward pass walks the steps in reverse: for
// step it it needs the state before and after the step, which the host
// recomputes segment by segment from checkpoints (every K = ceil(sqrt(T))
// steps).  Every site evaluates the per-site reverse functions of adjoint.hpp
// and adds the cotangents of the neighbour fields it read with FP64 atomics
// (gathers would need each site's stencil cotangents twice).
//
// A state slot is 72 planes of S doubles: U_x, U_y (18 each, full complex),
// E1, E2, phi, pi (9 each, Hermitian storage).  A cotangent set is 108
// planes: U_x, U_y, E1, E2, phi, pi, all full complex (18 each).
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

struct Lat {
    int N;
    long long S;
    __device__ __forceinline__ size_t at(int i, int j) const {
        return (size_t)clampi(i, N) * N + clampi(j, N);
    }
};

// ---- state copies between the live buffers and a slot -----------------------------
__global__ __launch_bounds__(IPG_BLOCK) void k_state_save(const EvDesc* __restrict__ evs, int parity,
                                                         double* __restrict__ slot) {
    const EvDesc& ev = evs[0];
    const long long S = ev.S;
    const long long t = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (t >= S) return;
    const double* src[6] = {ev.U[parity][0], ev.U[parity][1], ev.E1, ev.E2, ev.phi[parity], ev.pi};
    const int np[6] = {18, 18, 9, 9, 9, 9};
    int o = 0;
    for (int f = 0; f < 6; f++) {
        for (int k = 0; k < np[f]; k++) slot[(size_t)(o + k) * S + t] = src[f][(size_t)k * S + t];
        o += np[f];
    }
}
__global__ __launch_bounds__(IPG_BLOCK) void k_state_load(const EvDesc* __restrict__ evs, int parity,
                                                         const double* __restrict__ slot) {
    const EvDesc& ev = evs[0];
    const long long S = ev.S;
    const long long t = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (t >= S) return;
    double* dst[6] = {ev.U[parity][0], ev.U[parity][1], ev.E1, ev.E2, ev.phi[parity], ev.pi};
    const int np[6] = {18, 18, 9, 9, 9, 9};
    int o = 0;
    for (int f = 0; f < 6; f++) {
        for (int k = 0; k < np[f]; k++) dst[f][(size_t)k * S + t] = slot[(size_t)(o + k) * S + t];
        o += np[f];
    }
}

__global__ __launch_bounds__(IPG_BLOCK) void k_zero(double* __restrict__ p, long long n) {
    const long long t = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (t < n) p[t] = 0.;
}

// ---- loss: dL/dT^tautau(s) = tau_fm hbarc / a^2 for every site --------------------
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_loss(const EvDesc* __restrict__ evs, const double* __restrict__ xs,
                                                       int it, double yb, double* __restrict__ cout) {
    const EvDesc& ev = evs[0];
    const Lat L{ev.N, ev.S};
    const long long s = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (s >= L.S) return;
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const Slot X = slot_of(xs, S);
    const size_t pX = L.at(i + 1, j), pY = L.at(i, j + 1), pXY = L.at(i + 1, j + 1);
    TIn q;
    q.e1 = ldh(X.E1, S, s);
    q.e1py = ldh(X.E1, S, pY);
    q.e2 = ldh(X.E2, S, s);
    q.e2px = ldh(X.E2, S, pX);
    q.p = ldh(X.pi, S, s);
    q.ppx = ldh(X.pi, S, pX);
    q.ppy = ldh(X.pi, S, pY);
    q.ppxy = ldh(X.pi, S, pXY);
    q.ux = ldm(X.Ux, S, s);
    q.uy = ldm(X.Uy, S, s);
    q.uy_px = ldm(X.Uy, S, pX);
    q.ux_py = ldm(X.Ux, S, pY);
    q.ph = ldh(X.phi, S, s);
    q.phpx = ldh(X.phi, S, pX);
    q.phpy = ldh(X.phi, S, pY);
    q.phpxy = ldh(X.phi, S, pXY);
    q.plaq = i != N - 1;
    TBar b;
    M3* bb = reinterpret_cast<M3*>(&b);
    for (int k = 0; k < 16; k++) bb[k] = m_zero();
    tt_adj(q, it * ev.dtau, ev.g, yb, b);
    double* cU1 = cot(cout, S, 0);
    double* cU2 = cot(cout, S, 1);
    double* cE1 = cot(cout, S, 2);
    double* cE2 = cot(cout, S, 3);
    double* cPh = cot(cout, S, 4);
    double* cPi = cot(cout, S, 5);
    atomic_add_m3(cE1, S, s, b.e1);
    atomic_add_m3(cE1, S, pY, b.e1py);
    atomic_add_m3(cE2, S, s, b.e2);
    atomic_add_m3(cE2, S, pX, b.e2px);
    atomic_add_m3(cPi, S, s, b.p);
    atomic_add_m3(cPi, S, pX, b.ppx);
    atomic_add_m3(cPi, S, pY, b.ppy);
    atomic_add_m3(cPi, S, pXY, b.ppxy);
    atomic_add_m3(cU1, S, s, b.ux);
    atomic_add_m3(cU2, S, s, b.uy);
    atomic_add_m3(cU2, S, pX, b.uy_px);
    atomic_add_m3(cU1, S, pY, b.ux_py);
    atomic_add_m3(cPh, S, s, b.ph);
    atomic_add_m3(cPh, S, pX, b.phpx);
    atomic_add_m3(cPh, S, pY, b.phpy);
    atomic_add_m3(cPh, S, pXY, b.phpxy);
}

// ---- one leapfrog step in reverse ------------------------------------------------------
// xb: state before the step, xa: state after; cin: cotangent of the state after
// (E, phi, pi taken on their Hermitian part), cout: cotangent of the state
// before (zeroed by the caller)
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_step(const EvDesc* __restrict__ evs, const double* __restrict__ xb,
                                                       const double* __restrict__ xa, int it,
                                                       const double* __restrict__ cin, double* __restrict__ cout) {
    const EvDesc& ev = evs[0];
    const Lat L{ev.N, ev.S};
    const long long s = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (s >= L.S) return;
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const double dt = ev.dtau, tau = it * dt, g = ev.g;
    const Slot B = slot_of(xb, S), A = slot_of(xa, S);
    double* cU1 = cot(cout, S, 0);
    double* cU2 = cot(cout, S, 1);
    double* cE1 = cot(cout, S, 2);
    double* cE2 = cot(cout, S, 3);
    double* cPh = cot(cout, S, 4);
    double* cPi = cot(cout, S, 5);
    const double* iU1 = cot(const_cast<double*>(cin), S, 0);
    const double* iU2 = cot(const_cast<double*>(cin), S, 1);
    // local: U' = M(E') U and phi' = phi + c pi'
    const double kappa = g * g * dt / (tau + dt / 2.);
    M3 e1b = hproj(ldm(cot(const_cast<double*>(cin), S, 2), S, s));
    M3 e2b = hproj(ldm(cot(const_cast<double*>(cin), S, 3), S, s));
    const M3 phb = hproj(ldm(cot(const_cast<double*>(cin), S, 4), S, s));
    M3 pib = hproj(ldm(cot(const_cast<double*>(cin), S, 5), S, s));
    {
        M3 ub = m_zero();
        u_adj(ldm(B.Ux, S, s), ldh(A.E1, S, s), kappa, ldm(iU1, S, s), ub, e1b);
        atomic_add_m3(cU1, S, s, ub);
        ub = m_zero();
        u_adj(ldm(B.Uy, S, s), ldh(A.E2, S, s), kappa, ldm(iU2, S, s), ub, e2b);
        atomic_add_m3(cU2, S, s, ub);
        // E' is stored Hermitian
        e1b = hproj(e1b);
        e2b = hproj(e2b);
    }
    atomic_add_m3(cPh, S, s, phb);
    pib = pib + ((tau + dt / 2.) * dt) * phb;
    // E' stencil
    {
        EIn q;
        q.ux = ldm(B.Ux, S, s);
        q.uy = ldm(B.Uy, S, s);
        q.uy_px = ldm(B.Uy, S, L.at(i + 1, j));
        q.ux_py = ldm(B.Ux, S, L.at(i, j + 1));
        q.uy_pxmy = ldm(B.Uy, S, L.at(i + 1, j - 1));
        q.ux_my = ldm(B.Ux, S, L.at(i, j - 1));
        q.uy_my = ldm(B.Uy, S, L.at(i, j - 1));
        q.ux_mxpy = ldm(B.Ux, S, L.at(i - 1, j + 1));
        q.uy_mx = ldm(B.Uy, S, L.at(i - 1, j));
        q.ux_mx = ldm(B.Ux, S, L.at(i - 1, j));
        q.ph = ldh(B.phi, S, s);
        q.phpx = ldh(B.phi, S, L.at(i + 1, j));
        q.phpy = ldh(B.phi, S, L.at(i, j + 1));
        q.e1 = ldh(B.E1, S, s);
        q.e2 = ldh(B.E2, S, s);
        EBar b;
        M3* bb = reinterpret_cast<M3*>(&b);
        for (int k = 0; k < 15; k++) bb[k] = m_zero();
        e_adj(q, tau, dt, g, e1b, e2b, b);
        add_m3(cE1, S, s, b.e1);
        add_m3(cE2, S, s, b.e2);
        atomic_add_m3(cU1, S, s, b.ux);
        atomic_add_m3(cU2, S, s, b.uy);
        atomic_add_m3(cU2, S, L.at(i + 1, j), b.uy_px);
        atomic_add_m3(cU1, S, L.at(i, j + 1), b.ux_py);
        atomic_add_m3(cU2, S, L.at(i + 1, j - 1), b.uy_pxmy);
        atomic_add_m3(cU1, S, L.at(i, j - 1), b.ux_my);
        atomic_add_m3(cU2, S, L.at(i, j - 1), b.uy_my);
        atomic_add_m3(cU1, S, L.at(i - 1, j + 1), b.ux_mxpy);
        atomic_add_m3(cU2, S, L.at(i - 1, j), b.uy_mx);
        atomic_add_m3(cU1, S, L.at(i - 1, j), b.ux_mx);
        atomic_add_m3(cPh, S, s, b.ph);
        atomic_add_m3(cPh, S, L.at(i + 1, j), b.phpx);
        atomic_add_m3(cPh, S, L.at(i, j + 1), b.phpy);
    }
    // pi' stencil
    {
        PiIn q;
        q.ux = ldm(B.Ux, S, s);
        q.uy = ldm(B.Uy, S, s);
        q.uxm = ldm(B.Ux, S, L.at(i - 1, j));
        q.uym = ldm(B.Uy, S, L.at(i, j - 1));
        q.ph = ldh(B.phi, S, s);
        q.phpx = ldh(B.phi, S, L.at(i + 1, j));
        q.phmx = ldh(B.phi, S, L.at(i - 1, j));
        q.phpy = ldh(B.phi, S, L.at(i, j + 1));
        q.phmy = ldh(B.phi, S, L.at(i, j - 1));
        q.pi = ldh(B.pi, S, s);
        PiBar b;
        M3* bb = reinterpret_cast<M3*>(&b);
        for (int k = 0; k < 10; k++) bb[k] = m_zero();
        pi_adj(q, dt / tau, pib, b);
        add_m3(cPi, S, s, b.pi);
        atomic_add_m3(cU1, S, s, b.ux);
        atomic_add_m3(cU2, S, s, b.uy);
        atomic_add_m3(cU1, S, L.at(i - 1, j), b.uxm);
        atomic_add_m3(cU2, S, L.at(i, j - 1), b.uym);
        atomic_add_m3(cPh, S, s, b.ph);
        atomic_add_m3(cPh, S, L.at(i + 1, j), b.phpx);
        atomic_add_m3(cPh, S, L.at(i - 1, j), b.phmx);
        atomic_add_m3(cPh, S, L.at(i, j + 1), b.phpy);
        atomic_add_m3(cPh, S, L.at(i, j - 1), b.phmy);
    }
}

// ---- the tau = 0 kick (phi += dtau^2/2 pi, U = M(E) U) in reverse ----------------
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_kick(const EvDesc* __restrict__ evs, const double* __restrict__ x0,
                                                       const double* __restrict__ cin, double* __restrict__ cout) {
    const EvDesc& ev = evs[0];
    const long long s = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (s >= ev.S) return;
    const size_t S = ev.S;
    const double dt = ev.dtau, tau = 0., g = ev.g;
    const Slot X = slot_of(x0, S);
    const double kappa = g * g * dt / (tau + dt / 2.);
    M3 e1b = hproj(ldm(cot(const_cast<double*>(cin), S, 2), S, s));
    M3 e2b = hproj(ldm(cot(const_cast<double*>(cin), S, 3), S, s));
    const M3 phb = hproj(ldm(cot(const_cast<double*>(cin), S, 4), S, s));
    const M3 pib = hproj(ldm(cot(const_cast<double*>(cin), S, 5), S, s));
    M3 ub = m_zero();
    u_adj(ldm(X.Ux, S, s), ldh(X.E1, S, s), kappa, ldm(cot(const_cast<double*>(cin), S, 0), S, s), ub, e1b);
    stm(cot(cout, S, 0), S, s, ub);
    ub = m_zero();
    u_adj(ldm(X.Uy, S, s), ldh(X.E2, S, s), kappa, ldm(cot(const_cast<double*>(cin), S, 1), S, s), ub, e2b);
    stm(cot(cout, S, 1), S, s, ub);
    stm(cot(cout, S, 2), S, s, hproj(e1b));
    stm(cot(cout, S, 3), S, s, hproj(e2b));
    stm(cot(cout, S, 4), S, s, phb);
    stm(cot(cout, S, 5), S, s, hproj(pib + ((tau + dt / 2.) * dt) * phb));
}

// ---- back from the tau = 0 fields to the Wilson lines ------------------------------
// links from the saved Wilson lines (k_links, clamped neighbours)
__device__ __forceinline__ M3 link_of(const double* V, const Lat& L, int i, int j, int dir) {
    const size_t s = (size_t)i * L.N + j;
    const size_t n = dir == 0 ? L.at(i + 1, j) : L.at(i, j + 1);
    return mul_bd(ldm(V, L.S, s), ldm(V, L.S, n));
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
    const long long s = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (s >= L.S) return;
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const Slot X = slot_of(s0, S);
    const double* VA = vab;
    const double* VB = vab + 18 * S;
    const double c = -1. / (4. * ev.g);
    const M3 accb = cscale(cd{0., -c}, hproj(ldm(cot(const_cast<double*>(cs0), S, 5), S, s)));
    for (int dir = 0; dir < 2; dir++) {
        const int im = dir == 0 ? i - 1 : i, jm = dir == 0 ? j : j - 1;
        const size_t sm = L.at(im, jm);
        const int imc = (int)(sm / N), jmc = (int)(sm - (size_t)imc * N);
        const M3 D = link_of(VA, L, i, j, dir) - link_of(VB, L, i, j, dir);
        const M3 Dm = link_of(VA, L, imc, jmc, dir) - link_of(VB, L, imc, jmc, dir);
        const double* Uf = dir == 0 ? X.Ux : X.Uy;
        M3 Db = m_zero(), Ub = m_zero(), Dmb = m_zero(), Umb = m_zero();
        acc_dir_adj(D, ldm(Uf, S, s), Dm, ldm(Uf, S, sm), accb, Db, Ub, Dmb, Umb);
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
// L = V(s) V(s+d)^+ adds to the Wilson-line cotangents (vb, 2 x 18 planes)
__global__ __launch_bounds__(IPG_BLOCK) void k_adj_links_newton(const EvDesc* __restrict__ evs,
                                                               const double* __restrict__ s0,
                                                               const double* __restrict__ vab,
                                                               const double* __restrict__ ulb,
                                                               const double* __restrict__ db,
                                                               double* __restrict__ vb) {
    const EvDesc& ev = evs[0];
    const Lat L{ev.N, ev.S};
    const long long s = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    if (s >= L.S) return;
    const size_t S = L.S;
    const int N = L.N, i = (int)(s / N), j = (int)(s - (long long)i * N);
    const Slot X = slot_of(s0, S);
    const double* VA = vab;
    const double* VB = vab + 18 * S;
    for (int dir = 0; dir < 2; dir++) {
        const size_t n = dir == 0 ? L.at(i + 1, j) : L.at(i, j + 1);
        const M3 va = ldm(VA, S, s), van = ldm(VA, S, n), vbs = ldm(VB, S, s), vbn = ldm(VB, S, n);
        const M3 U1 = mul_bd(va, van), U2 = mul_bd(vbs, vbn);
        const M3 U = ldm(dir == 0 ? X.Ux : X.Uy, S, s);
        M3 U1b = m_zero(), U2b = m_zero();
        newton_adj(U1, U2, U, ldm(ulb + (size_t)dir * 18 * S, S, s), U1b, U2b);
        const M3 Db = ldm(db + (size_t)dir * 18 * S, S, s);
        U1b = U1b + Db;
        U2b = U2b - Db;
        // L = V(s) V(n)^+:  V(s) b += Lb V(n),  V(n) b += Lb^+ V(s)
        atomic_add_m3(vb, S, s, U1b * van);
        atomic_add_m3(vb, S, n, mul_ad(U1b, va));
        atomic_add_m3(vb + 18 * S, S, s, U2b * vbn);
        atomic_add_m3(vb + 18 * S, S, n, mul_ad(U2b, vbs));
    }
}

}  // namespace

void ipg_adj_pi_init(const EvDesc* d, long long S, const double* s0, const double* vab, const double* cs0,
                     double* ulb, double* db, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_pi_init, dim3((unsigned)((S + IPG_BLOCK - 1) / IPG_BLOCK)), dim3(IPG_BLOCK), 0, st, d,
                       s0, vab, cs0, ulb, db);
}
void ipg_adj_links_newton(const EvDesc* d, long long S, const double* s0, const double* vab, const double* ulb,
                          const double* db, double* vb, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_links_newton, dim3((unsigned)((S + IPG_BLOCK - 1) / IPG_BLOCK)), dim3(IPG_BLOCK), 0,
                       st, d, s0, vab, ulb, db, vb);
}

static unsigned nb_of(long long n) { return (unsigned)((n + IPG_BLOCK - 1) / IPG_BLOCK); }

void ipg_adj_state_save(const EvDesc* d, long long S, int parity, double* slot, hipStream_t st) {
    hipLaunchKernelGGL(k_state_save, dim3(nb_of(S)), dim3(IPG_BLOCK), 0, st, d, parity, slot);
}
void ipg_adj_state_load(const EvDesc* d, long long S, int parity, const double* slot, hipStream_t st) {
    hipLaunchKernelGGL(k_state_load, dim3(nb_of(S)), dim3(IPG_BLOCK), 0, st, d, parity, slot);
}
void ipg_adj_zero(double* p, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3(nb_of(n)), dim3(IPG_BLOCK), 0, st, p, n);
}
void ipg_adj_loss(const EvDesc* d, long long S, const double* xs, int it, double yb, double* cout, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_loss, dim3(nb_of(S)), dim3(IPG_BLOCK), 0, st, d, xs, it, yb, cout);
}
void ipg_adj_step(const EvDesc* d, long long S, const double* xb, const double* xa, int it, const double* cin,
                  double* cout, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_step, dim3(nb_of(S)), dim3(IPG_BLOCK), 0, st, d, xb, xa, it, cin, cout);
}
void ipg_adj_kick(const EvDesc* d, long long S, const double* x0, const double* cin, double* cout, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_kick, dim3(nb_of(S)), dim3(IPG_BLOCK), 0, st, d, x0, cin, cout);
}
