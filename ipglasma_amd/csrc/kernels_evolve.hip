// kernels_evolve.hip -- classical Yang-Mills leapfrog, T^{mu nu}, Gauss law.
//
// One fused site kernel per leapfrog step replaces the four OpenMP sweeps of
// Evolution::run (Evolution.cpp:405-431):
//   evolvePi  (Evolution.cpp:117-178)  pi += dtau/tau [U phi(+) U^+ + ... - 4 phi]
//   evolveE   (Evolution.cpp:180-293)  E_i += i tau dtau/(2g^2) [plaquettes]_TL
//                                             + i dtau/tau [phi~_i, phi], then TL
//   evolvePhi (Evolution.cpp:90-115)   phi += (tau + dtau/2) dtau pi
//   evolveU   (Evolution.cpp:35-88)    U_i <- (1 + X + X^2/2) U_i,
//                                      X = i g^2 dtau/(tau+dtau/2) E_i
// pi and E depend on the old U and phi only, so the kernel updates pi and E in
// place, then phi and U from the new pi and E, writing phi and U to the other
// half of a ping-pong pair (neighbours still read the old values).  Each site
// reads and writes the six fields once: 1728 algorithmic bytes per site-step
// in the reference's representation (six full 3x3 complex matrices); E1, E2,
// phi and pi are Hermitian and stored as 9 reals, so the kernel moves
// 2 x (2 x 144 + 4 x 72) = 1152 bytes per site-step.
// Neighbours are clamped (Lattice.cpp:19-28).
#include "ipg_device.hpp"
#include "ipg_internal.hpp"

namespace {

struct Nb {
    int s, pX, pY, mX, mY, pXmY, mXpY;
};

__device__ __forceinline__ Nb neighbours(int s, int N) {
    int i = s / N, j = s - (s / N) * N;
    Nb n;
    n.s = s;
    n.pX = clampi(i + 1, N) * N + j;
    n.mX = clampi(i - 1, N) * N + j;
    n.pY = i * N + clampi(j + 1, N);
    n.mY = i * N + clampi(j - 1, N);
    n.pXmY = clampi(i + 1, N) * N + clampi(j - 1, N);
    n.mXpY = clampi(i - 1, N) * N + clampi(j + 1, N);
    return n;
}

// Anti-Hermitian 3x3 in compact form: imaginary parts of the diagonal and the
// upper off-diagonal entries (the lower ones are -conj of them).
struct AH {
    double d[3];
    cd o[3];  // (0,1), (0,2), (1,2)
};
// X - X^+
__device__ __forceinline__ AH ah_of(const M3& X) {
    AH a;
    a.d[0] = 2. * X.e[0].i;
    a.d[1] = 2. * X.e[4].i;
    a.d[2] = 2. * X.e[8].i;
    a.o[0] = X.e[1] - cconj(X.e[3]);
    a.o[1] = X.e[2] - cconj(X.e[6]);
    a.o[2] = X.e[5] - cconj(X.e[7]);
    return a;
}
// traceless part of an anti-Hermitian matrix
__device__ __forceinline__ AH ah_tl(AH a) {
    const double t = (a.d[0] + a.d[1] + a.d[2]) / 3.0;
    a.d[0] -= t;
    a.d[1] -= t;
    a.d[2] -= t;
    return a;
}
// traceless Hermitian part of  E + i c1 A + i c2 B  (E Hermitian, A, B anti-Hermitian)
__device__ __forceinline__ M3 e_update(const M3& E, double c1, const AH& A, double c2, const AH& B) {
    M3 h;
    double dd[3];
#pragma unroll
    for (int k = 0; k < 3; k++) dd[k] = E.e[4 * k].r - c1 * A.d[k] - c2 * B.d[k];
    const double t = (dd[0] + dd[1] + dd[2]) / 3.0;
    h.e[0] = cd{dd[0] - t, 0.};
    h.e[4] = cd{dd[1] - t, 0.};
    h.e[8] = cd{dd[2] - t, 0.};
    const int up[3] = {1, 2, 5}, lo[3] = {3, 6, 7};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        cd v = E.e[up[q]] + cd{-c1 * A.o[q].i - c2 * B.o[q].i, c1 * A.o[q].r + c2 * B.o[q].r};
        h.e[up[q]] = v;
        h.e[lo[q]] = cconj(v);
    }
    return h;
}

// (1 + X + X^2/2) with X = i c E  (the n=2 loop of Evolution.cpp:54-68)
__device__ __forceinline__ M3 taylor2(double c, const M3& E) {
    M3 X = mul_i(c, E);
    M3 t = m_one() + 0.5 * X;
    return m_one() + X * t;
}

#define IPG_PHASE() __builtin_amdgcn_sched_barrier(0)

// OCC: waves per SIMD the register budget is sized for (1: no spill, 2:
// half the registers).  TILED: a block takes a tile of 4 rows x 64 sites (one
// wavefront per row, coalesced along y as in the row-major mapping) instead
// of 256 consecutive sites, so that the x-neighbour rows a block reads are
// mostly its own (events with N a multiple of 64; others keep the row-major
// mapping); tiles follow each other along y (TILED 1) or down a 64-wide
// column strip along x (TILED 2), so that the blocks in flight on an XCD
// share their halo rows.  The autotuner picks the build; the results are
// identical.
template <int OCC, int TILED = 0>
__global__ __launch_bounds__(IPG_BLOCK, OCC) void k_step(const EvDesc* __restrict__ evs, int nev, int it, int b0) {
    // b0: first block of a launch that covers one event's blocks only (a
    // multiple of 8, so the XCD of a block is unchanged)
    const int bg = (int)blockIdx.x + b0;
    const int e = find_event(evs, nev, bg);
    const EvDesc& ev = evs[e];
    if (!ev.active || it > ev.itmax) return;
    const int lb = xcd_remap(bg - ev.block0, ev.nblocks);
    long long s;
    if (TILED && (ev.N & 63) == 0) {
        static_assert(IPG_BLOCK == 256, "4 x 64 tiles");
        const int tpr = ev.N >> 6;  // tiles along y
        const int tpc = ev.N >> 2;  // tiles along x
        if (lb >= tpc * tpr) return;
        const int ti = TILED == 1 ? lb / tpr : lb % tpc;
        const int tj = TILED == 1 ? lb - ti * tpr : lb / tpc;
        s = (long long)(ti * 4 + (int)(threadIdx.x >> 6)) * ev.N + tj * 64 + (int)(threadIdx.x & 63);
    } else {
        s = (long long)lb * IPG_BLOCK + threadIdx.x;
        if (s >= ev.S) return;
    }
    const size_t S = ev.S;
    const int p = (it - 1) & 1;
    const double tau = it * ev.dtau;
    const bool half = it == ev.itmax;
    const double dt = half ? ev.dtau / 2. : ev.dtau;
    const Nb n = neighbours((int)s, ev.N);
    const double* __restrict__ Ux = ev.U[p][0];
    const double* __restrict__ Uy = ev.U[p][1];
    const double* __restrict__ phi = ev.phi[p];

    // phase 1: pi (evolvePi) and the commutators [phi~_i, phi] of evolveE
    M3 ux = ldm(Ux, S, s), uy;
    AH c1, c2;
    M3 pin;
    {
        const M3 ph = ldh(phi, S, s);
        M3 t = ux * mul_bd(ldh(phi, S, n.pX), ux);  // phi~_x
        c1 = ah_of(t * ph);
        M3 br = t;
        {
            M3 u = ldm(Ux, S, n.mX);
            br = br + mul_ad(u, ldh(phi, S, n.mX)) * u;
        }
        uy = ldm(Uy, S, s);
        t = uy * mul_bd(ldh(phi, S, n.pY), uy);  // phi~_y
        c2 = ah_of(t * ph);
        br = br + t;
        {
            M3 u = ldm(Uy, S, n.mY);
            br = br + mul_ad(u, ldh(phi, S, n.mY)) * u;
        }
        br = br - 4. * ph;
        pin = herm_part(ldh(ev.pi, S, s) + (dt / tau) * br);
        sth(ev.pio, S, s, pin);
        if (!half) sth(ev.phi[p ^ 1], S, s, ph + ((tau + dt / 2.) * dt) * pin);  // evolvePhi
    }
    IPG_PHASE();
    const double a1 = tau * dt / (2. * ev.g * ev.g);
    const double a2 = dt / tau;
    const double cu = ev.g * ev.g * dt / (tau + dt / 2.);
    // phase 2: E1 from the plaquettes U_12 and U_{1,-2} (evolveE), then U_x
    M3 U12 = (ux * ldm(Uy, S, n.pX)) * mul_bd(dag(ldm(Ux, S, n.pY)), uy);
    {
        M3 P = mul_bd(ux, ldm(Uy, S, n.pXmY)) * mul_ad(ldm(Ux, S, n.mY), ldm(Uy, S, n.mY));
        const AH t1 = ah_tl(ah_of(U12 + P));
        const M3 e1n = e_update(ldh(ev.E1, S, s), a1, t1, a2, c1);
        sth(ev.E1o, S, s, e1n);
        if (!half) stm(ev.U[p ^ 1][0], S, s, taylor2(cu, e1n) * ux);  // evolveU
    }
    IPG_PHASE();
    // phase 3: E2 from U_21 = U_12^+ and U_{2,-1}, then U_y
    {
        M3 P = mul_bd(uy, ldm(Ux, S, n.mXpY)) * mul_ad(ldm(Uy, S, n.mX), ldm(Ux, S, n.mX));
        const AH t1 = ah_tl(ah_of(dag(U12) + P));
        const M3 e2n = e_update(ldh(ev.E2, S, s), a1, t1, a2, c2);
        sth(ev.E2o, S, s, e2n);
        if (!half) stm(ev.U[p ^ 1][1], S, s, taylor2(cu, e2n) * uy);
    }
}

// evolvePhi(dtau, 0) and evolveU(dtau, 0) before the loop (Evolution.cpp:393-394)
__global__ __launch_bounds__(IPG_BLOCK) void k_step0(const EvDesc* __restrict__ evs, int nev) {
    const int e = find_event(evs, nev, blockIdx.x);
    const EvDesc& ev = evs[e];
    if (!ev.active) return;
    const int lb = xcd_remap(blockIdx.x - ev.block0, ev.nblocks);
    const long long s = (long long)lb * IPG_BLOCK + threadIdx.x;
    if (s >= ev.S) return;
    const size_t S = ev.S;
    const double dt = ev.dtau, tau = 0.;
    M3 ph = ldh(ev.phi[0], S, s) + ((tau + dt / 2.) * dt) * ldh(ev.pi, S, s);
    sth(ev.phi[0], S, s, ph);
    const double cu = ev.g * ev.g * dt / (tau + dt / 2.);
    stm(ev.U[0][0], S, s, taylor2(cu, ldh(ev.E1, S, s)) * ldm(ev.U[0][0], S, s));
    stm(ev.U[0][1], S, s, taylor2(cu, ldh(ev.E2, S, s)) * ldm(ev.U[0][1], S, s));
}

// ---- reductions ---------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}
// deterministic block sum (fixed tree): result valid in thread 0
__device__ double block_sum(double v) {
    __shared__ double w[IPG_BLOCK / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0)
        for (int k = 0; k < IPG_BLOCK / 64; k++) r += w[k];
    __syncthreads();
    return r;
}
__device__ double block_max(double v) {
    __shared__ double w[IPG_BLOCK / 64];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0)
        for (int k = 0; k < IPG_BLOCK / 64; k++) r = fmax(r, w[k]);
    __syncthreads();
    return r;
}

// ---- T^{mu nu} (Evolution.cpp:659-1246) ---------------------------------------
__device__ __forceinline__ M3 m4(const M3& a, const M3& b, const M3& c, const M3& d) { return ((a * b) * c) * d; }
__device__ __forceinline__ M3 m3(const M3& a, const M3& b, const M3& c) { return (a * b) * c; }
// ((P1-P2) - tr/3) + P3 - P4 - tr(P3-P4)/3
__device__ __forceinline__ M3 tlquad(const M3& P1, const M3& P2, const M3& P3, const M3& P4) {
    return traceless(P1 - P2) + traceless(P3 - P4);
}

// Each block below reloads its inputs (L1/L2 hits) after a compiler memory
// barrier, so that only the matrices of one term are live at a time.
#define IPG_FENCE() asm volatile("" ::: "memory")

// T^{mu nu} in three passes (register pressure): diagonal components,
// T^{tau eta} and T^{xy}; the block sums of T^tautau
__global__ __launch_bounds__(IPG_BLOCK) void k_tmunu_a(const EvDesc* __restrict__ evs, int nev, int it) {
    const int e = find_event(evs, nev, blockIdx.x);
    const EvDesc& ev = evs[e];
    if (!ev.active || it != ev.itmax) return;
    const int lb = xcd_remap(blockIdx.x - ev.block0, ev.nblocks);
    const long long s = (long long)lb * IPG_BLOCK + threadIdx.x;
    double ttsum = 0.;
    if (s < ev.S) {
        const size_t S = ev.S;
        const int N = ev.N;
        const int p = (it - 1) & 1;
        const double* Ux = ev.U[p][0];
        const double* Uy = ev.U[p][1];
        const double* phi = ev.phi[p];
        const double* E1 = ev.E1;
        const double* E2 = ev.E2;
        [[maybe_unused]] const double* PI = ev.pi;
        [[maybe_unused]] const double tau = it * ev.dtau, g = ev.g, a = ev.a;
        const int i = (int)(s / N), j = (int)(s - (s / N) * N);
        auto at = [&](int di, int dj) { return (size_t)clampi(i + di, N) * N + clampi(j + dj, N); };
        // the neighbours of the whole T^{mu nu} stencil (each pass uses a subset)
        [[maybe_unused]] const size_t pX = at(1, 0), pY = at(0, 1), pXY = at(1, 1), mX = at(-1, 0), mY = at(0, -1);
        [[maybe_unused]] const size_t mXpY = at(-1, 1), pXmY = at(1, -1), p2X = at(2, 0), p2Y = at(0, 2);
        [[maybe_unused]] const size_t p2XY = at(2, 1), pX2Y = at(1, 2);
        double* T = ev.T;
        [[maybe_unused]] const double a2 = a * a, a4 = a2 * a2, a5 = a4 * a, a6 = a4 * a2;
        // phi gradients: gX = Ux phi(x+x) Ux^+ - phi, gY, gXY1 = Ux(+y) phi(+x+y) Ux(+y)^+ - phi(+y),
        // gXY2 = Uy(+x) phi(+x+y) Uy(+x)^+ - phi(+x)
        auto gX = [&]() { M3 u = ldm<true>(Ux, S, s); return m3(u, ldh<true>(phi, S, pX), dag(u)) - ldh<true>(phi, S, s); };
        auto gY = [&]() { M3 u = ldm<true>(Uy, S, s); return m3(u, ldh<true>(phi, S, pY), dag(u)) - ldh<true>(phi, S, s); };
        auto gXY1 = [&]() { M3 u = ldm<true>(Ux, S, pY); return m3(u, ldh<true>(phi, S, pXY), dag(u)) - ldh<true>(phi, S, pY); };
        auto gXY2 = [&]() { M3 u = ldm<true>(Uy, S, pX); return m3(u, ldh<true>(phi, S, pXY), dag(u)) - ldh<true>(phi, S, pX); };

        // electric and longitudinal (Evolution.cpp:754-814)
        double tt, xx, yy, ee;
        {
            const M3 e1 = ldh<true>(E1, S, s), e1p = ldh<true>(E1, S, pY), e2 = ldh<true>(E2, S, s), e2p = ldh<true>(E2, S, pX);
            const double t11 = retrprod(e1, e1), t1p = retrprod(e1p, e1p), t22 = retrprod(e2, e2),
                         t2p = retrprod(e2p, e2p);
            double lng = 0.;
            {
                M3 q = ldh<true>(PI, S, s);
                lng += retrprod(q, q);
                q = ldh<true>(PI, S, pX);
                lng += retrprod(q, q);
                q = ldh<true>(PI, S, pY);
                lng += retrprod(q, q);
                q = ldh<true>(PI, S, pXY);
                lng += retrprod(q, q);
                lng /= 4.;
            }
            const double gt = g * g / tau / tau;
            tt = gt * (t11 + t1p + t22 + t2p) / 2. + lng;
            xx = gt * (-t11 - t1p + t22 + t2p) / 2. + lng;
            yy = gt * (t11 + t1p - t22 - t2p) / 2. + lng;
            ee = 1. / tau / tau * (gt * (t11 + t1p + t22 + t2p) / 2. - lng);
        }
        IPG_FENCE();
        // magnetic and transverse phi gradients (Evolution.cpp:816-905)
        {
            double rplaq = 3.0;
            if (i != N - 1) {
                const M3 ux = ldm<true>(Ux, S, s);
                rplaq = retrprod(ux, ldm<true>(Uy, S, pX) * mul_bd(dag(ldm<true>(Ux, S, pY)), ldm<true>(Uy, S, s)));
            }
            double A, B, C, D;
            {
                M3 q = gX();
                A = retrprod(q, q);
            }
            {
                M3 q = gXY1();
                B = retrprod(q, q);
            }
            {
                M3 q = gY();
                C = retrprod(q, q);
            }
            {
                M3 q = gXY2();
                D = retrprod(q, q);
            }
            const double mag = 2. / (g * g) * (3.0 - rplaq);
            const double h = 0.5 / tau / tau;
            tt = (tt + mag) + h * (A + B + C + D);
            xx = (xx + mag) + h * (A + B - C - D);
            yy = (yy + mag) + h * (-A - B + C + D);
            ee = ee + 1. / tau / tau * (-mag + h * (A + B + C + D));
        }
        T[0 * S + s] = tt / a4;
        T[1 * S + s] = xx / a4;
        T[2 * S + s] = yy / a4;
        T[3 * S + s] = ee / a6;
        ttsum = tt / a4;
        IPG_FENCE();
        // T^{tau eta}
        {
            double re = retrprod(ldh<true>(E1, S, s), gX());
            re += retrprod(ldh<true>(E1, S, pY), gXY1());
            re += retrprod(ldh<true>(E2, S, s), gY());
            re += retrprod(ldh<true>(E2, S, pX), gXY2());
            T[6 * S + s] = (g / tau / tau / tau * re) / a5;
        }
        IPG_FENCE();
        // T^{xy}
        {
            const M3 ux = ldm<true>(Ux, S, s), uy = ldm<true>(Uy, S, s);
            double b;
            {
                const M3 el = (-1. / 4. * g * g * (ldh<true>(E1, S, s) + m3(uy, ldh<true>(E1, S, pY), dag(uy)))) *
                              (ldh<true>(E2, S, s) + m3(ux, ldh<true>(E2, S, pX), dag(ux)));
                b = 4. * tr(el).r;
            }
            const M3 yl = m3(uy, gXY1(), dag(uy));  // U_y gXY1 U_y^+
            const M3 xl = m3(ux, gXY2(), dag(ux));  // U_x gXY2 U_x^+
            const M3 gx = gX(), gy = gY();
            b += retrprod(gx, gy) + retrprod(yl, gy) + retrprod(gx, xl) + retrprod(yl, xl);
            T[7 * S + s] = (2. / tau / tau * (0.25 * b)) / a4;
        }
        IPG_FENCE();
    }
    double bs = block_sum(ttsum);
    if (threadIdx.x == 0) ev.partial[lb] = bs;
}

// clover parts of T^{tau x}, T^{tau y}, T^{x eta}, T^{y eta}
__global__ __launch_bounds__(IPG_BLOCK) void k_tmunu_b(const EvDesc* __restrict__ evs, int nev, int it) {
    const int e = find_event(evs, nev, blockIdx.x);
    const EvDesc& ev = evs[e];
    if (!ev.active || it != ev.itmax) return;
    const int lb = xcd_remap(blockIdx.x - ev.block0, ev.nblocks);
    const long long s = (long long)lb * IPG_BLOCK + threadIdx.x;
    if (s < ev.S) {
        const size_t S = ev.S;
        const int N = ev.N;
        const int p = (it - 1) & 1;
        const double* Ux = ev.U[p][0];
        const double* Uy = ev.U[p][1];
        const double* phi = ev.phi[p];
        const double* E1 = ev.E1;
        const double* E2 = ev.E2;
        [[maybe_unused]] const double* PI = ev.pi;
        [[maybe_unused]] const double tau = it * ev.dtau, g = ev.g, a = ev.a;
        const int i = (int)(s / N), j = (int)(s - (s / N) * N);
        auto at = [&](int di, int dj) { return (size_t)clampi(i + di, N) * N + clampi(j + dj, N); };
        // the neighbours of the whole T^{mu nu} stencil (each pass uses a subset)
        [[maybe_unused]] const size_t pX = at(1, 0), pY = at(0, 1), pXY = at(1, 1), mX = at(-1, 0), mY = at(0, -1);
        [[maybe_unused]] const size_t mXpY = at(-1, 1), pXmY = at(1, -1), p2X = at(2, 0), p2Y = at(0, 2);
        [[maybe_unused]] const size_t p2XY = at(2, 1), pX2Y = at(1, 2);
        double* T = ev.T;
        [[maybe_unused]] const double a2 = a * a, a4 = a2 * a2, a5 = a4 * a, a6 = a4 * a2;
        // phi gradients: gX = Ux phi(x+x) Ux^+ - phi, gY, gXY1 = Ux(+y) phi(+x+y) Ux(+y)^+ - phi(+y),
        // gXY2 = Uy(+x) phi(+x+y) Uy(+x)^+ - phi(+x)
        auto gX = [&]() { M3 u = ldm<true>(Ux, S, s); return m3(u, ldh<true>(phi, S, pX), dag(u)) - ldh<true>(phi, S, s); };
        auto gY = [&]() { M3 u = ldm<true>(Uy, S, s); return m3(u, ldh<true>(phi, S, pY), dag(u)) - ldh<true>(phi, S, s); };
        auto gXY1 = [&]() { M3 u = ldm<true>(Ux, S, pY); return m3(u, ldh<true>(phi, S, pXY), dag(u)) - ldh<true>(phi, S, pY); };
        auto gXY2 = [&]() { M3 u = ldm<true>(Uy, S, pX); return m3(u, ldh<true>(phi, S, pXY), dag(u)) - ldh<true>(phi, S, pX); };

        // T^{tau x}, T^{x eta}, T^{tau y}, T^{y eta} (Evolution.cpp:1047-1234).
        // The clover combinations enter with opposite signs, e.g. the T^{x eta}
        // clover (Pb - Pa)_TL + (Pd - Pc)_TL is -X1.
        double ttx_im, tty_im, xe_im, ye_im;
        {
            // X1 = (Pa - Pb)_TL + (Pc - Pd)_TL,
            // Pa = Uy Ux(+y) Uy(+x)^+ Ux^+, Pb = Ux Uy(+x) Ux(+y)^+ Uy^+
            const M3 ux = ldm<true>(Ux, S, s), uy = ldm<true>(Uy, S, s);
            M3 X = traceless(m4(uy, ldm<true>(Ux, S, pY), dag(ldm<true>(Uy, S, pX)), dag(ux)) -
                             m4(ux, ldm<true>(Uy, S, pX), dag(ldm<true>(Ux, S, pY)), dag(uy)));
            {
                const M3 UxmX = ldm<true>(Ux, S, mX), UxmXpY = ldm<true>(Ux, S, mXpY), UymX = ldm<true>(Uy, S, mX);
                X = X + traceless(m4(dag(UxmX), UymX, UxmXpY, dag(uy)) - m4(uy, dag(UxmXpY), dag(UymX), UxmX));
            }
            ttx_im = imtrprod(ldh<true>(E2, S, s), X);
            xe_im = -imtrprod(X, gY());
        }
        IPG_FENCE();
        {
            // Y1 = (Pb - Pa)_TL + (Qc - Qd)_TL
            const M3 ux = ldm<true>(Ux, S, s), uy = ldm<true>(Uy, S, s);
            M3 Y = traceless(m4(ux, ldm<true>(Uy, S, pX), dag(ldm<true>(Ux, S, pY)), dag(uy)) -
                             m4(uy, ldm<true>(Ux, S, pY), dag(ldm<true>(Uy, S, pX)), dag(ux)));
            {
                const M3 UymY = ldm<true>(Uy, S, mY), UxmY = ldm<true>(Ux, S, mY), UypXmY = ldm<true>(Uy, S, pXmY);
                Y = Y + traceless(m4(dag(UymY), UxmY, UypXmY, dag(ux)) - m4(ux, dag(UypXmY), dag(UxmY), UymY));
            }
            tty_im = imtrprod(ldh<true>(E1, S, s), Y);
            ye_im = -imtrprod(Y, gX());
        }
        IPG_FENCE();
        {
            // X2 = (Pe - Pf)_TL + (Pg - Ph)_TL
            const M3 UxpX = ldm<true>(Ux, S, pX), UxpXpY = ldm<true>(Ux, S, pXY), Uyp2X = ldm<true>(Uy, S, p2X);
            const M3 UypX = ldm<true>(Uy, S, pX);
            M3 X = traceless(m4(UypX, UxpXpY, dag(Uyp2X), dag(UxpX)) - m4(UxpX, Uyp2X, dag(UxpXpY), dag(UypX)));
            {
                const M3 ux = ldm<true>(Ux, S, s), uy = ldm<true>(Uy, S, s), UxpY = ldm<true>(Ux, S, pY);
                X = X + traceless(m4(dag(ux), uy, UxpY, dag(UypX)) - m4(UypX, dag(UxpY), dag(uy), ux));
            }
            ttx_im += imtrprod(ldh<true>(E2, S, pX), X);
            xe_im -= imtrprod(X, gXY2());
        }
        IPG_FENCE();
        {
            // Y2 = (Qe - Qf)_TL + (Qg - Qh)_TL
            const M3 UxpY = ldm<true>(Ux, S, pY), UypXpY = ldm<true>(Uy, S, pXY), Uxp2Y = ldm<true>(Ux, S, p2Y);
            const M3 UypY = ldm<true>(Uy, S, pY);
            M3 Y = traceless(m4(UxpY, UypXpY, dag(Uxp2Y), dag(UypY)) - m4(UypY, Uxp2Y, dag(UypXpY), dag(UxpY)));
            {
                const M3 ux = ldm<true>(Ux, S, s), uy = ldm<true>(Uy, S, s), UypX = ldm<true>(Uy, S, pX);
                Y = Y + traceless(m4(dag(uy), ux, UypX, dag(UxpY)) - m4(UxpY, dag(UypX), dag(ux), uy));
            }
            tty_im += imtrprod(ldh<true>(E1, S, pY), Y);
            ye_im -= imtrprod(Y, gXY1());
        }
        // raw clover sums, finished by k_tmunu_c
        T[4 * S + s] = ttx_im;
        T[5 * S + s] = tty_im;
        T[8 * S + s] = xe_im;
        T[9 * S + s] = ye_im;
    }
}

// pi grad-phi and E pi parts; finishes the four flux components
__global__ __launch_bounds__(IPG_BLOCK) void k_tmunu_c(const EvDesc* __restrict__ evs, int nev, int it) {
    const int e = find_event(evs, nev, blockIdx.x);
    const EvDesc& ev = evs[e];
    if (!ev.active || it != ev.itmax) return;
    const int lb = xcd_remap(blockIdx.x - ev.block0, ev.nblocks);
    const long long s = (long long)lb * IPG_BLOCK + threadIdx.x;
    if (s < ev.S) {
        const size_t S = ev.S;
        const int N = ev.N;
        const int p = (it - 1) & 1;
        const double* Ux = ev.U[p][0];
        const double* Uy = ev.U[p][1];
        const double* phi = ev.phi[p];
        const double* E1 = ev.E1;
        const double* E2 = ev.E2;
        [[maybe_unused]] const double* PI = ev.pi;
        [[maybe_unused]] const double tau = it * ev.dtau, g = ev.g, a = ev.a;
        const int i = (int)(s / N), j = (int)(s - (s / N) * N);
        auto at = [&](int di, int dj) { return (size_t)clampi(i + di, N) * N + clampi(j + dj, N); };
        // the neighbours of the whole T^{mu nu} stencil (each pass uses a subset)
        [[maybe_unused]] const size_t pX = at(1, 0), pY = at(0, 1), pXY = at(1, 1), mX = at(-1, 0), mY = at(0, -1);
        [[maybe_unused]] const size_t mXpY = at(-1, 1), pXmY = at(1, -1), p2X = at(2, 0), p2Y = at(0, 2);
        [[maybe_unused]] const size_t p2XY = at(2, 1), pX2Y = at(1, 2);
        double* T = ev.T;
        [[maybe_unused]] const double a2 = a * a, a4 = a2 * a2, a5 = a4 * a, a6 = a4 * a2;
        // pi phi-gradient parts of T^{tau x} and T^{tau y}
        double ttx_re, tty_re;
        const double ttx_im = T[4 * S + s], tty_im = T[5 * S + s], xe_im = T[8 * S + s], ye_im = T[9 * S + s];
        IPG_FENCE();
        {
            M3 u = ldm<true>(Ux, S, s), v = ldm<true>(Ux, S, mX);
            ttx_re = retrprod(ldh<true>(PI, S, s), m3(u, ldh<true>(phi, S, pX), dag(u)) - m3(dag(v), ldh<true>(phi, S, mX), v));
            u = ldm<true>(Ux, S, pY);
            v = ldm<true>(Ux, S, mXpY);
            ttx_re += retrprod(ldh<true>(PI, S, pY), m3(u, ldh<true>(phi, S, pXY), dag(u)) - m3(dag(v), ldh<true>(phi, S, mXpY), v));
            u = ldm<true>(Ux, S, pX);
            v = ldm<true>(Ux, S, s);
            ttx_re += retrprod(ldh<true>(PI, S, pX), m3(u, ldh<true>(phi, S, p2X), dag(u)) - m3(dag(v), ldh<true>(phi, S, s), v));
            u = ldm<true>(Ux, S, pXY);
            v = ldm<true>(Ux, S, pY);
            ttx_re += retrprod(ldh<true>(PI, S, pXY), m3(u, ldh<true>(phi, S, p2XY), dag(u)) - m3(dag(v), ldh<true>(phi, S, pY), v));
        }
        IPG_FENCE();
        {
            M3 u = ldm<true>(Uy, S, s), v = ldm<true>(Uy, S, mY);
            tty_re = retrprod(ldh<true>(PI, S, s), m3(u, ldh<true>(phi, S, pY), dag(u)) - m3(dag(v), ldh<true>(phi, S, mY), v));
            u = ldm<true>(Uy, S, pX);
            v = ldm<true>(Uy, S, pXmY);
            tty_re += retrprod(ldh<true>(PI, S, pX), m3(u, ldh<true>(phi, S, pXY), dag(u)) - m3(dag(v), ldh<true>(phi, S, pXmY), v));
            u = ldm<true>(Uy, S, pY);
            v = ldm<true>(Uy, S, s);
            tty_re += retrprod(ldh<true>(PI, S, pY), m3(u, ldh<true>(phi, S, p2Y), dag(u)) - m3(dag(v), ldh<true>(phi, S, s), v));
            u = ldm<true>(Uy, S, pXY);
            v = ldm<true>(Uy, S, pX);
            tty_re += retrprod(ldh<true>(PI, S, pXY), m3(u, ldh<true>(phi, S, pX2Y), dag(u)) - m3(dag(v), ldh<true>(phi, S, pX), v));
        }
        T[4 * S + s] = (2. / tau / 8. * ttx_im - 2. / 8. / tau * ttx_re) / a4;
        T[5 * S + s] = (2. / tau / 8. * tty_im - 2. / 8. / tau * tty_re) / a4;
        IPG_FENCE();
        // E pi parts of T^{x eta}, T^{y eta}
        {
            M3 u = ldm<true>(Ux, S, s);
            double r1 = retrprod(ldh<true>(E1, S, s), ldh<true>(PI, S, s) + m3(u, ldh<true>(PI, S, pX), dag(u)));
            u = ldm<true>(Ux, S, pY);
            r1 += retrprod(ldh<true>(E1, S, pY), ldh<true>(PI, S, pY) + m3(u, ldh<true>(PI, S, pXY), dag(u)));
            T[8 * S + s] = (-2. / tau / tau * (1. / 4. * g * r1 + 1. / 8. / g * xe_im)) / a5;
        }
        IPG_FENCE();
        {
            M3 u = ldm<true>(Uy, S, s);
            double r1 = retrprod(ldh<true>(E2, S, s), ldh<true>(PI, S, s) + m3(u, ldh<true>(PI, S, pY), dag(u)));
            u = ldm<true>(Uy, S, pX);
            r1 += retrprod(ldh<true>(E2, S, pX), ldh<true>(PI, S, pX) + m3(u, ldh<true>(PI, S, pXY), dag(u)));
            T[9 * S + s] = (-2. / tau / tau * (1. / 4. * g * r1 + 1. / 8. / g * ye_im)) / a5;
        }
    }
}

__global__ __launch_bounds__(IPG_BLOCK) void k_gauss(const EvDesc* __restrict__ evs, int nev, int it) {
    const int e = find_event(evs, nev, blockIdx.x);
    const EvDesc& ev = evs[e];
    if (!ev.active || it != ev.itmax) return;
    const int lb = xcd_remap(blockIdx.x - ev.block0, ev.nblocks);
    const long long s = (long long)lb * IPG_BLOCK + threadIdx.x;
    double v = 0.;
    if (s < ev.S) {
        const size_t S = ev.S;
        const int p = (it - 1) & 1;
        const Nb n = neighbours((int)s, ev.N);
        const M3 uxm = ldm(ev.U[p][0], S, n.mX), uym = ldm(ev.U[p][1], S, n.mY);
        const M3 ph = ldh(ev.phi[p], S, s), pp = ldh(ev.pi, S, s);
        M3 G = mul_ad(uxm, ldh(ev.E1, S, n.mX)) * uxm - ldh(ev.E1, S, s);
        G = G + mul_ad(uym, ldh(ev.E2, S, n.mY)) * uym - ldh(ev.E2, S, s);
        G = G - mul_i(1.0, ph * pp - pp * ph);
        v = square(G);
    }
    double bm = block_max(v);
    if (threadIdx.x == 0) ev.partial[ev.nblocks + lb] = bm;
}

// per event: out[4e+0] = sum T^tautau (fm^-4), out[4e+1] = Gauss max
__global__ __launch_bounds__(IPG_BLOCK) void k_finish(const EvDesc* __restrict__ evs, int nev, int it,
                                                     double* __restrict__ out) {
    const int e = blockIdx.x;
    if (e >= nev) return;
    const EvDesc& ev = evs[e];
    if (!ev.active || it != ev.itmax) return;
    const int nb = (int)((ev.S + IPG_BLOCK - 1) / IPG_BLOCK);
    // fixed-order per-thread partial sums, then a fixed tree
    double sm = 0., mx = 0.;
    for (int b = threadIdx.x; b < ev.nblocks; b += IPG_BLOCK) {
        int l = b;  // logical block index
        if (l < nb) {
            sm += ev.partial[l];
            mx = fmax(mx, ev.partial[ev.nblocks + l]);
        }
    }
    double ts = block_sum(sm);
    double tm = block_max(mx);
    if (threadIdx.x == 0) {
        out[4 * e + 0] = ts;
        out[4 * e + 1] = tm;
    }
}

}  // namespace

void ipg_launch_step0(const EvDesc* d_evs, int nev, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_step0, dim3(nblk), dim3(IPG_BLOCK), 0, st, d_evs, nev);
}
// variant 0: k_step<1>, 1: k_step<2>, 2: k_step<1> on 4 x 64 tiles along y,
// 3: k_step<1> on 4 x 64 tiles down column strips (k_step<2> with opaque
// lane offsets spills 284 B per lane instead of 508, and still takes 158 us
// against 83 at N = 512)
void ipg_launch_step(const EvDesc* d_evs, int nev, int nblk, int it, hipStream_t st, int variant) {
    ipg_launch_step_event(d_evs, nev, 0, nblk, it, st, variant);
}
void ipg_launch_step_event(const EvDesc* d_evs, int nev, int block0, int nblocks, int it, hipStream_t st,
                           int variant) {
    if (variant == 1)
        hipLaunchKernelGGL(k_step<2>, dim3(nblocks), dim3(IPG_BLOCK), 0, st, d_evs, nev, it, block0);
    else if (variant == 2)
        hipLaunchKernelGGL((k_step<1, 1>), dim3(nblocks), dim3(IPG_BLOCK), 0, st, d_evs, nev, it, block0);
    else if (variant == 3)
        hipLaunchKernelGGL((k_step<1, 2>), dim3(nblocks), dim3(IPG_BLOCK), 0, st, d_evs, nev, it, block0);

    else
        hipLaunchKernelGGL(k_step<1>, dim3(nblocks), dim3(IPG_BLOCK), 0, st, d_evs, nev, it, block0);
}
void ipg_launch_tmunu(const EvDesc* d_evs, int nev, int nblk, int it, hipStream_t st) {
    hipLaunchKernelGGL(k_tmunu_a, dim3(nblk), dim3(IPG_BLOCK), 0, st, d_evs, nev, it);
    hipLaunchKernelGGL(k_tmunu_b, dim3(nblk), dim3(IPG_BLOCK), 0, st, d_evs, nev, it);
    hipLaunchKernelGGL(k_tmunu_c, dim3(nblk), dim3(IPG_BLOCK), 0, st, d_evs, nev, it);
}
void ipg_launch_gauss(const EvDesc* d_evs, int nev, int nblk, int it, hipStream_t st) {
    hipLaunchKernelGGL(k_gauss, dim3(nblk), dim3(IPG_BLOCK), 0, st, d_evs, nev, it);
}
void ipg_launch_finish(const EvDesc* d_evs, int nev, int it, double* d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_finish, dim3(nev), dim3(IPG_BLOCK), 0, st, d_evs, nev, it, d_out);
}
