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

__global__ __launch_bounds__(IPG_BLOCK) void k_step(const EvDesc* __restrict__ evs, int nev, int it) {
    const int e = find_event(evs, nev, blockIdx.x);
    const EvDesc& ev = evs[e];
    if (!ev.active || it > ev.itmax) return;
    const int lb = xcd_remap(blockIdx.x - ev.block0, ev.nblocks);
    const long long s = (long long)lb * IPG_BLOCK + threadIdx.x;
    if (s >= ev.S) return;
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
        sth(ev.pi, S, s, pin);
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
        sth(ev.E1, S, s, e1n);
        if (!half) stm(ev.U[p ^ 1][0], S, s, taylor2(cu, e1n) * ux);  // evolveU
    }
    IPG_PHASE();
    // phase 3: E2 from U_21 = U_12^+ and U_{2,-1}, then U_y
    {
        M3 P = mul_bd(uy, ldm(Ux, S, n.mXpY)) * mul_ad(ldm(Uy, S, n.mX), ldm(Ux, S, n.mX));
        const AH t1 = ah_tl(ah_of(dag(U12) + P));
        const M3 e2n = e_update(ldh(ev.E2, S, s), a1, t1, a2, c2);
        sth(ev.E2, S, s, e2n);
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

__global__ __launch_bounds__(IPG_BLOCK) void k_tmunu(const EvDesc* __restrict__ evs, int nev, int it) {
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
        const double tau = it * ev.dtau, g = ev.g, a = ev.a;
        const int i = (int)(s / N), j = (int)(s - (s / N) * N);
        auto at = [&](int di, int dj) { return (size_t)clampi(i + di, N) * N + clampi(j + dj, N); };
        const size_t pX = at(1, 0), pY = at(0, 1), pXY = at(1, 1), mX = at(-1, 0), mY = at(0, -1);
        const size_t mXpY = at(-1, 1), pXmY = at(1, -1), p2X = at(2, 0), p2Y = at(0, 2);
        const size_t p2XY = at(2, 1), pX2Y = at(1, 2);
        double* T = ev.T;
        // electric and longitudinal
        const M3 e1 = ldh(ev.E1, S, s), e2 = ldh(ev.E2, S, s), e1p = ldh(ev.E1, S, pY), e2p = ldh(ev.E2, S, pX);
        const M3 p0 = ldh(ev.pi, S, s), pXm = ldh(ev.pi, S, pX), pYm = ldh(ev.pi, S, pY), pXYm = ldh(ev.pi, S, pXY);
        const double t11 = retrprod(e1, e1), t1p = retrprod(e1p, e1p), t22 = retrprod(e2, e2), t2p = retrprod(e2p, e2p);
        const double lng = (retrprod(p0, p0) + retrprod(pXm, pXm) + retrprod(pYm, pYm) + retrprod(pXYm, pXYm)) / 4.;
        const double gt = g * g / tau / tau;
        double tt = gt * (t11 + t1p + t22 + t2p) / 2. + lng;
        double xx = gt * (-t11 - t1p + t22 + t2p) / 2. + lng;
        double yy = gt * (t11 + t1p - t22 - t2p) / 2. + lng;
        double ee = 1. / tau / tau * (gt * (t11 + t1p + t22 + t2p) / 2. - lng);
        // magnetic and transverse phi gradients
        const M3 ux = ldm(Ux, S, s), uy = ldm(Uy, S, s);
        const M3 UxpY = ldm(Ux, S, pY), UypX = ldm(Uy, S, pX);
        double rplaq = 3.0;
        if (i != N - 1) rplaq = retrprod(ux, UypX * (mul_bd(dag(UxpY), uy)));
        const M3 ph = ldh(phi, S, s), phX = ldh(phi, S, pX), phY = ldh(phi, S, pY), phXY = ldh(phi, S, pXY);
        const M3 Udx = dag(ux), Udy = dag(uy), UDxpY = dag(UxpY), UDypX = dag(UypX);
        const M3 gX = m3(ux, phX, Udx) - ph;   // Ux phiX Ux^+ - phi
        const M3 gY = m3(uy, phY, Udy) - ph;
        const M3 gXY1 = m3(UxpY, phXY, UDxpY) - phY;
        const M3 gXY2 = m3(UypX, phXY, UDypX) - phX;
        const double A = retrprod(gX, gX), B = retrprod(gXY1, gXY1), C = retrprod(gY, gY), D = retrprod(gXY2, gXY2);
        const double mag = 2. / (g * g) * (3.0 - rplaq);
        const double h = 0.5 / tau / tau;
        tt = (tt + mag) + h * (A + B + C + D);
        xx = (xx + mag) + h * (A + B - C - D);
        yy = (yy + mag) + h * (-A - B + C + D);
        ee = ee + 1. / tau / tau * (-mag + h * (A + B + C + D));
        const double a2 = a * a, a4 = a2 * a2, a5 = a4 * a, a6 = a4 * a2;
        T[0 * S + s] = tt / a4;
        T[1 * S + s] = xx / a4;
        T[2 * S + s] = yy / a4;
        T[3 * S + s] = ee / a6;
        ttsum = tt / a4;

        // flux components
        const M3 UxmX = ldm(Ux, S, mX), UDxmX = dag(UxmX);
        const M3 UxmXpY = ldm(Ux, S, mXpY), UDxmXpY = dag(UxmXpY);
        const M3 UxpX = ldm(Ux, S, pX), UDxpX = dag(UxpX);
        const M3 UxpXpY = ldm(Ux, S, pXY), UDxpXpY = dag(UxpXpY);
        const M3 UymY = ldm(Uy, S, mY), UDymY = dag(UymY);
        const M3 UypXmY = ldm(Uy, S, pXmY), UDypXmY = dag(UypXmY);
        const M3 UypY = ldm(Uy, S, pY), UDypY = dag(UypY);
        const M3 Uyp2X = ldm(Uy, S, p2X), UDyp2X = dag(Uyp2X);
        const M3 Uxp2Y = ldm(Ux, S, p2Y), UDxp2Y = dag(Uxp2Y);
        const M3 UypXpY = ldm(Uy, S, pXY), UDypXpY = dag(UypXpY);
        const M3 UymX = ldm(Uy, S, mX), UDymX = dag(UymX);
        const M3 UxmY = ldm(Ux, S, mY), UDxmY = dag(UxmY);

        // plaquette-type clovers shared by several components
        const M3 Pa = m4(uy, UxpY, UDypX, Udx);       // U_y U_x(+y) U_y(+x)^+ U_x^+
        const M3 Pb = m4(ux, UypX, UDxpY, Udy);       // U_x U_y(+x) U_x(+y)^+ U_y^+
        const M3 Pc = m4(UDxmX, UymX, UxmXpY, Udy);
        const M3 Pd = m4(uy, UDxmXpY, UDymX, UxmX);
        const M3 Pe = m4(UypX, UxpXpY, UDyp2X, UDxpX);
        const M3 Pf = m4(UxpX, Uyp2X, UDxpXpY, UDypX);
        const M3 Pg = m4(Udx, uy, UxpY, UDypX);
        const M3 Ph = m4(UypX, UDxpY, Udy, ux);
        const M3 Qc = m4(UDymY, UxmY, UypXmY, Udx);
        const M3 Qd = m4(ux, UDypXmY, UDxmY, UymY);
        const M3 Qe = m4(UxpY, UypXpY, UDxp2Y, UDypY);
        const M3 Qf = m4(UypY, Uxp2Y, UDypXpY, UDxpY);
        const M3 Qg = m4(Udy, ux, UypX, UDxpY);
        const M3 Qh = m4(UxpY, UDypX, Udx, uy);

        // T^{tau x}
        {
            const M3 phmX = ldh(phi, S, mX), phmXpY = ldh(phi, S, mXpY), ph2X = ldh(phi, S, p2X);
            const M3 ph2XY = ldh(phi, S, p2XY);
            double im = imtrprod(e2, tlquad(Pa, Pb, Pc, Pd)) + imtrprod(e2p, tlquad(Pe, Pf, Pg, Ph));
            double re = retrprod(p0, m3(ux, phX, Udx) - m3(UDxmX, phmX, UxmX)) +
                        retrprod(pYm, m3(UxpY, phXY, UDxpY) - m3(UDxmXpY, phmXpY, UxmXpY)) +
                        retrprod(pXm, m3(UxpX, ph2X, UDxpX) - m3(Udx, ph, ux)) +
                        retrprod(pXYm, m3(UxpXpY, ph2XY, UDxpXpY) - m3(UDxpY, phY, UxpY));
            T[4 * S + s] = (2. / tau / 8. * im - 2. / 8. / tau * re) / a4;
        }
        // T^{tau y}
        {
            const M3 phmY = ldh(phi, S, mY), phpXmY = ldh(phi, S, pXmY), ph2Y = ldh(phi, S, p2Y);
            const M3 phX2Y = ldh(phi, S, pX2Y);
            double im = imtrprod(e1, tlquad(Pb, Pa, Qc, Qd)) + imtrprod(e1p, tlquad(Qe, Qf, Qg, Qh));
            double re = retrprod(p0, m3(uy, phY, Udy) - m3(UDymY, phmY, UymY)) +
                        retrprod(pXm, m3(UypX, phXY, UDypX) - m3(UDypXmY, phpXmY, UypXmY)) +
                        retrprod(pYm, m3(UypY, ph2Y, UDypY) - m3(Udy, ph, uy)) +
                        retrprod(pXYm, m3(UypXpY, phX2Y, UDypXpY) - m3(UDypX, phX, UypX));
            T[5 * S + s] = (2. / tau / 8. * im - 2. / 8. / tau * re) / a4;
        }
        // T^{tau eta}
        {
            double re = retrprod(e1, gX) + retrprod(e1p, gXY1) + retrprod(e2, gY) + retrprod(e2p, gXY2);
            T[6 * S + s] = (g / tau / tau / tau * re) / a5;
        }
        // T^{xy}
        {
            M3 el = (-1. / 4. * g * g * (e1 + m3(uy, e1p, Udy))) * (e2 + m3(ux, e2p, Udx));
            M3 yl = (uy * gXY1) * Udy;        // U_y gXY1 U_y^+
            M3 xl = (ux * gXY2) * Udx;        // U_x gXY2 U_x^+
            M3 b = gX * gY + yl * gY + gX * xl + yl * xl;
            T[7 * S + s] = (2. / tau / tau * (retrprod(el, m_one()) + 0.25 * retrprod(b, m_one()))) / a4;
        }
        // T^{x eta}
        {
            double r1 = 1. / 4. * g * (retrprod(e1, p0 + m3(ux, pXm, Udx)) + retrprod(e1p, pYm + m3(UxpY, pXYm, UDxpY)));
            double r2 = 1. / 8. / g * (imtrprod(tlquad(Pb, Pa, Pd, Pc), gY) + imtrprod(tlquad(Ph, Pg, Pf, Pe), gXY2));
            T[8 * S + s] = (-2. / tau / tau * (r1 + r2)) / a5;
        }
        // T^{y eta}
        {
            double r1 = 1. / 4. * g * (retrprod(e2, p0 + m3(uy, pYm, Udy)) + retrprod(e2p, pXm + m3(UypX, pXYm, UDypX)));
            double r2 = 1. / 8. / g * (imtrprod(tlquad(Pa, Pb, Qd, Qc), gX) + imtrprod(tlquad(Qh, Qg, Qf, Qe), gXY1));
            T[9 * S + s] = (-2. / tau / tau * (r1 + r2)) / a5;
        }
    }
    double bs = block_sum(ttsum);
    if (threadIdx.x == 0) ev.partial[lb] = bs;
}

// Gauss-law violation max_x 1/2 sum |G|^2 (Evolution.cpp:295-364)
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
void ipg_launch_step(const EvDesc* d_evs, int nev, int nblk, int it, hipStream_t st) {
    hipLaunchKernelGGL(k_step, dim3(nblk), dim3(IPG_BLOCK), 0, st, d_evs, nev, it);
}
void ipg_launch_tmunu(const EvDesc* d_evs, int nev, int nblk, int it, hipStream_t st) {
    hipLaunchKernelGGL(k_tmunu, dim3(nblk), dim3(IPG_BLOCK), 0, st, d_evs, nev, it);
}
void ipg_launch_gauss(const EvDesc* d_evs, int nev, int nblk, int it, hipStream_t st) {
    hipLaunchKernelGGL(k_gauss, dim3(nblk), dim3(IPG_BLOCK), 0, st, d_evs, nev, it);
}
void ipg_launch_finish(const EvDesc* d_evs, int nev, int it, double* d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_finish, dim3(nev), dim3(IPG_BLOCK), 0, st, d_evs, nev, it, d_out);
}
