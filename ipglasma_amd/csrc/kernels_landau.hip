// kernels_landau.hip -- Landau frame and spatial eccentricities at the end of
// the evolution.
//
//   k_landau        MyEigen::flowVelocity4D (MyEigen.cpp:25-321): per site the
//                   4x4 eigenproblem of T^mu_nu, eps, u^mu and pi^{mu nu}; block
//                   partials of the u^mu averages (MyEigen.cpp:199-206) and of
//                   the first eccentricity pass (Evolution.cpp:1346-1460)
//   k_landau_carry  pi^{mu nu} at sites whose eps or u^mu the reference takes
//                   from an earlier site (its running variables): a walk back
//                   through the site's block, then through per-block maxima
//   k_ecc_mid       centroid and the per-event sums of the first pass
//   k_ecc           moments about the centroid (Evolution.cpp:1478-1660) and
//                   the second centroid pass (1670-1790)
//   k_ecc_finish    eps_n, Psi_n, Rbar, sqrt<r^2>
//
// Built without floating-point contraction (see Makefile): the per-site Landau
// results are the bit pattern of the same code on the host.
#include "ipg_device.hpp"
#include "ipg_internal.hpp"
#include "landau.hpp"

namespace {

constexpr double kHbarc = 0.1973269718;
// Evolution.cpp:646 calls eccentricity(lat, param, it, 0.0, 0)
constexpr double kCutoff = 0.;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
// deterministic block sum (fixed tree), valid in thread 0
template <int NT>
__device__ double bsum(double v) {
    __shared__ double w[NT / 64];
    v = wsum(v);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0)
        for (int k = 0; k < NT / 64; k++) r += w[k];
    __syncthreads();
    return r;
}

// block maximum of an int, valid in thread 0
__device__ int bmax_int(int v) {
    __shared__ int w[IPG_BLOCK / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_down(v, o, 64));
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    int r = -1;
    if (threadIdx.x == 0)
        for (int k = 0; k < IPG_BLOCK / 64; k++) r = max(r, w[k]);
    __syncthreads();
    return r;
}

// partial slots of k_landau: 0-11 sums, 12/13 the last site of the block that
// set eps / u^mu (read by k_landau_carry before k_ecc reuses 12-30)
constexpr int kSlotEps = 12, kSlotU = 13;

__device__ __forceinline__ void load_T(const EvDesc& ev, long long s, double (&T)[10]) {
#pragma unroll
    for (int k = 0; k < 10; k++) T[k] = ev.T[k * ev.S + s];
}

__global__ __launch_bounds__(IPG_BLOCK) void k_landau(const EvDesc* __restrict__ evs, int nev, int it, int all) {
    const int e = find_event(evs, nev, blockIdx.x);
    const EvDesc& ev = evs[e];
    if (!ev.active || (!all && it != ev.itmax)) return;
    const int lb = xcd_remap(blockIdx.x - ev.block0, ev.nblocks);
    const long long s = (long long)lb * IPG_BLOCK + threadIdx.x;
    const long long S = ev.S;
    double acc[12];
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = 0.;
    int flags = 0;
    if (s < S) {
        const int N = ev.N;
        const int si = (int)(s / N), sj = (int)(s - (long long)si * N);
        double T[10];
        load_T(ev, s, T);
        ipgl::LandauSite o;
        ipgl::landau_site(T, N, ev.a, ev.dtau, it, si, sj, o);
        ev.land[0 * S + s] = o.eps;
#pragma unroll
        for (int k = 0; k < 4; k++) ev.land[(1 + k) * S + s] = o.u[k];
        ev.land_epsv[s] = o.eps_var;
        ev.land_flag[s] = o.flags;
        flags = o.flags;
        if ((o.flags & 3) == 3) {
            double pi[10];
            ipgl::landau_pi(T, o.eps_var, o.u[0], o.u[1], o.u[2], o.u[3], it, ev.dtau, ev.a, pi);
#pragma unroll
            for (int k = 0; k < 10; k++) ev.land[(5 + k) * S + s] = pi[k];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) acc[k] = o.acc[k];
        acc[4] = (o.flags & 4) ? 1. : 0.;
        acc[5] = (o.flags & 2) ? 0. : 1.;
        // first eccentricity pass: weight eps u^tau unless eps < cutoff
        const double L = ev.L, a = ev.a;
        const double x = -L / 2. + a * si, y = -L / 2. + a * sj;
        double w = 0.;
        if (!(o.eps < kCutoff)) {
            w = o.eps * o.u[0];
            acc[8] = 1.;
            acc[9] = o.eps * kHbarc;
            // <Q_sA^2 Q_sB^2> needs the colour-charge densities (0 without them)
            if (ev.g2A && ev.g2B)
                acc[10] = ev.g2A[s] * ev.QsmuRatio * ev.QsmuRatio * ev.g2B[s] * ev.QsmuRatioB * ev.QsmuRatioB / a /
                          a / a / a;
        }
        acc[6] = x * w;
        acc[7] = y * w;
        acc[11] = w;
    }
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const double r = bsum<IPG_BLOCK>(acc[k]);
        if (threadIdx.x == 0) ev.partial[(size_t)k * ev.nblocks + lb] = r;
    }
    const int le = bmax_int((flags & 1) ? (int)s : -1);
    const int lu = bmax_int((flags & 2) ? (int)s : -1);
    if (threadIdx.x == 0) {
        ev.partial[(size_t)kSlotEps * ev.nblocks + lb] = (double)le;
        ev.partial[(size_t)kSlotU * ev.nblocks + lb] = (double)lu;
    }
}

// last site q <= s whose flag has `bit` (-1: none)
__device__ long long last_set(const EvDesc& ev, long long s, int bit, int slot) {
    const long long b0 = (s / IPG_BLOCK) * IPG_BLOCK;
    for (long long q = s; q >= b0; q--)
        if (ev.land_flag[q] & bit) return q;
    for (long long b = s / IPG_BLOCK - 1; b >= 0; b--) {
        const double m = ev.partial[(size_t)slot * ev.nblocks + b];
        if (m >= 0.) return (long long)m;
    }
    return -1;
}

// pi^{mu nu} at sites without their own (eps, u^mu): the reference forms it
// from its running variables, last set at an earlier site in row-major order
__global__ __launch_bounds__(IPG_BLOCK) void k_landau_carry(const EvDesc* __restrict__ evs, int nev, int it,
                                                           int all) {
    const int e = find_event(evs, nev, blockIdx.x);
    const EvDesc& ev = evs[e];
    if (!ev.active || (!all && it != ev.itmax)) return;
    const long long s = (long long)xcd_remap(blockIdx.x - ev.block0, ev.nblocks) * IPG_BLOCK + threadIdx.x;
    const long long S = ev.S;
    if (s >= S) return;
    if ((ev.land_flag[s] & 3) == 3) return;
    const long long re = last_set(ev, s, 1, kSlotEps), ru = last_set(ev, s, 2, kSlotU);
    const double eps = re >= 0 ? ev.land_epsv[re] : 0.;
    double u[4] = {0., 0., 0., 0.};
    if (ru >= 0)
        for (int k = 0; k < 4; k++) u[k] = ev.land[(1 + k) * S + ru];
    double T[10], pi[10];
    load_T(ev, s, T);
    ipgl::landau_pi(T, eps, u[0], u[1], u[2], u[3], it, ev.dtau, ev.a, pi);
    for (int k = 0; k < 10; k++) ev.land[(5 + k) * S + s] = pi[k];
}

// per-event sum of partial slot k over the event's (unpadded) blocks
__device__ double slot_sum(const EvDesc& ev, int k) {
    const int nb = (int)((ev.S + IPG_BLOCK - 1) / IPG_BLOCK);
    double v = 0.;
    for (int l = threadIdx.x; l < nb; l += IPG_BLOCK) v += ev.partial[(size_t)k * ev.nblocks + l];
    return bsum<IPG_BLOCK>(v);
}

// obs layout: 0-5 eps_1..6, 6-11 Psi_1..6, 12 area, 13 Rbar, 14 <eps> hbarc,
// 15 <Q_sA^2 Q_sB^2> hbarc, 16 sqrt<r^2>, 17-18 centroid, 19-22 u^mu average
// sums, 23 QR failures, 24 sites without a timelike eigenvector, 25 sum w
__global__ __launch_bounds__(IPG_BLOCK) void k_ecc_mid(const EvDesc* __restrict__ evs, int nev, int it, int all) {
    const EvDesc& ev = evs[blockIdx.x];
    if (!ev.active || (!all && it != ev.itmax)) return;
    double r[12];
    for (int k = 0; k < 12; k++) r[k] = slot_sum(ev, k);
    if (threadIdx.x == 0) {
        double* o = ev.obs;
        for (int k = 0; k < 4; k++) o[19 + k] = r[k];
        o[23] = r[4];
        o[24] = r[5];
        const double tot = r[11];
        o[25] = tot;
        o[17] = r[6] / tot;
        o[18] = r[7] / tot;
        o[12] = r[8] * ev.a * ev.a;
        o[14] = r[9] / r[8];
        o[15] = r[10] / r[8] * kHbarc;
    }
}

__global__ __launch_bounds__(IPG_BLOCK) void k_ecc(const EvDesc* __restrict__ evs, int nev, int it, int all) {
    const int e = find_event(evs, nev, blockIdx.x);
    const EvDesc& ev = evs[e];
    if (!ev.active || (!all && it != ev.itmax)) return;
    const int lb = xcd_remap(blockIdx.x - ev.block0, ev.nblocks);
    const long long s = (long long)lb * IPG_BLOCK + threadIdx.x;
    const long long S = ev.S;
    double acc[19];
#pragma unroll
    for (int k = 0; k < 19; k++) acc[k] = 0.;
    if (s < S) {
        const int N = ev.N;
        const int si = (int)(s / N), sj = (int)(s - (long long)si * N);
        const double L = ev.L, a = ev.a;
        const double avx = ev.obs[17], avy = ev.obs[18];
        const double x = -L / 2. + a * si - avx, y = -L / 2. + a * sj - avy;
        const double eps = ev.land[s], ut = ev.land[S + s];
        const double w = eps < kCutoff ? 0. : eps * ut;
        acc[17] = x * x * w;
        acc[18] = y * y * w;
        if (si >= 2 && si < N - 2 && sj >= 2 && sj < N - 2) {
            // r^k e^{i n phi} = r^(k-n) z^n, z = x + i y
            const double r2 = x * x + y * y, r = sqrt(r2);
            double zr = x, zi = y;  // z^n
            double p[7][2];
            p[1][0] = x;
            p[1][1] = y;
#pragma unroll
            for (int n = 2; n <= 6; n++) {
                const double nr = zr * x - zi * y, ni = zr * y + zi * x;
                zr = nr;
                zi = ni;
                p[n][0] = zr;
                p[n][1] = zi;
            }
            acc[0] = r2 * r * w;        // r^3 (n = 1 and n = 3)
            acc[1] = r2 * w;            // r^2
            acc[2] = r2 * r2 * w;       // r^4
            acc[3] = r2 * r2 * r * w;   // r^5
            acc[4] = r2 * r2 * r2 * w;  // r^6
            acc[5] = r2 * p[1][0] * w;  // r^3 cos(phi)
            acc[6] = r2 * p[1][1] * w;
#pragma unroll
            for (int n = 2; n <= 6; n++) {
                acc[7 + 2 * (n - 2)] = p[n][0] * w;
                acc[8 + 2 * (n - 2)] = p[n][1] * w;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 19; k++) {
        const double r = bsum<IPG_BLOCK>(acc[k]);
        if (threadIdx.x == 0) ev.partial[(size_t)(12 + k) * ev.nblocks + lb] = r;
    }
}

__global__ __launch_bounds__(IPG_BLOCK) void k_ecc_finish(const EvDesc* __restrict__ evs, int nev, int it,
                                                         int all) {
    const EvDesc& ev = evs[blockIdx.x];
    if (!ev.active || (!all && it != ev.itmax)) return;
    double r[19];
    for (int k = 0; k < 19; k++) r[k] = slot_sum(ev, 12 + k);
    if (threadIdx.x == 0) {
        double* o = ev.obs;
        const double rn[7] = {0., r[0], r[1], r[0], r[2], r[3], r[4]};
        const double c[7] = {0., r[5], r[7], r[9], r[11], r[13], r[15]};
        const double sn[7] = {0., r[6], r[8], r[10], r[12], r[14], r[16]};
        for (int n = 1; n <= 6; n++) {
            o[n - 1] = sqrt(c[n] * c[n] + sn[n] * sn[n]) / rn[n];
            o[5 + n] = (atan(sn[n] / c[n]) + M_PI) / n;
        }
        const double tot = o[25];
        const double axx = r[17] / tot, ayy = r[18] / tot;
        o[13] = 1. / sqrt(1. / axx + 1. / ayy);
        o[16] = sqrt(r[1] / tot);
    }
}

// ---- eps_2 at itmax in reverse ------------------------------------------------------
// Pass 1: block sums of C, Sn, R, Mx, My (partial slots 0-4, one block per 256
// sites, after the forward eccentricity kernels have run).  Pass 2: per site,
// dL/dw -> (dL/d eps, dL/d u^tau) -> dL/dT (ipgl::landau_site_adj) into tb[10][S].
__global__ __launch_bounds__(IPG_BLOCK) void k_e2a_sums(const EvDesc* __restrict__ evs) {
    const EvDesc& ev = evs[0];
    const long long s = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    const long long S = ev.S;
    double acc[5] = {0., 0., 0., 0., 0.};
    if (s < S) {
        const int N = ev.N;
        const int si = (int)(s / N), sj = (int)(s - (long long)si * N);
        if (si >= 2 && si < N - 2 && sj >= 2 && sj < N - 2) {
            const double dx = -ev.L / 2. + ev.a * si - ev.obs[17], dy = -ev.L / 2. + ev.a * sj - ev.obs[18];
            const double eps = ev.land[s], ut = ev.land[S + s];
            const double w = eps < kCutoff ? 0. : eps * ut;
            acc[0] = (dx * dx - dy * dy) * w;
            acc[1] = 2. * dx * dy * w;
            acc[2] = (dx * dx + dy * dy) * w;
            acc[3] = dx * w;
            acc[4] = dy * w;
        }
    }
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const double r = bsum<IPG_BLOCK>(acc[k]);
        if (threadIdx.x == 0) ev.partial[(size_t)k * ev.nblocks + blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(IPG_BLOCK) void k_e2a_site(const EvDesc* __restrict__ evs, int it,
                                                       double* __restrict__ tb) {
    const EvDesc& ev = evs[0];
    // slot_sum is valid in thread 0: broadcast the five sums through LDS
    __shared__ double qs[5];
    for (int k = 0; k < 5; k++) {
        const double v = slot_sum(ev, k);
        if (threadIdx.x == 0) qs[k] = v;
    }
    __syncthreads();
    double q[5];
    for (int k = 0; k < 5; k++) q[k] = qs[k];
    const long long s = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    const long long S = ev.S;
    if (s >= S) return;
    const int N = ev.N;
    const int si = (int)(s / N), sj = (int)(s - (long long)si * N);
    const double dx = -ev.L / 2. + ev.a * si - ev.obs[17], dy = -ev.L / 2. + ev.a * sj - ev.obs[18];
    const bool in = si >= 2 && si < N - 2 && sj >= 2 && sj < N - 2;
    const double wb = ipgl::ecc2_wbar(q[0], q[1], q[2], q[3], q[4], ev.obs[25], dx, dy, in, 1.);
    const double eps = ev.land[s], ut = ev.land[S + s];
    double T[10], Tb[10];
    load_T(ev, s, T);
#pragma unroll
    for (int k = 0; k < 10; k++) Tb[k] = 0.;
    if (!(eps < kCutoff)) ipgl::landau_site_adj(T, N, ev.a, ev.dtau, it, si, sj, wb * ut, wb * eps, Tb);
#pragma unroll
    for (int k = 0; k < 10; k++) tb[(size_t)k * S + s] = Tb[k];
}

}  // namespace

void ipg_launch_landau(const EvDesc* d, int nev, int nblk, int it, int all, hipStream_t st) {
    hipLaunchKernelGGL(k_landau, dim3(nblk), dim3(IPG_BLOCK), 0, st, d, nev, it, all);
    hipLaunchKernelGGL(k_landau_carry, dim3(nblk), dim3(IPG_BLOCK), 0, st, d, nev, it, all);
    hipLaunchKernelGGL(k_ecc_mid, dim3(nev), dim3(IPG_BLOCK), 0, st, d, nev, it, all);
    hipLaunchKernelGGL(k_ecc, dim3(nblk), dim3(IPG_BLOCK), 0, st, d, nev, it, all);
    hipLaunchKernelGGL(k_ecc_finish, dim3(nev), dim3(IPG_BLOCK), 0, st, d, nev, it, all);
}

void ipg_adj_ecc2(const EvDesc* d, long long S, int it, double* tb, hipStream_t st) {
    const unsigned nb = (unsigned)((S + IPG_BLOCK - 1) / IPG_BLOCK);
    hipLaunchKernelGGL(k_e2a_sums, dim3(nb), dim3(IPG_BLOCK), 0, st, d);
    hipLaunchKernelGGL(k_e2a_site, dim3(nb), dim3(IPG_BLOCK), 0, st, d, it, tb);
}
