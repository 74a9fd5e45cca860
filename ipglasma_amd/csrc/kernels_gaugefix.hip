// kernels_gaugefix.hip -- Coulomb gauge fixing and the gluon spectra of
// Evolution::multiplicity, for one event.
//
//   k_gf_chi<N>       GaugeFix::FFTChi (GaugeFix.cpp:17-165), one row per
//                     block: div A = U_x - U_x(x-1) + U_y - U_y(y-1) with
//                     periodic neighbours, chi^a = Im tr(div A t^a), the row
//                     partial of the residual sum_a (chi^a)^2 / 2 / Nc, and
//                     the forward transform along y of chi packed as four
//                     complex rows (chi^{2p} + i chi^{2p+1})
//   k_gf_cols<N>      forward transform along x, times -1.5 / (k~^2 + 1e-9) / N^2,
//                     inverse transform along x
//   k_gf_rows_inv<N>  inverse transform along y, g = reu(exp(i chi'^a t^a))
//                     (Pade 6, Matrix.cpp:431-560; Matrix.h:67-90)
//   k_gf_transform    U_i(x) <- g(x) U_i(x) g(x+i)^+ (periodic), E, phi, pi
//                     <- g . g^+ (GaugeFix::gaugeTransform, GaugeFix.cpp:167-210)
//   k_mult_rows<N>,
//   k_mult_cols<N>    the 2D transform of a Hermitian field (E1, E2 or pi):
//                     its nine real planes packed in five complex fields
//   k_mult_accum      per k: Re tr(F(k) F(-k)) = sum_m w_m |F_m(k)|^2, the
//                     multiplicity n(k) = 2 / (omega N^2) c |...|, its energy,
//                     and the k_T histogram (Evolution.cpp:2454-2527)
//   k_gf_decide       the stopping rule of FFTChi on the device: the residual
//                     from the row partials (summed in row order), the value
//                     of every tenth step kept, stop below 1e-9 or when it
//                     rises while below 1e-6; the gauge-fixing kernels of a
//                     stopped iteration return at once, so the host launches
//                     several steps between reads of the state
//
// Stopping state gs (doubles): [0] stopped (0 or 1), [1] the step index,
// [2] the residual of the last tenth step, [3] the last residual.
//
// The reference transforms with fftshift around FFTW (FFT.cpp:492-610) and
// reads momenta in the shifted order; the kernels work in the unshifted
// order, where lattice momentum q and the shifted index i = q + N/2 (mod N)
// carry the same 4 sin^2(pi q / N), and -k of shifted index (N-i, N-j) is
// unshifted (-q mod N).  The shifted indices 0 (momentum N/2) have no
// partner in the reference and are left out here too.
#include "gaugefix.hpp"
#include "fft_lds.hpp"
#include "ipg_device.hpp"
#include "ipg_internal.hpp"

namespace {

__device__ __forceinline__ double wsum_gf(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
template <int NT>
__device__ double bsum_gf(double v) {
    __shared__ double w[NT / 64];
    v = wsum_gf(v);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0)
        for (int k = 0; k < NT / 64; k++) r += w[k];
    __syncthreads();
    return r;
}

template <int N>
__global__ __launch_bounds__(IPG_BLOCK) void k_gf_chi(const EvDesc* __restrict__ evs, FftPlanDev pl, int par,
                                                     double2* __restrict__ buf, double* __restrict__ part,
                                                     const double* __restrict__ gs) {
    __shared__ double2 lds[ipgfft::lds_slots(N, 4)];
    if (gs[0] != 0.) return;
    const EvDesc& ev = evs[0];
    const int i = blockIdx.x;
    const size_t S = ev.S;
    double res = 0.;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const size_t s = (size_t)i * N + j;
        const size_t mX = (size_t)((i + N - 1) % N) * N + j, mY = (size_t)i * N + (j + N - 1) % N;
        const M3 divA = ldm(ev.U[par][0], S, s) - ldm(ev.U[par][0], S, mX) + ldm(ev.U[par][1], S, s) -
                        ldm(ev.U[par][1], S, mY);
        cd t[8];
        tr_gen(divA, t);
        double chi[8];
#pragma unroll
        for (int a = 0; a < 8; a++) {
            chi[a] = t[a].i;
            res += chi[a] * chi[a] * 0.5 / 3.0;
        }
#pragma unroll
        for (int p = 0; p < 4; p++) lds[ipgfft::at(N, p, j)] = make_double2(chi[2 * p], chi[2 * p + 1]);
    }
    const double r = bsum_gf<IPG_BLOCK>(res);
    if (threadIdx.x == 0) part[i] = r;
    __syncthreads();
    ipgfft::lds_fft<N, 4, IPG_BLOCK, false>(lds, pl.tw);
    for (int q = threadIdx.x; q < 4 * N; q += blockDim.x) {
        const int p = q / N, j = q - p * N;
        buf[(size_t)p * S + (size_t)i * N + j] = lds[ipgfft::at(N, p, j)];
    }
}

// columns: NF packed complex planes; mode 0 multiplies by the gauge-fixing
// kernel and transforms back, mode 1 leaves the forward transform
template <int N>
__global__ __launch_bounds__(IPG_BLOCK) void k_gf_cols(const EvDesc* __restrict__ evs, FftPlanDev pl,
                                                      double2* __restrict__ buf, int nf, int mode,
                                                      const double* __restrict__ gs) {
    constexpr int C = ipgfft::cols_per_block(N);
    __shared__ double2 lds[ipgfft::lds_slots(N, C)];
    if (gs && gs[0] != 0.) return;
    constexpr int per_pack = N / C;
    const int p = blockIdx.x / per_pack;
    const int j0 = (blockIdx.x % per_pack) * C;
    const EvDesc& ev = evs[0];
    const size_t S = ev.S;
    if (p >= nf) return;
    double2* b = buf + (size_t)p * S;
    for (int q = threadIdx.x; q < C * N; q += blockDim.x) {
        const int r = q / C, c = q - r * C;
        lds[ipgfft::at(N, c, r)] = b[(size_t)r * N + j0 + c];
    }
    __syncthreads();
    ipgfft::lds_fft<N, C, IPG_BLOCK, false>(lds, pl.tw);
    if (mode == 0) {
        const double invN2 = 1.0 / ((double)N * (double)N);
        for (int q = threadIdx.x; q < C * N; q += blockDim.x) {
            const int c = q / N, r = q - c * N;
            const double kt2 = pl.s2[r] + pl.s2[j0 + c];
            const double f = -1.5 * (1. / (kt2 + 1e-9)) * invN2;
            const int a = ipgfft::at(N, c, r);
            lds[a] = make_double2(lds[a].x * f, lds[a].y * f);
        }
        __syncthreads();
        ipgfft::lds_fft<N, C, IPG_BLOCK, true>(lds, pl.tw);
    }
    for (int q = threadIdx.x; q < C * N; q += blockDim.x) {
        const int r = q / C, c = q - r * C;
        b[(size_t)r * N + j0 + c] = lds[ipgfft::at(N, c, r)];
    }
}

template <int N>
__global__ __launch_bounds__(IPG_BLOCK) void k_gf_rows_inv(const EvDesc* __restrict__ evs, FftPlanDev pl,
                                                          const double2* __restrict__ buf, double* __restrict__ G,
                                                          const double* __restrict__ gs) {
    __shared__ double2 lds[ipgfft::lds_slots(N, 4)];
    if (gs[0] != 0.) return;
    const EvDesc& ev = evs[0];
    const int i = blockIdx.x;
    const size_t S = ev.S;
    for (int q = threadIdx.x; q < 4 * N; q += blockDim.x) {
        const int p = q / N, j = q - p * N;
        lds[ipgfft::at(N, p, j)] = buf[(size_t)p * S + (size_t)i * N + j];
    }
    __syncthreads();
    ipgfft::lds_fft<N, 4, IPG_BLOCK, true>(lds, pl.tw);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        double chi[8];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const double2 v = lds[ipgfft::at(N, p, j)];
            chi[2 * p] = v.x;
            chi[2 * p + 1] = v.y;
        }
        stm(G, S, (size_t)i * N + j, ipggf::gauge_rotation(chi));
    }
}

__global__ __launch_bounds__(IPG_BLOCK) void k_gf_transform(const EvDesc* __restrict__ evs, int par,
                                                           const double* __restrict__ G,
                                                           const double* __restrict__ gs) {
    if (gs[0] != 0.) return;
    const EvDesc& ev = evs[0];
    // XCD-aware block order (grid padded to a multiple of 8): the blocks of
    // neighbouring rows share one XCD's L2 for the rotations of row i + 1
    const long long s = (long long)xcd_remap((int)blockIdx.x, (int)gridDim.x) * IPG_BLOCK + threadIdx.x;
    const size_t S = ev.S;
    if (s >= ev.S) return;
    const int N = ev.N;
    const int i = (int)(s / N), j = (int)(s - (long long)i * N);
    const size_t pX = (size_t)((i + 1) % N) * N + j, pY = (size_t)i * N + (j + 1) % N;
    const M3 g = ldm(G, S, s);
    stm(ev.U[par][0], S, s, mul_bd(g * ldm(ev.U[par][0], S, s), ldm(G, S, pX)));
    stm(ev.U[par][1], S, s, mul_bd(g * ldm(ev.U[par][1], S, s), ldm(G, S, pY)));
    sth(ev.E1, S, s, mul_bd(g * ldh(ev.E1, S, s), g));
    sth(ev.E2, S, s, mul_bd(g * ldh(ev.E2, S, s), g));
    sth(ev.phi[par], S, s, mul_bd(g * ldh(ev.phi[par], S, s), g));
    sth(ev.pi, S, s, mul_bd(g * ldh(ev.pi, S, s), g));
}

// one wavefront: GaugeFix.cpp:101-116 on the row partials of this step
// (lane l sums rows l, l + 64, ...; then a fixed shuffle tree)
__global__ void k_gf_decide(int N, const double* __restrict__ part, double* __restrict__ gs) {
    if (gs[0] != 0.) return;
    double v = 0.;
    for (int i = (int)threadIdx.x; i < N; i += 64) v += part[i];
    v = wsum_gf(v);
    if (threadIdx.x != 0) return;
    const double res = v / ((double)N * (double)N);
    const int k = (int)gs[1];
    double prev = gs[2];
    if (k % 10 == 0) prev = res;
    const bool stop = res < 1e-9 || (res > prev && res < 1e-6);
    gs[2] = prev;
    gs[3] = res;
    if (stop)
        gs[0] = 1.;
    else
        gs[1] = (double)(k + 1);
}
__global__ void k_gf_reset(double* __restrict__ gs) {
    if (threadIdx.x == 0) {
        gs[0] = 0.;
        gs[1] = 0.;
        gs[2] = 10000.;
        gs[3] = 0.;
    }
}

// Hermitian storage planes: 3 real diagonals, then (re, im) of the upper
// entries 01, 02, 12; tr(X Y) = sum_m w_m X_m Y_m with w = 1 on the
// diagonal and 2 off it.  Packed pairs: (d0, d1), (d2, 0), (01), (02), (12).
template <int N>
__global__ __launch_bounds__(IPG_BLOCK) void k_mult_rows(const EvDesc* __restrict__ evs, FftPlanDev pl,
                                                        const double* __restrict__ F, double2* __restrict__ buf) {
    __shared__ double2 lds[ipgfft::lds_slots(N, 5)];
    const EvDesc& ev = evs[0];
    const int i = blockIdx.x;
    const size_t S = ev.S;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const size_t s = (size_t)i * N + j;
        double h[9];
#pragma unroll
        for (int m = 0; m < 9; m++) h[m] = F[(size_t)m * S + s];
        lds[ipgfft::at(N, 0, j)] = make_double2(h[0], h[1]);
        lds[ipgfft::at(N, 1, j)] = make_double2(h[2], 0.);
        lds[ipgfft::at(N, 2, j)] = make_double2(h[3], h[4]);
        lds[ipgfft::at(N, 3, j)] = make_double2(h[5], h[6]);
        lds[ipgfft::at(N, 4, j)] = make_double2(h[7], h[8]);
    }
    __syncthreads();
    ipgfft::lds_fft<N, 5, IPG_BLOCK, false>(lds, pl.tw);
    for (int q = threadIdx.x; q < 5 * N; q += blockDim.x) {
        const int p = q / N, j = q - p * N;
        buf[(size_t)p * S + (size_t)i * N + j] = lds[ipgfft::at(N, p, j)];
    }
}

constexpr int kBins = 100;
// partial layout per block: [0] dN, [1] dE, [2..101] n[ik], [102..201] E[ik],
// [202..301] counts
constexpr int kMultSlots = 2 + 3 * kBins;

__global__ __launch_bounds__(IPG_BLOCK) void k_mult_accum(const EvDesc* __restrict__ evs, FftPlanDev pl,
                                                         const double2* __restrict__ buf, double pref, double ahbarc,
                                                         double* __restrict__ part) {
    __shared__ double hn[kBins], hE[kBins], hc[kBins];
    const EvDesc& ev = evs[0];
    const int N = ev.N;
    const size_t S = ev.S;
    for (int k = threadIdx.x; k < kBins; k += blockDim.x) hn[k] = hE[k] = hc[k] = 0.;
    __syncthreads();
    const long long s = (long long)blockIdx.x * IPG_BLOCK + threadIdx.x;
    double dn = 0., de = 0.;
    const double dkt = 2.83 / (double)kBins;
    if (s < ev.S) {
        const int qx = (int)(s / N), qy = (int)(s - (long long)qx * N);
        const double kt2 = pl.s2[qx] + pl.s2[qy];
        if (qx != N / 2 && qy != N / 2 && kt2 != 0.) {
            const size_t sm = (size_t)((N - qx) % N) * N + (size_t)((N - qy) % N);
            const double w[5] = {1., 1., 2., 2., 2.};
            double v = 0.;
#pragma unroll
            for (int p = 0; p < 5; p++) {
                const double2 a = buf[(size_t)p * S + s], b = buf[(size_t)p * S + sm];
                v += w[p] * 0.5 * (a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y);
            }
            const double om = sqrt(kt2);
            const double nkt = 2. / om / (double)((long long)N * N) * (pref * v);
            dn = nkt;
            de = nkt * om * ahbarc;
            const double r = sqrt(kt2);
            int ik = (int)(r / dkt);
            if (ik > 0 && !(r > ik * dkt)) ik--;
            if (r > (ik + 1) * dkt) ik++;
            if (ik >= 0 && ik < kBins && r > ik * dkt && r <= (ik + 1) * dkt) {
                const double fn = nkt / dkt / 2 / M_PI / sqrt(kt2) * 2 * M_PI * sqrt(kt2) * dkt * N * N / M_PI /
                                  M_PI / 2. / 2.;
                const double fe = sqrt(kt2) * ahbarc * nkt / dkt / 2 / M_PI / sqrt(kt2) * 2 * M_PI * sqrt(kt2) *
                                  dkt * N * N / M_PI / M_PI / 2. / 2.;
                atomicAdd(&hn[ik], fn);
                atomicAdd(&hE[ik], fe);
                atomicAdd(&hc[ik], 1.);
            }
        }
    }
    const double rn = bsum_gf<IPG_BLOCK>(dn);
    const double re = bsum_gf<IPG_BLOCK>(de);
    __syncthreads();
    double* o = part + (size_t)blockIdx.x * kMultSlots;
    if (threadIdx.x == 0) {
        o[0] = rn;
        o[1] = re;
    }
    for (int k = threadIdx.x; k < kBins; k += blockDim.x) {
        o[2 + k] = hn[k];
        o[2 + kBins + k] = hE[k];
        o[2 + 2 * kBins + k] = hc[k];
    }
}

template <int N>
void launch_gf_step_n(const EvDesc* d, int par, long long S, const FftPlanDev& pl, double2* buf, double* G,
                      double* part, double* gs, hipStream_t st) {
    constexpr int C = ipgfft::cols_per_block(N);
    hipLaunchKernelGGL(k_gf_chi<N>, dim3(N), dim3(IPG_BLOCK), 0, st, d, pl, par, buf, part, (const double*)gs);
    hipLaunchKernelGGL(k_gf_decide, dim3(1), dim3(64), 0, st, N, (const double*)part, gs);
    hipLaunchKernelGGL(k_gf_cols<N>, dim3(4 * (N / C)), dim3(IPG_BLOCK), 0, st, d, pl, buf, 4, 0,
                       (const double*)gs);
    hipLaunchKernelGGL(k_gf_rows_inv<N>, dim3(N), dim3(IPG_BLOCK), 0, st, d, pl, buf, G, (const double*)gs);
    hipLaunchKernelGGL(k_gf_transform, dim3((unsigned)(((S + IPG_BLOCK - 1) / IPG_BLOCK + 7) / 8 * 8)), dim3(IPG_BLOCK), 0, st,
                       d, par, (const double*)G, (const double*)gs);
}
template <int N>
void launch_mult_fft_n(const EvDesc* d, const FftPlanDev& pl, const double* F, double2* buf, hipStream_t st) {
    constexpr int C = ipgfft::cols_per_block(N);
    hipLaunchKernelGGL(k_mult_rows<N>, dim3(N), dim3(IPG_BLOCK), 0, st, d, pl, F, buf);
    hipLaunchKernelGGL(k_gf_cols<N>, dim3(5 * (N / C)), dim3(IPG_BLOCK), 0, st, d, pl, buf, 5, 1,
                       (const double*)nullptr);
}

#define IPG_FOR_SIZES(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

}  // namespace

int ipg_mult_slots() { return kMultSlots; }

void ipg_launch_gf_reset(double* gs, hipStream_t st) { hipLaunchKernelGGL(k_gf_reset, dim3(1), dim3(64), 0, st, gs); }
// one gauge-fixing step: chi with its row transform and residual partials,
// the stopping rule, the column pass, the inverse rows into the rotations G,
// the transform of the fields; every kernel returns at once once gs says
// stopped
void ipg_launch_gf_step(const EvDesc* d, int par, int N, const FftPlanDev& pl, double2* buf, double* G,
                        double* part, double* gs, hipStream_t st) {
    const long long S = (long long)N * N;
    switch (N) {
#define X(n) \
    case n: launch_gf_step_n<n>(d, par, S, pl, buf, G, part, gs, st); break;
        IPG_FOR_SIZES(X)
#undef X
        default: break;
    }
}
void ipg_launch_mult(const EvDesc* d, int N, const FftPlanDev& pl, const double* F, double2* buf, double pref,
                     double ahbarc, double* part, hipStream_t st) {
    switch (N) {
#define X(n) \
    case n: launch_mult_fft_n<n>(d, pl, F, buf, st); break;
        IPG_FOR_SIZES(X)
#undef X
        default: return;
    }
    const long long S = (long long)N * N;
    hipLaunchKernelGGL(k_mult_accum, dim3((unsigned)((S + IPG_BLOCK - 1) / IPG_BLOCK)), dim3(IPG_BLOCK), 0, st, d,
                       pl, buf, pref, ahbarc, part);
}
