// fft_lds.hpp -- length-N complex transforms of rows held in LDS.
//
// Stockham auto-sort formulation: every stage reads its R inputs of one
// butterfly at stride N/R, twiddles them, takes an R-point DFT in registers
// and writes the outputs at stride Ns (the length already transformed).
// Natural order in and out, so no bit reversal, and any N = 2^a 3^b works:
// the plan takes the radix-3 stages first, then radix 8 (with one or two
// radix-4 stages, or one radix-2 stage, for the leftover power of two).
// N = 512 is three radix-8 passes instead of nine radix-2 passes.
//
// Forward = exp(-2 pi i jk/N); inverse = exp(+2 pi i jk/N), unnormalised.
// The twiddle table holds exp(-2 pi i m/N) for m < N.
//
// The per-butterfly functions are host+device so that a host build of the
// same code can be checked against numpy.
#pragma once
#include <hip/hip_runtime.h>

#define IPGF_HD __host__ __device__ inline

namespace ipgfft {

// ---- compile-time plan ------------------------------------------------------------
IPGF_HD constexpr int count3(int n) {
    int c = 0;
    while (n % 3 == 0) {
        n /= 3;
        c++;
    }
    return c;
}
IPGF_HD constexpr int log2c(int n) {
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}
IPGF_HD constexpr int pow3(int b) {
    int p = 1;
    for (int i = 0; i < b; i++) p *= 3;
    return p;
}
// the radices of the power-of-two part 2^m in stage order
IPGF_HD constexpr int twos_count(int m) {
    if (m == 0) return 0;
    if (m % 3 == 0) return m / 3;
    if (m % 3 == 2) return m / 3 + 1;
    if (m == 1) return 1;
    return (m - 4) / 3 + 2;
}
IPGF_HD constexpr int twos_radix(int m, int i) {
    if (m % 3 == 0) return 8;
    if (m % 3 == 2) return i < m / 3 ? 8 : 4;
    if (m == 1) return 2;
    return i < (m - 4) / 3 ? 8 : 4;
}
IPGF_HD constexpr int nstages(int N) {
    const int b = count3(N);
    return b + twos_count(log2c(N / pow3(b)));
}
IPGF_HD constexpr int radix(int N, int s) {
    const int b = count3(N);
    return s < b ? 3 : twos_radix(log2c(N / pow3(b)), s - b);
}
// length already transformed before stage s
IPGF_HD constexpr int stride_ns(int N, int s) {
    int p = 1;
    for (int i = 0; i < s; i++) p *= radix(N, i);
    return p;
}
// supported lengths: 2^a 3^b with b <= 1, 16 <= N <= 1024
IPGF_HD constexpr bool supported(int N) {
    if (N < 16 || N > 1024) return false;
    int n = N;
    if (n % 3 == 0) n /= 3;
    return (n & (n - 1)) == 0;
}

// columns per block of the column pass: the largest divisor of N with
// C N complex values in 64 KiB of LDS
IPGF_HD constexpr int cols_per_block(int N) {
    int c = 4096 / N < N ? 4096 / N : N;
    while (c > 1 && N % c != 0) c--;
    return c;
}

// ---- butterflies ------------------------------------------------------------------
IPGF_HD double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
IPGF_HD double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
IPGF_HD double2 cmulf(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
IPGF_HD double2 rot90(double2 a) {
    return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
    IPGF_HD static void run(double2 (&v)[2]) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV>
struct Dft<3, INV> {
    IPGF_HD static void run(double2 (&v)[3]) {
        const double c = 0.86602540378443864676;  // sqrt(3)/2
        const double2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        const double2 m = make_double2(v[0].x - 0.5 * s.x, v[0].y - 0.5 * s.y);
        const double2 t = rot90<INV>(make_double2(c * d.x, c * d.y));
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, t);
        v[2] = csub(m, t);
    }
};

template <bool INV>
struct Dft<4, INV> {
    IPGF_HD static void run(double2 (&v)[4]) {
        const double2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        const double2 c = cadd(v[1], v[3]), d = rot90<INV>(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, d);
        v[3] = csub(b, d);
    }
};

template <bool INV>
struct Dft<8, INV> {
    IPGF_HD static void run(double2 (&v)[8]) {
        const double h = 0.70710678118654752440;  // 1/sqrt(2)
        double2 a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            a[k] = cadd(v[k], v[k + 4]);
            b[k] = csub(v[k], v[k + 4]);
        }
        // b_k *= w^k, w = exp(-+ 2 pi i / 8)
        b[1] = INV ? make_double2(h * (b[1].x - b[1].y), h * (b[1].x + b[1].y))
                   : make_double2(h * (b[1].x + b[1].y), h * (b[1].y - b[1].x));
        b[2] = rot90<INV>(b[2]);
        b[3] = INV ? make_double2(-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y))
                   : make_double2(h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y));
        Dft<4, INV>::run(a);
        Dft<4, INV>::run(b);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[2 * k] = a[k];
            v[2 * k + 1] = b[k];
        }
    }
};

// butterfly j (0 <= j < N/R) of the stage with Ns already transformed:
// inputs at j + r N/R, outputs at (j / Ns) Ns R + j % Ns + q Ns
template <int N, int R, int NS, bool INV>
IPGF_HD void butterfly(double2 (&v)[R], int j, const double2* __restrict__ tw) {
    if (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; r++) {
            double2 w = tw[r * k * (N / (NS * R))];
            if (INV) w.y = -w.y;
            v[r] = cmulf(v[r], w);
        }
    }
    Dft<R, INV>::run(v);
}
template <int N, int R, int NS>
IPGF_HD int out_index(int j) {
    return (j / NS) * NS * R + j % NS;
}

// ---- LDS layout -------------------------------------------------------------------
// Element i of row `row` sits at row * row_stride(N) + pad(i) (double2 slots).
// pad() skips one slot in every 8, so that the first Stockham stage's writes
// at stride R = 8 or 4 (and the second stage's runs of 8 that are 64 apart)
// fall in distinct banks; the row stride is 9 mod 16 slots, so that the
// column pass's transposed accesses (one lane per row) do too.  Unpadded,
// about half of the LDS cycles of the transform kernels were bank conflicts.
IPGF_HD constexpr int pad(int i) { return i + (i >> 3); }
IPGF_HD constexpr int row_stride(int N) {
    const int b = pad(N - 1) + 1;
    return b + ((9 - b % 16) + 16) % 16;
}
// LDS slots of NROW rows
IPGF_HD constexpr int lds_slots(int N, int NROW) { return NROW * row_stride(N); }
IPGF_HD constexpr int at(int N, int row, int i) { return row * row_stride(N) + pad(i); }

// ---- block-cooperative transform of NROW rows of length N in LDS ---------------
// (device only; BLK threads; element i of row `row` at x[at(N, row, i)]).
// WAVE: the rows belong to one wavefront (BLK = 64, lane = threadIdx.x % 64)
// and the stages synchronise that wavefront only, so the waves of a block
// run their own rows without waiting for each other.
template <int N, int NROW, int BLK, bool INV, int S, bool WAVE = false>
__device__ __forceinline__ void stages(double2* x, const double2* __restrict__ tw) {
    if constexpr (S < nstages(N)) {
        constexpr int R = radix(N, S);
        constexpr int NS = stride_ns(N, S);
        constexpr int NB = N / R;
        constexpr int TOT = NROW * NB;
        constexpr int K = (TOT + BLK - 1) / BLK;
        constexpr int RS = row_stride(N);
        static_assert(!WAVE || BLK == 64, "a wavefront transform runs on 64 lanes");
        const int tid = WAVE ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
        double2 v[K][R];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int q = tid + k * BLK;
            if (q < TOT) {
                const int row = q / NB, j = q - row * NB;
#pragma unroll
                for (int r = 0; r < R; r++) v[k][r] = x[row * RS + pad(j + r * NB)];
            }
        }
        if constexpr (WAVE)
            __syncwarp();
        else
            __syncthreads();
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int q = tid + k * BLK;
            if (q < TOT) {
                const int row = q / NB, j = q - row * NB;
                butterfly<N, R, NS, INV>(v[k], j, tw);
                const int o = out_index<N, R, NS>(j);
#pragma unroll
                for (int r = 0; r < R; r++) x[row * RS + pad(o + r * NS)] = v[k][r];
            }
        }
        if constexpr (WAVE)
            __syncwarp();
        else
            __syncthreads();
        stages<N, NROW, BLK, INV, S + 1, WAVE>(x, tw);
    }
}

template <int N, int NROW, int BLK, bool INV>
__device__ __forceinline__ void lds_fft(double2* x, const double2* __restrict__ tw) {
    stages<N, NROW, BLK, INV, 0>(x, tw);
}
// NROW rows starting at x (row stride row_stride(N)) transformed by the
// calling wavefront alone
template <int N, int NROW, bool INV>
__device__ __forceinline__ void wave_fft(double2* x, const double2* __restrict__ tw) {
    stages<N, NROW, 64, INV, 0, true>(x, tw);
}

}  // namespace ipgfft
