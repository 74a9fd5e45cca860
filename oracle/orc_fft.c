/* orc_fft.c -- oracle 2D DFT.  TEST INFRASTRUCTURE ONLY.
 *
 * Stands in for FFT::fftnComplex (FFT.cpp:492-610), which wraps an FFTW
 * 2D plan between two fftshifts and divides the inverse by N^2.  For even N
 * the shifted transform times a multiplier evaluated at k = 2pi(-1/2 + i/N)
 * equals the unshifted transform times the same multiplier at k = 2pi q/N
 * (the lattice momentum 4 sin^2(k/2) is 2pi-periodic), so the oracle uses
 * the plain unshifted DFT.  Textbook iterative radix-2 Cooley-Tukey for
 * powers of two; recursive mixed radix (2,3,5, naive DFT otherwise) else.
 * Pinned by numpy.fft.fft2 / ifft2 and a plane-wave eigenfunction test.
 */
#include <math.h>
#include <stdlib.h>

#include "ipg_oracle.h"
#include "orc_internal.h"

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* in-place radix-2 on contiguous x[n]; sign=+1 -> exp(-i ...) */
static void fft_pow2(int n, cplx *x, const cplx *tw /* n/2 forward */, int sign) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            cplx t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; k++) {
                cplx w = tw[k * step];
                if (sign < 0) w = conj(w);
                cplx a = x[i + k], b = x[i + k + half] * w;
                x[i + k] = a + b;
                x[i + k + half] = a - b;
            }
        }
    }
}

/* recursive mixed radix, out-of-place: y[k] = sum_j x[j*stride] w^{jk} */
static void fft_mixed(int n, const cplx *x, int stride, cplx *y, int sign) {
    if (n == 1) {
        y[0] = x[0];
        return;
    }
    int p = 0;
    for (int f = 2; f * f <= n; f++)
        if (n % f == 0) {
            p = f;
            break;
        }
    if (p == 0) p = n; /* prime: naive DFT */
    int m = n / p;
    cplx *sub = (cplx *)malloc(sizeof(cplx) * n);
    for (int r = 0; r < p; r++) fft_mixed(m, x + r * stride, stride * p, sub + r * m, sign);
    double s = sign > 0 ? -1.0 : 1.0;
    for (int k = 0; k < n; k++) {
        cplx acc = 0.0;
        for (int r = 0; r < p; r++) {
            double ang = s * 2.0 * M_PI * (double)((long)r * k % n) / (double)n;
            acc += sub[r * m + (k % m)] * CMPLX(cos(ang), sin(ang));
        }
        y[k] = acc;
    }
    free(sub);
}

static void fft1d(int n, cplx *x, const cplx *tw, cplx *tmp, int sign) {
    if (is_pow2(n)) {
        fft_pow2(n, x, tw, sign);
    } else {
        fft_mixed(n, x, 1, tmp, sign);
        for (int i = 0; i < n; i++) x[i] = tmp[i];
    }
}

int orc_fft2d(int N, double *data, int sign) {
    cplx *d = (cplx *)data; /* interleaved re,im has the layout of cplx */
    cplx *tw = (cplx *)malloc(sizeof(cplx) * (N / 2 + 1));
    for (int k = 0; k <= N / 2; k++) {
        double ang = -2.0 * M_PI * (double)k / (double)N;
        tw[k] = CMPLX(cos(ang), sin(ang));
    }
#pragma omp parallel
    {
        cplx *col = (cplx *)malloc(sizeof(cplx) * N);
        cplx *tmp = (cplx *)malloc(sizeof(cplx) * N);
#pragma omp for
        for (int i = 0; i < N; i++) fft1d(N, d + (size_t)i * N, tw, tmp, sign);
#pragma omp for
        for (int j = 0; j < N; j++) {
            for (int i = 0; i < N; i++) col[i] = d[(size_t)i * N + j];
            fft1d(N, col, tw, tmp, sign);
            for (int i = 0; i < N; i++) d[(size_t)i * N + j] = col[i];
        }
        free(col);
        free(tmp);
    }
    if (sign < 0) {
        double inv = (double)N * (double)N;
        for (size_t i = 0; i < (size_t)N * N; i++) d[i] = CMPLX(creal(d[i]) / inv, cimag(d[i]) / inv);
    }
    free(tw);
    return 0;
}
