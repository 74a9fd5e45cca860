// su3.hpp -- 3x3 complex FP64 algebra for the CDNA4 kernels.
//
// Matrices live in registers as 9 (re, im) pairs, row-major e[3*i+j]
// (Matrix.h:45).  In HBM a field is structure-of-arrays: 18 planes of S
// doubles, plane 2k = Re e[k], plane 2k+1 = Im e[k], so that a wavefront of
// 64 consecutive sites reads 512 contiguous bytes per plane.
//
// The operations follow Matrix.cpp:100-306 (product, conjg, prodABconj,
// prodAconjB) and the closed-form exponential Matrix::expmCoeff
// (Matrix.cpp:310-426, Morningstar & Peardon, PRD 69 (2004) 054501) with
// Init::getUfromExponent (Init.cpp:3483-3499).
#pragma once
#include <hip/hip_runtime.h>

#define IPG_HD __host__ __device__ __forceinline__

struct cd {
    double r, i;
};

IPG_HD cd cmk(double r, double i) { return cd{r, i}; }
IPG_HD cd operator+(cd a, cd b) { return cd{a.r + b.r, a.i + b.i}; }
IPG_HD cd operator-(cd a, cd b) { return cd{a.r - b.r, a.i - b.i}; }
IPG_HD cd operator-(cd a) { return cd{-a.r, -a.i}; }
IPG_HD cd operator*(cd a, cd b) { return cd{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
IPG_HD cd operator*(cd a, double s) { return cd{a.r * s, a.i * s}; }
IPG_HD cd operator*(double s, cd a) { return cd{a.r * s, a.i * s}; }
IPG_HD cd cconj(cd a) { return cd{a.r, -a.i}; }
// a * conj(b)
IPG_HD cd mulc(cd a, cd b) { return cd{a.r * b.r + a.i * b.i, a.i * b.r - a.r * b.i}; }
// conj(a) * b
IPG_HD cd cmul(cd a, cd b) { return cd{a.r * b.r + a.i * b.i, a.r * b.i - a.i * b.r}; }
IPG_HD cd cdiv(cd a, cd b) {
    double d = b.r * b.r + b.i * b.i;
    return cd{(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}

struct M3 {
    cd e[9];
};

IPG_HD M3 m_zero() {
    M3 m;
#pragma unroll
    for (int k = 0; k < 9; k++) m.e[k] = cd{0., 0.};
    return m;
}
IPG_HD M3 m_one() {
    M3 m = m_zero();
    m.e[0].r = m.e[4].r = m.e[8].r = 1.;
    return m;
}
IPG_HD M3 operator*(const M3& a, const M3& b) {
    M3 c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            c.e[3 * i + j] = a.e[3 * i] * b.e[j] + a.e[3 * i + 1] * b.e[3 + j] + a.e[3 * i + 2] * b.e[6 + j];
    return c;
}
// a * b^dagger
IPG_HD M3 mul_bd(const M3& a, const M3& b) {
    M3 c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            c.e[3 * i + j] = mulc(a.e[3 * i], b.e[3 * j]) + mulc(a.e[3 * i + 1], b.e[3 * j + 1]) +
                             mulc(a.e[3 * i + 2], b.e[3 * j + 2]);
    return c;
}
// a^dagger * b
IPG_HD M3 mul_ad(const M3& a, const M3& b) {
    M3 c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            c.e[3 * i + j] = cmul(a.e[i], b.e[j]) + cmul(a.e[3 + i], b.e[3 + j]) + cmul(a.e[6 + i], b.e[6 + j]);
    return c;
}
IPG_HD M3 dag(const M3& a) {
    M3 c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) c.e[3 * j + i] = cconj(a.e[3 * i + j]);
    return c;
}
IPG_HD M3 operator+(const M3& a, const M3& b) {
    M3 c;
#pragma unroll
    for (int k = 0; k < 9; k++) c.e[k] = a.e[k] + b.e[k];
    return c;
}
IPG_HD M3 operator-(const M3& a, const M3& b) {
    M3 c;
#pragma unroll
    for (int k = 0; k < 9; k++) c.e[k] = a.e[k] - b.e[k];
    return c;
}
IPG_HD M3 operator*(double s, const M3& a) {
    M3 c;
#pragma unroll
    for (int k = 0; k < 9; k++) c.e[k] = a.e[k] * s;
    return c;
}
// multiply by i*s
IPG_HD M3 mul_i(double s, const M3& a) {
    M3 c;
#pragma unroll
    for (int k = 0; k < 9; k++) c.e[k] = cd{-a.e[k].i * s, a.e[k].r * s};
    return c;
}
IPG_HD cd tr(const M3& a) { return a.e[0] + a.e[4] + a.e[8]; }
// tr(a*b)
IPG_HD cd trprod(const M3& a, const M3& b) {
    cd t{0., 0.};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) t = t + a.e[3 * i + j] * b.e[3 * j + i];
    return t;
}
// Re tr(a*b)
IPG_HD double retrprod(const M3& a, const M3& b) {
    double t = 0.;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) t += a.e[3 * i + j].r * b.e[3 * j + i].r - a.e[3 * i + j].i * b.e[3 * j + i].i;
    return t;
}
IPG_HD double imtrprod(const M3& a, const M3& b) {
    double t = 0.;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) t += a.e[3 * i + j].r * b.e[3 * j + i].i + a.e[3 * i + j].i * b.e[3 * j + i].r;
    return t;
}
// a - tr(a)/3
IPG_HD M3 traceless(M3 a) {
    cd t = tr(a);
    cd t3{t.r / 3.0, t.i / 3.0};
    a.e[0] = a.e[0] - t3;
    a.e[4] = a.e[4] - t3;
    a.e[8] = a.e[8] - t3;
    return a;
}
// 0.5 * sum |e|^2  (Matrix.h:172-178)
IPG_HD double square(const M3& a) {
    double t = 0.;
#pragma unroll
    for (int k = 0; k < 9; k++) t += a.e[k].r * a.e[k].r + a.e[k].i * a.e[k].i;
    return 0.5 * t;
}

// ---- SoA field access ------------------------------------------------------
// Field pointers are cast to the global address space so that each plane is
// addressed as (uniform plane base in SGPRs) + (32-bit lane offset): one
// offset register per site instead of a 64-bit address per plane.  With
// Opaque the offset passes through an empty asm statement: left visible, the
// compiler may widen it and fold it into 64-bit per-plane addresses in
// VGPRs, which in the reverse-mode kernels costs 50-100 registers per lane
// or spills (the forward kernels are faster with it visible: neighbour
// offsets fold into the instructions' immediate offsets).
template <bool Opaque>
__device__ __forceinline__ unsigned lane_offset(size_t s) {
    unsigned ob = (unsigned)s * 8u;
    if (Opaque) asm volatile("" : "+v"(ob));
    return ob;
}
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ const gdouble* gp(const double* p) { return (const gdouble*)p; }
__device__ __forceinline__ gdouble* gp(double* p) { return (gdouble*)p; }
// plane k at byte offset ob (< 4 GiB, one plane of at most 2^29 sites)
__device__ __forceinline__ double gld(const double* f, size_t plane, unsigned ob) {
    return *(const gdouble*)((const gchar*)gp(f + plane) + ob);
}
__device__ __forceinline__ void gst(double* f, size_t plane, unsigned ob, double v) {
    *(gdouble*)((gchar*)gp(f + plane) + ob) = v;
}

template <bool Opaque = false>
__device__ __forceinline__ M3 ldm(const double* f, size_t S, size_t s) {
    const unsigned ob = lane_offset<Opaque>(s);
    M3 m;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        m.e[k].r = gld(f, (2 * k) * S, ob);
        m.e[k].i = gld(f, (2 * k + 1) * S, ob);
    }
    return m;
}
template <bool Opaque = false>
__device__ __forceinline__ void stm(double* f, size_t S, size_t s, const M3& m) {
    const unsigned ob = lane_offset<Opaque>(s);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        gst(f, (2 * k) * S, ob, m.e[k].r);
        gst(f, (2 * k + 1) * S, ob, m.e[k].i);
    }
}

// Hermitian fields (E^i, phi, pi) are stored as 9 planes: the real diagonal
// h00 h11 h22, then Re/Im of h01, h02, h12.  They are Hermitian by
// construction (Evolution.cpp:245-290 builds E from i(X - X^+) and
// commutators of Hermitian matrices; pi = -(2i/g) A^eta with A^eta
// anti-Hermitian, Init.cpp:2669-2673); storing the upper triangle drops only
// the round-off-level anti-Hermitian part.  pi and phi keep their trace.
template <bool Opaque = false>
__device__ __forceinline__ M3 ldh(const double* f, size_t S, size_t s0) {
    const unsigned s = lane_offset<Opaque>(s0);
    M3 m;
    const double r01 = gld(f, 3 * S, s), i01 = gld(f, 4 * S, s);
    const double r02 = gld(f, 5 * S, s), i02 = gld(f, 6 * S, s);
    const double r12 = gld(f, 7 * S, s), i12 = gld(f, 8 * S, s);
    m.e[0] = cd{gld(f, 0, s), 0.};
    m.e[4] = cd{gld(f, S, s), 0.};
    m.e[8] = cd{gld(f, 2 * S, s), 0.};
    m.e[1] = cd{r01, i01};
    m.e[3] = cd{r01, -i01};
    m.e[2] = cd{r02, i02};
    m.e[6] = cd{r02, -i02};
    m.e[5] = cd{r12, i12};
    m.e[7] = cd{r12, -i12};
    return m;
}
template <bool Opaque = false>
__device__ __forceinline__ void sth(double* f, size_t S, size_t s0, const M3& m) {
    const unsigned s = lane_offset<Opaque>(s0);
    gst(f, 0, s, m.e[0].r);
    gst(f, S, s, m.e[4].r);
    gst(f, 2 * S, s, m.e[8].r);
    gst(f, 3 * S, s, m.e[1].r);
    gst(f, 4 * S, s, m.e[1].i);
    gst(f, 5 * S, s, m.e[2].r);
    gst(f, 6 * S, s, m.e[2].i);
    gst(f, 7 * S, s, m.e[5].r);
    gst(f, 8 * S, s, m.e[5].i);
}
// the stored (upper-triangle) form of a matrix, expanded again
IPG_HD M3 herm_part(const M3& m) {
    M3 h;
    h.e[0] = cd{m.e[0].r, 0.};
    h.e[4] = cd{m.e[4].r, 0.};
    h.e[8] = cd{m.e[8].r, 0.};
    h.e[1] = m.e[1];
    h.e[2] = m.e[2];
    h.e[5] = m.e[5];
    h.e[3] = cconj(m.e[1]);
    h.e[6] = cconj(m.e[2]);
    h.e[7] = cconj(m.e[5]);
    return h;
}

// ---- SU(3) exponential -------------------------------------------------------
// Coefficients (u0, u_a) of exp(i Q^a t^a) = u0 1 + u_a t^a; NaN -> 0.
IPG_HD void expm_coeff(const double Q[8], cd out[9]) {
    const double r3 = sqrt(3.0);
    double c0 = r3 * (Q[0] * Q[0] * Q[7] + Q[1] * Q[1] * Q[7] + Q[2] * Q[2] * Q[7]);
    c0 -= Q[7] * Q[7] * Q[7] / r3;
    c0 -= (r3 / 2.0) * (Q[3] * Q[3] * Q[7] + Q[4] * Q[4] * Q[7] + Q[5] * Q[5] * Q[7] + Q[6] * Q[6] * Q[7]);
    c0 += 3.0 * (Q[0] * Q[3] * Q[5] + Q[0] * Q[4] * Q[6] + Q[1] * Q[4] * Q[5] - Q[1] * Q[3] * Q[6]);
    c0 += 1.5 * (Q[2] * Q[3] * Q[3] + Q[2] * Q[4] * Q[4] - Q[2] * Q[5] * Q[5] - Q[2] * Q[6] * Q[6]);
    c0 /= 12.0;
    double c1 = 0.;
#pragma unroll
    for (int a = 0; a < 8; a++) c1 += Q[a] * Q[a];
    c1 *= 0.25;
    // pow(c1/3, 1.5) as x*sqrt(x): two correctly rounded operations, so the
    // degenerate case |c0/c0max| ~ 1 (acos -> NaN) is decided as on the host
    const double x3 = c1 / 3.0;
    double c0max = fmax(1e-15, 2.0 * (x3 * sqrt(x3)));
    double th3 = acos(c0 / c0max) / 3.0;
    double sth, cth, sw, cw;
    sincos(th3, &sth, &cth);
    double u = sqrt(c1 / 3.0) * cth;
    double w = sqrt(c1) * sth;
    sincos(w, &sw, &cw);
    double xi0 = sw / w;
    double den = 9.0 * u * u - w * w;
    double s2u, c2u, su, cu;
    sincos(2.0 * u, &s2u, &c2u);
    sincos(u, &su, &cu);
    cd e2iu{c2u, s2u};
    cd emiu{cu, -su};
    // f0 = (u^2-w^2) e^{2iu} + e^{-iu} (8u^2 cos w + 2iu xi0 (3u^2+w^2))
    cd f0 = e2iu * (u * u - w * w) + emiu * cd{8.0 * u * u * cw, 2.0 * u * xi0 * (3.0 * u * u + w * w)};
    f0 = cd{f0.r / den, f0.i / den};
    // f1 = 2u e^{2iu} - e^{-iu} (2u cos w - i (3u^2 - w^2) xi0)
    cd f1 = e2iu * (2.0 * u) - emiu * cd{2.0 * u * cw, -(3.0 * u * u - w * w) * xi0};
    f1 = cd{f1.r / den, f1.i / den};
    // f2 = e^{2iu} - e^{-iu} (cos w + 3 i u xi0)
    cd f2 = e2iu - emiu * cd{cw, 3.0 * u * xi0};
    f2 = cd{f2.r / den, f2.i / den};
    cd u0 = f0 + f2 * (2.0 / 3.0 * c1);
    cd hf2 = f2 * 0.5;
    f1 = cdiv(f1, hf2);
    double d[8];
    d[0] = Q[3] * Q[5] + Q[4] * Q[6] + 2.0 / r3 * Q[0] * Q[7];
    d[1] = 2.0 * Q[1] * Q[7] / r3 - Q[3] * Q[6] + Q[4] * Q[5];
    d[2] = 2.0 * Q[2] * Q[7] / r3 + 0.5 * Q[3] * Q[3] + 0.5 * Q[4] * Q[4] - 0.5 * Q[5] * Q[5] - 0.5 * Q[6] * Q[6];
    d[3] = -1.0 / r3 * Q[3] * Q[7] + Q[0] * Q[5] - Q[1] * Q[6] + Q[2] * Q[3];
    d[4] = -1.0 / r3 * Q[4] * Q[7] + Q[0] * Q[6] + Q[1] * Q[5] + Q[2] * Q[4];
    d[5] = -1.0 / r3 * Q[5] * Q[7] + Q[0] * Q[3] + Q[1] * Q[4] - Q[2] * Q[5];
    d[6] = -1.0 / r3 * Q[6] * Q[7] + Q[0] * Q[4] - Q[1] * Q[3] - Q[2] * Q[6];
    d[7] = (Q[0] * Q[0] + Q[1] * Q[1] + Q[2] * Q[2] - Q[7] * Q[7] - 0.5 * Q[3] * Q[3] - 0.5 * Q[4] * Q[4] -
            0.5 * Q[5] * Q[5] - 0.5 * Q[6] * Q[6]) /
           r3;
    out[0] = u0;
#pragma unroll
    for (int a = 0; a < 8; a++) {
        cd ua = f1 * Q[a] + cd{d[a], 0.};
        out[a + 1] = ua * hf2;
    }
#pragma unroll
    for (int k = 0; k < 9; k++)
        if (out[k].r != out[k].r || out[k].i != out[k].i) out[k] = cd{0., 0.};
}

// exp(i Q^a t^a); the identity when |Re u0| < 1e-15 (Init.cpp:3488-3489)
IPG_HD M3 exp_iQt(const double Q[8]) {
    cd u[9];
    expm_coeff(Q, u);
    if (fabs(u[0].r) < 1e-15) return m_one();
    const double s3 = 1.0 / (2.0 * sqrt(3.0));
    const double t8 = -1.0 / sqrt(3.0);
    M3 U;
    U.e[0] = (u[0] + u[3] * 0.5) + u[8] * s3;
    U.e[1] = cd{0.5 * u[1].r + 0.5 * u[2].i, 0.5 * u[1].i - 0.5 * u[2].r};
    U.e[2] = cd{0.5 * u[4].r + 0.5 * u[5].i, 0.5 * u[4].i - 0.5 * u[5].r};
    U.e[3] = cd{0.5 * u[1].r - 0.5 * u[2].i, 0.5 * u[1].i + 0.5 * u[2].r};
    U.e[4] = (u[0] - u[3] * 0.5) + u[8] * s3;
    U.e[5] = cd{0.5 * u[6].r + 0.5 * u[7].i, 0.5 * u[6].i - 0.5 * u[7].r};
    U.e[6] = cd{0.5 * u[4].r - 0.5 * u[5].i, 0.5 * u[4].i + 0.5 * u[5].r};
    U.e[7] = cd{0.5 * u[6].r - 0.5 * u[7].i, 0.5 * u[6].i + 0.5 * u[7].r};
    U.e[8] = u[0] + u[8] * t8;
    return U;
}

// t^a (Group.cpp:68-149) applied on the left: (t^a M)
IPG_HD M3 gen_mul(int a, const M3& M) {
    M3 c = m_zero();
    const double s3 = 1.0 / (2.0 * sqrt(3.0));
    switch (a) {
        case 0:
            for (int j = 0; j < 3; j++) { c.e[j] = M.e[3 + j] * 0.5; c.e[3 + j] = M.e[j] * 0.5; }
            break;
        case 1:
            for (int j = 0; j < 3; j++) {
                c.e[j] = cd{0.5 * M.e[3 + j].i, -0.5 * M.e[3 + j].r};
                c.e[3 + j] = cd{-0.5 * M.e[j].i, 0.5 * M.e[j].r};
            }
            break;
        case 2:
            for (int j = 0; j < 3; j++) { c.e[j] = M.e[j] * 0.5; c.e[3 + j] = M.e[3 + j] * -0.5; }
            break;
        case 3:
            for (int j = 0; j < 3; j++) { c.e[j] = M.e[6 + j] * 0.5; c.e[6 + j] = M.e[j] * 0.5; }
            break;
        case 4:
            for (int j = 0; j < 3; j++) {
                c.e[j] = cd{0.5 * M.e[6 + j].i, -0.5 * M.e[6 + j].r};
                c.e[6 + j] = cd{-0.5 * M.e[j].i, 0.5 * M.e[j].r};
            }
            break;
        case 5:
            for (int j = 0; j < 3; j++) { c.e[3 + j] = M.e[6 + j] * 0.5; c.e[6 + j] = M.e[3 + j] * 0.5; }
            break;
        case 6:
            for (int j = 0; j < 3; j++) {
                c.e[3 + j] = cd{0.5 * M.e[6 + j].i, -0.5 * M.e[6 + j].r};
                c.e[6 + j] = cd{-0.5 * M.e[3 + j].i, 0.5 * M.e[3 + j].r};
            }
            break;
        default:
            for (int j = 0; j < 3; j++) {
                c.e[j] = M.e[j] * s3;
                c.e[3 + j] = M.e[3 + j] * s3;
                c.e[6 + j] = M.e[6 + j] * (-2.0 * s3);
            }
            break;
    }
    return c;
}

// tr(t^a M): the traces of the generators against M, all eight at once
IPG_HD void tr_gen(const M3& M, cd t[8]) {
    const double s3 = 1.0 / (2.0 * sqrt(3.0));
    // t0: 0.5(M10 + M01)
    t[0] = (M.e[3] + M.e[1]) * 0.5;
    // t1: (-0.5i) M10 + (0.5i) M01
    t[1] = cd{0.5 * (M.e[3].i - M.e[1].i), 0.5 * (M.e[1].r - M.e[3].r)};
    t[2] = (M.e[0] - M.e[4]) * 0.5;
    t[3] = (M.e[6] + M.e[2]) * 0.5;
    t[4] = cd{0.5 * (M.e[6].i - M.e[2].i), 0.5 * (M.e[2].r - M.e[6].r)};
    t[5] = (M.e[7] + M.e[5]) * 0.5;
    t[6] = cd{0.5 * (M.e[7].i - M.e[5].i), 0.5 * (M.e[5].r - M.e[7].r)};
    t[7] = (M.e[0] + M.e[4]) * s3 + M.e[8] * (-2.0 * s3);
}
