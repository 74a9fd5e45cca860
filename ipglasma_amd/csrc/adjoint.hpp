// adjoint.hpp -- reverse mode of one leapfrog step and of T^tautau at one
// site (Evolution.cpp:35-293 and 754-905), host+device.
//
// Convention: for a real loss L and a complex matrix Z the cotangent is
// Zb = dL/dRe Z + i dL/dIm Z, so that dL = Re tr(Zb^+ dZ).  Rules used:
//   Y = A B      ->  Ab += Yb B^+,  Bb += A^+ Yb
//   Y = A^+      ->  Ab += Yb^+
//   Y = c A      ->  Ab += conj(c) Yb
//   Y = TL(A)    ->  Ab += TL(Yb)              (traceless part: self-adjoint)
//   y = Re tr(Z Z) -> Zb += 2 yb Z^+
// Links are general complex matrices (the second-order update is not
// unitary); E, phi and pi are Hermitian, and their cotangents are projected
// on the Hermitian matrices where they are stored.
//
// Forward step at tau (lattice units), the reference's order:
//   pi'  = pi + (dtau/tau) [Ux phi(x+x) Ux^+ + Ux(x-x)^+ phi(x-x) Ux(x-x)
//                           + (same in y) - 4 phi]
//   E1'  = TL(E1 + c1 TL(U12 + U1m2 - h.c.) + c2 [Ux phi(x+x) Ux^+, phi])
//   E2'  = TL(E2 + c1 TL(U12^+ + U2m1 - h.c.) + c2 [Uy phi(x+y) Uy^+, phi])
//   phi' = phi + (tau + dtau/2) dtau pi'
//   U'   = (1 + X + X^2/2) U,  X = i g^2 dtau / (tau + dtau/2) E'
// with c1 = i tau dtau / (2 g^2), c2 = i dtau / tau.
#pragma once
#include "su3.hpp"

namespace ipgad {

IPG_HD M3 hproj(const M3& z) { return 0.5 * (z + dag(z)); }
IPG_HD M3 cscale(cd c, const M3& a) {
    M3 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.e[k] = c * a.e[k];
    return r;
}

// Y = A F A^+ (F Hermitian):  Ab += (Yb + Yb^+) A F,  Fb += A^+ Yb A
IPG_HD void adj_conj(const M3& A, const M3& F, const M3& Yb, M3& Ab, M3& Fb) {
    Ab = Ab + (Yb + dag(Yb)) * (A * F);
    Fb = Fb + mul_ad(A, Yb) * A;
}
// Y = V^+ G V (G Hermitian):  Vb += G V (Yb + Yb^+),  Gb += V Yb V^+
IPG_HD void adj_conjd(const M3& V, const M3& G, const M3& Yb, M3& Vb, M3& Gb) {
    Vb = Vb + (G * V) * (Yb + dag(Yb));
    Gb = Gb + mul_bd(V * Yb, V);
}
// Y = M1 M2 M3 M4: cotangents of the four factors (each product is formed
// just before its last use)
IPG_HD void adj_prod4(const M3& M1, const M3& M2, const M3& M3_, const M3& M4, const M3& Yb, M3& B1, M3& B2,
                      M3& B3, M3& B4) {
    const M3 P34 = M3_ * M4;
    B1 = mul_bd(Yb, M2 * P34);               // Yb (M2 M3 M4)^+
    B2 = mul_bd(mul_ad(M1, Yb), P34);        // M1^+ Yb (M3 M4)^+
    const M3 P12 = M1 * M2;
    B3 = mul_bd(mul_ad(P12, Yb), M4);        // (M1 M2)^+ Yb M4^+
    B4 = mul_ad(P12 * M3_, Yb);              // (M1 M2 M3)^+ Yb
}

// ---- pi update -------------------------------------------------------------------
// inputs of site s: ux = Ux(s), uy = Uy(s), uxm = Ux(s-x), uym = Uy(s-y),
// phi at s, s+x, s-x, s+y, s-y; cp = dtau/tau
struct PiIn {
    M3 ux, uy, uxm, uym, ph, phpx, phmx, phpy, phmy, pi;
};
IPG_HD M3 pi_fwd(const PiIn& q, double cp) {
    const M3 br = (((q.ux * mul_bd(q.phpx, q.ux) + mul_ad(q.uxm, q.phmx) * q.uxm) + q.uy * mul_bd(q.phpy, q.uy)) +
                   mul_ad(q.uym, q.phmy) * q.uym) -
                  4. * q.ph;
    return q.pi + cp * br;
}
struct PiBar {
    M3 ux, uy, uxm, uym, ph, phpx, phmx, phpy, phmy, pi;
};
IPG_HD void pi_adj(const PiIn& q, double cp, const M3& Yb, PiBar& b) {
    b.pi = b.pi + Yb;
    const M3 Bb = cp * Yb;
    adj_conj(q.ux, q.phpx, Bb, b.ux, b.phpx);
    adj_conjd(q.uxm, q.phmx, Bb, b.uxm, b.phmx);
    adj_conj(q.uy, q.phpy, Bb, b.uy, b.phpy);
    adj_conjd(q.uym, q.phmy, Bb, b.uym, b.phmy);
    b.ph = b.ph - 4. * Bb;
}

// ---- E update -----------------------------------------------------------------------
// links: ux, uy at s; uy_px = Uy(s+x), ux_py = Ux(s+y), uy_pxmy = Uy(s+x-y),
// ux_my = Ux(s-y), uy_my = Uy(s-y), ux_mxpy = Ux(s-x+y), uy_mx = Uy(s-x),
// ux_mx = Ux(s-x); phi at s, s+x, s+y
struct EIn {
    M3 ux, uy, uy_px, ux_py, uy_pxmy, ux_my, uy_my, ux_mxpy, uy_mx, ux_mx, ph, phpx, phpy, e1, e2;
};
struct EBar {
    M3 ux, uy, uy_px, ux_py, uy_pxmy, ux_my, uy_my, ux_mxpy, uy_mx, ux_mx, ph, phpx, phpy, e1, e2;
};
IPG_HD void e_fwd(const EIn& q, double tau, double dtau, double g, M3& e1n, M3& e2n) {
    const cd c1{0., tau * dtau / (2. * g * g)}, c2{0., dtau / tau};
    const M3 U12 = (q.ux * q.uy_px) * mul_bd(dag(q.ux_py), q.uy);
    const M3 U1m2 = mul_bd(q.ux, q.uy_pxmy) * mul_ad(q.ux_my, q.uy_my);
    const M3 U2m1 = mul_bd(q.uy, q.ux_mxpy) * mul_ad(q.uy_mx, q.ux_mx);
    const M3 tx = q.ux * mul_bd(q.phpx, q.ux), ty = q.uy * mul_bd(q.phpy, q.uy);
    M3 t1 = traceless(((U12 + U1m2) - dag(U12)) - dag(U1m2));
    M3 t2 = tx * q.ph - q.ph * tx;
    e1n = traceless(q.e1 + (cscale(c1, t1) + cscale(c2, t2)));
    t1 = traceless(((dag(U12) + U2m1) - U12) - dag(U2m1));
    t2 = ty * q.ph - q.ph * ty;
    e2n = traceless(q.e2 + (cscale(c1, t1) + cscale(c2, t2)));
}
// the E' update in reverse, in pieces: Z1, Z2 = TL(E1'b), TL(E2'b) give the
// cotangents of the three plaquette products U12, U1m2, U2m1 ...
IPG_HD void plaq_bars(const M3& Z1, const M3& Z2, double tau, double dtau, double g, M3& U12b, M3& U1m2b,
                      M3& U2m1b) {
    const cd c1c{0., -tau * dtau / (2. * g * g)};  // conj(c1)
    const M3 t1b = traceless(cscale(c1c, Z1)), t1pb = traceless(cscale(c1c, Z2));
    // t1 = U12 + U1m2 - U12^+ - U1m2^+, t1' = U12^+ + U2m1 - U12 - U2m1^+
    U12b = (t1b - dag(t1b)) + (dag(t1pb) - t1pb);
    U1m2b = t1b - dag(t1b);
    U2m1b = t1pb - dag(t1pb);
}
// ... and, with t2b = conj(c2) Z, the commutator t2 = tx phi - phi tx,
// tx = U phn U^+ (tx, phi Hermitian): cotangents of phi, U and phn
IPG_HD void comm_adj(const M3& U, const M3& phn, const M3& ph, const M3& t2b, M3& phb, M3& Ub, M3& phnb) {
    const M3 tx = U * mul_bd(phn, U);
    const M3 txb = t2b * ph - ph * t2b;
    phb = phb + (tx * t2b - t2b * tx);
    adj_conj(U, phn, txb, Ub, phnb);
}
IPG_HD void e_adj(const EIn& q, double tau, double dtau, double g, const M3& Y1b, const M3& Y2b, EBar& b) {
    const cd c2c{0., -dtau / tau};  // conj(c2)
    const M3 Z1 = traceless(Y1b), Z2 = traceless(Y2b);
    b.e1 = b.e1 + Z1;
    b.e2 = b.e2 + Z2;
    M3 U12b, U1m2b, U2m1b;
    plaq_bars(Z1, Z2, tau, dtau, g, U12b, U1m2b, U2m1b);
    M3 B1, B2, B3, B4;
    // U12 = ux uy_px ux_py^+ uy^+
    adj_prod4(q.ux, q.uy_px, dag(q.ux_py), dag(q.uy), U12b, B1, B2, B3, B4);
    b.ux = b.ux + B1;
    b.uy_px = b.uy_px + B2;
    b.ux_py = b.ux_py + dag(B3);
    b.uy = b.uy + dag(B4);
    // U1m2 = ux uy_pxmy^+ ux_my^+ uy_my
    adj_prod4(q.ux, dag(q.uy_pxmy), dag(q.ux_my), q.uy_my, U1m2b, B1, B2, B3, B4);
    b.ux = b.ux + B1;
    b.uy_pxmy = b.uy_pxmy + dag(B2);
    b.ux_my = b.ux_my + dag(B3);
    b.uy_my = b.uy_my + B4;
    // U2m1 = uy ux_mxpy^+ uy_mx^+ ux_mx
    adj_prod4(q.uy, dag(q.ux_mxpy), dag(q.uy_mx), q.ux_mx, U2m1b, B1, B2, B3, B4);
    b.uy = b.uy + B1;
    b.ux_mxpy = b.ux_mxpy + dag(B2);
    b.uy_mx = b.uy_mx + dag(B3);
    b.ux_mx = b.ux_mx + B4;
    comm_adj(q.ux, q.phpx, q.ph, cscale(c2c, Z1), b.ph, b.ux, b.phpx);
    comm_adj(q.uy, q.phpy, q.ph, cscale(c2c, Z2), b.ph, b.uy, b.phpy);
}

// ---- link update U' = (1 + X + X^2/2) U, X = i kappa E' ---------------------------
IPG_HD M3 taylor_m(const M3& X) { return m_one() + X * (m_one() + 0.5 * X); }
IPG_HD M3 u_fwd(const M3& U, const M3& E, double kappa) { return taylor_m(cscale(cd{0., kappa}, E)) * U; }
// Ub += M^+ Yb,  Eb += -i kappa (Mb + (Mb X^+ + X^+ Mb)/2),  Mb = Yb U^+
IPG_HD void u_adj(const M3& U, const M3& E, double kappa, const M3& Yb, M3& Ub, M3& Eb) {
    const M3 X = cscale(cd{0., kappa}, E);
    const M3 M = taylor_m(X);
    Ub = Ub + mul_ad(M, Yb);
    const M3 Mb = mul_bd(Yb, U);
    const M3 Xd = dag(X);
    const M3 Xb = Mb + 0.5 * (Mb * Xd + Xd * Mb);
    Eb = Eb + cscale(cd{0., -kappa}, Xb);
}

// ---- T^tautau at one site (lattice units, before the a^-4) ----------------------------
// e1, e1py = E1(s), E1(s+y); e2, e2px = E2(s), E2(s+x); pi at s, s+x, s+y,
// s+x+y; links for the plaquette and the phi gradients; phi at s, s+x, s+y,
// s+x+y; plaq = 0 on the last row (the reference sets U_plaq = 1 there)
struct TIn {
    M3 e1, e1py, e2, e2px, p, ppx, ppy, ppxy, ux, uy, uy_px, ux_py, ph, phpx, phpy, phpxy;
    int plaq;
};
struct TBar {
    M3 e1, e1py, e2, e2px, p, ppx, ppy, ppxy, ux, uy, uy_px, ux_py, ph, phpx, phpy, phpxy;
};
IPG_HD double tt_fwd(const TIn& q, double tau, double g) {
    const double gt = g * g / tau / tau;
    const double el = retrprod(q.e1, q.e1) + retrprod(q.e1py, q.e1py) + retrprod(q.e2, q.e2) +
                      retrprod(q.e2px, q.e2px);
    const double lng = (retrprod(q.p, q.p) + retrprod(q.ppx, q.ppx) + retrprod(q.ppy, q.ppy) +
                        retrprod(q.ppxy, q.ppxy)) /
                       4.;
    double rp = 3.;
    if (q.plaq) {
        const M3 P = (q.ux * q.uy_px) * mul_bd(dag(q.ux_py), q.uy);
        rp = tr(P).r;
    }
    const M3 dA = q.ph - q.ux * mul_bd(q.phpx, q.ux);
    const M3 dB = q.phpy - q.ux_py * mul_bd(q.phpxy, q.ux_py);
    const M3 dC = q.ph - q.uy * mul_bd(q.phpy, q.uy);
    const M3 dD = q.phpx - q.uy_px * mul_bd(q.phpxy, q.uy_px);
    const double h = 0.5 / tau / tau;
    return gt * el / 2. + lng + 2. / (g * g) * (3. - rp) +
           h * (retrprod(dA, dA) + retrprod(dB, dB) + retrprod(dC, dC) + retrprod(dD, dD));
}
IPG_HD void tt_adj(const TIn& q, double tau, double g, double yb, TBar& b) {
    const double gt = g * g / tau / tau;
    b.e1 = b.e1 + (yb * gt) * dag(q.e1);
    b.e1py = b.e1py + (yb * gt) * dag(q.e1py);
    b.e2 = b.e2 + (yb * gt) * dag(q.e2);
    b.e2px = b.e2px + (yb * gt) * dag(q.e2px);
    b.p = b.p + (yb * 0.5) * dag(q.p);
    b.ppx = b.ppx + (yb * 0.5) * dag(q.ppx);
    b.ppy = b.ppy + (yb * 0.5) * dag(q.ppy);
    b.ppxy = b.ppxy + (yb * 0.5) * dag(q.ppxy);
    if (q.plaq) {
        // d/dP of -2/g^2 Re tr P: Pb = -2/g^2 yb 1
        const M3 Pb = (-2. / (g * g) * yb) * m_one();
        M3 B1, B2, B3, B4;
        adj_prod4(q.ux, q.uy_px, dag(q.ux_py), dag(q.uy), Pb, B1, B2, B3, B4);
        b.ux = b.ux + B1;
        b.uy_px = b.uy_px + B2;
        b.ux_py = b.ux_py + dag(B3);
        b.uy = b.uy + dag(B4);
    }
    const double h2 = yb * 2. * (0.5 / tau / tau);  // d Re tr(d d) = 2 Re tr(d^+^+ dd)
    {
        const M3 d = q.ph - q.ux * mul_bd(q.phpx, q.ux);
        const M3 db = h2 * dag(d);
        b.ph = b.ph + db;
        M3 nb = -1. * db;
        adj_conj(q.ux, q.phpx, nb, b.ux, b.phpx);
    }
    {
        const M3 d = q.phpy - q.ux_py * mul_bd(q.phpxy, q.ux_py);
        const M3 db = h2 * dag(d);
        b.phpy = b.phpy + db;
        M3 nb = -1. * db;
        adj_conj(q.ux_py, q.phpxy, nb, b.ux_py, b.phpxy);
    }
    {
        const M3 d = q.ph - q.uy * mul_bd(q.phpy, q.uy);
        const M3 db = h2 * dag(d);
        b.ph = b.ph + db;
        M3 nb = -1. * db;
        adj_conj(q.uy, q.phpy, nb, b.uy, b.phpy);
    }
    {
        const M3 d = q.phpx - q.uy_px * mul_bd(q.phpxy, q.uy_px);
        const M3 db = h2 * dag(d);
        b.phpx = b.phpx + db;
        M3 nb = -1. * db;
        adj_conj(q.uy_px, q.phpxy, nb, b.uy_px, b.phpxy);
    }
}

// ---- tau = 0 root finder by implicit differentiation (Init.cpp:3344-3481) ----
// The solution U of F(U; W) = 0, F_a = -Im tr t^a [(W - W^+) + (W U^+ - U W^+)],
// lies on the coset SU(3) U0 (U = exp(i alpha t) U0, U0 = U1 U2, W = U1 + U2).
// Near it, U(beta) = exp(i beta t) E* U0 with E* = U U0^{-1}; then
//   dF_a/dbeta_b = Re tr(t^a W U^+ t^b) + Re tr(t^a t^b U W^+),
//   dU = i (dbeta . t) U + E* dU0,
// and the reverse pass is  betab_b = Re tr(Ub^+ i t^b U),
// lambda = -J^{-T} betab,  T = sum_a lambda_a t^a,
//   Wb = -2i T (1 + U) ... (from F_lambda = Re tr(i T M)),  Usb = 2i T W,
//   U0b = E*^+ (Ub + Usb),  U1b = Wb + U0b U2^+,  U2b = Wb + U1^+ U0b.
// E* = U U0^+ (U0 is unitary: products of Wilson lines).
IPG_HD void tgen(int a, M3& t) {
    t = m_zero();
    const double s3 = 1.0 / (2.0 * sqrt(3.0));
    switch (a) {
        case 0: t.e[1] = cd{0.5, 0.}; t.e[3] = cd{0.5, 0.}; break;
        case 1: t.e[1] = cd{0., -0.5}; t.e[3] = cd{0., 0.5}; break;
        case 2: t.e[0] = cd{0.5, 0.}; t.e[4] = cd{-0.5, 0.}; break;
        case 3: t.e[2] = cd{0.5, 0.}; t.e[6] = cd{0.5, 0.}; break;
        case 4: t.e[2] = cd{0., -0.5}; t.e[6] = cd{0., 0.5}; break;
        case 5: t.e[5] = cd{0.5, 0.}; t.e[7] = cd{0.5, 0.}; break;
        case 6: t.e[5] = cd{0., -0.5}; t.e[7] = cd{0., 0.5}; break;
        default: t.e[0] = cd{s3, 0.}; t.e[4] = cd{s3, 0.}; t.e[8] = cd{-2. * s3, 0.}; break;
    }
}
// solve A x = b (8x8, partial pivoting), A row-major, overwritten.  The
// loops are unrolled and the pivot rows exchanged by selects, so that A and
// b stay in registers (an exchange through a run-time row index puts them
// in scratch memory on the GPU)
IPG_HD void solve8(double (&A)[64], double (&b)[8]) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
        int p = j;
        double mx = fabs(A[8 * j + j]);
#pragma unroll
        for (int r = j + 1; r < 8; r++)
            if (fabs(A[8 * r + j]) > mx) {
                mx = fabs(A[8 * r + j]);
                p = r;
            }
#pragma unroll
        for (int r = j + 1; r < 8; r++) {
            const bool sw = r == p;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const double x = A[8 * j + c], y = A[8 * r + c];
                A[8 * j + c] = sw ? y : x;
                A[8 * r + c] = sw ? x : y;
            }
            const double x = b[j], y = b[r];
            b[j] = sw ? y : x;
            b[r] = sw ? x : y;
        }
#pragma unroll
        for (int r = j + 1; r < 8; r++) {
            const double f = A[8 * r + j] / A[8 * j + j];
#pragma unroll
            for (int c = j; c < 8; c++) A[8 * r + c] -= f * A[8 * j + c];
            b[r] -= f * b[j];
        }
    }
#pragma unroll
    for (int r = 7; r >= 0; r--) {
        double t = b[r];
#pragma unroll
        for (int c = r + 1; c < 8; c++) t -= A[8 * r + c] * b[c];
        b[r] = t / A[8 * r + r];
    }
}
IPG_HD void newton_adj(const M3& U1, const M3& U2, const M3& U, const M3& Ub, M3& U1b, M3& U2b) {
    const M3 W = U1 + U2, U0 = U1 * U2;
    const M3 A = mul_bd(W, U), B = mul_bd(U, W);  // W U^+, U W^+
    // unrolled: the generators are constants, their zeros fold away
    M3 t[8];
#pragma unroll
    for (int a = 0; a < 8; a++) tgen(a, t[a]);
    double JT[64], rhs[8];  // J^T: JT[8 b + a] = J_ab
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const M3 ta_A = t[a] * A;
#pragma unroll
        for (int b = 0; b < 8; b++) JT[8 * b + a] = retrprod(ta_A, t[b]) + retrprod(t[a] * t[b], B);
    }
    // betab_b = Re tr(Ub^+ i t^b U); solve J^T lambda = -betab
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const M3 d = mul_i(1., t[b] * U);
        rhs[b] = -retrprod(dag(Ub), d);
    }
    solve8(JT, rhs);
    M3 T = m_zero();
#pragma unroll
    for (int a = 0; a < 8; a++) T = T + rhs[a] * t[a];
    const M3 iT = mul_i(1., T);
    const M3 Wb = (-2.) * (iT * (m_one() + U));
    const M3 Usb = 2. * (iT * W);
    const M3 Es = mul_bd(U, U0);  // E* = U U0^+
    const M3 U0b = mul_ad(Es, Ub + Usb);
    U1b = U1b + Wb + mul_bd(U0b, U2);
    U2b = U2b + Wb + mul_ad(U1, U0b);
}

// ---- pi at tau = 0: pi = -(i / 4g) acc (Init.cpp:2521-2584, 2669-2673) --------------
// x part of acc at site s from D = L1x(s) - L2x(s), U = Ux(s) and the site s-x:
//   (D U^+ - D - U D^+ + D^+) + (-U'^+ D' + D' + D'^+ U' - D'^+)
IPG_HD M3 acc_dir(const M3& D, const M3& U, const M3& Dm, const M3& Um) {
    const M3 Dd = dag(D), Dmd = dag(Dm);
    M3 acc = ((mul_bd(D, U) - D) - U * Dd) + Dd;
    return (((acc - mul_ad(Um, Dm)) + Dm) + Dmd * Um) - Dmd;
}
IPG_HD void acc_dir_adj(const M3& D, const M3& U, const M3& Dm, const M3& Um, const M3& ab, M3& Db, M3& Ub,
                        M3& Dmb, M3& Umb) {
    const M3 abd = dag(ab);
    Db = Db + ab * U - ab - abd * U + abd;
    Ub = Ub + abd * D - ab * D;
    Dmb = Dmb - Um * ab + ab + Um * abd - abd;
    Umb = Umb - Dm * abd + Dm * ab;
}

// ---- E = exp(i Q^a t^a) (Matrix.cpp:310-426): Qb_a = Re tr(Eb^+ dE/dQ_a) -------------
// By Cayley-Hamilton, exp(iQ) = f0 + f1 Q + f2 Q^2 for traceless Hermitian Q,
// the f_j being functions of c0 = det Q and c1 = tr(Q^2)/2 only
// (Morningstar and Peardon, Phys. Rev. D 69 (2004) 054501: eigenvalues 2u,
// -u+w, -u-w).  Then
//   Qb_a = Re tr(G t^a),  G = Re(b0) Q^2 + Re(b1) Q + f1 Eb^+ + f2 (Q Eb^+ + Eb^+ Q),
//   b_i = sum_j (d f_j / d c_i) tr(Eb^+ Q^j),
// from d c0 / d Q_a = tr(Q^2 t^a) and d c1 / d Q_a = tr(Q t^a).

// xi0(w) = sin(w)/w and xi1(w) = (cos w - xi0)/w^2 (Taylor series near 0)
IPG_HD void expq_xi(double w, double& x0, double& x1) {
    const double w2 = w * w;
    if (fabs(w) < 0.05) {
        x0 = 1. - w2 / 6. * (1. - w2 / 20. * (1. - w2 / 42.));
        x1 = -1. / 3. + w2 * (1. / 30. - w2 * (1. / 840. - w2 / 45360.));
    } else {
        double sw, cw;
        sincos(w, &sw, &cw);
        x0 = sw / w;
        x1 = (cw - x0) / w2;
    }
}

// f_j(c0, c1) and their partial derivatives fa_j = d/dc0, fb_j = d/dc1.
// Small c1 (all |eigenvalues| <= sqrt(4 c1 / 3) < 0.82): the exponential
// series reduced with Q^3 = c1 Q + c0; otherwise the closed form at |c0|,
// with f_j(-c0, c1) = (-1)^j conj(f_j(c0, c1)) for negative c0, which keeps
// 9u^2 - w^2 >= 2 c1 away from zero.
IPG_HD void expq_fcoef(double c0, double c1, cd (&f)[3], cd (&fa)[3], cd (&fb)[3]) {
    if (c1 < 0.5) {
        // Q^n = a I + b Q + d Q^2; (a, b, d) <- (c0 d, a + c1 d, b)
        double a = 1., b = 0., d = 0.;
        double a0 = 0., b0 = 0., d0 = 0., a1 = 0., b1 = 0., d1 = 0.;  // d/dc0, d/dc1
        double fr[3] = {0., 0., 0.}, fi[3] = {0., 0., 0.};
        double ar[3] = {0., 0., 0.}, ai[3] = {0., 0., 0.}, br[3] = {0., 0., 0.}, bi[3] = {0., 0., 0.};
        double coef = 1.;  // 1/n!
        for (int n = 0; n < 26; n++) {
            // i^n / n! = coef * (1, i, -1, -i)[n mod 4]
            const int m = n & 3;
            const double cr = m == 0 ? coef : (m == 2 ? -coef : 0.);
            const double ci = m == 1 ? coef : (m == 3 ? -coef : 0.);
            const double v[3] = {a, b, d}, v0[3] = {a0, b0, d0}, v1[3] = {a1, b1, d1};
#pragma unroll
            for (int j = 0; j < 3; j++) {
                fr[j] += cr * v[j];
                fi[j] += ci * v[j];
                ar[j] += cr * v0[j];
                ai[j] += ci * v0[j];
                br[j] += cr * v1[j];
                bi[j] += ci * v1[j];
            }
            // the recursion's matrix has 1-norm max(1, |c0| + c1) = 1 here
            // (c1 < 0.5), and its derivatives add |d| to a0 and b1, so the
            // sum s of |v|, |v0|, |v1| at most triples per term while 1/n!
            // falls at least 5-fold from n = 4: once coef s < 1e-19 the terms
            // left add less than 1.5e-19 (|f_j| is of order 1)
            const double sz = fabs(a) + fabs(b) + fabs(d) + fabs(a0) + fabs(b0) + fabs(d0) + fabs(a1) + fabs(b1) +
                              fabs(d1);
            if (n >= 4 && coef * sz < 1e-19) break;
            const double na = c0 * d, nb = a + c1 * d, nd = b;
            const double na0 = d + c0 * d0, nb0 = a0 + c1 * d0, nd0 = b0;
            const double na1 = c0 * d1, nb1 = a1 + d + c1 * d1, nd1 = b1;
            a = na, b = nb, d = nd;
            a0 = na0, b0 = nb0, d0 = nd0;
            a1 = na1, b1 = nb1, d1 = nd1;
            coef /= (double)(n + 1);
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            f[j] = cd{fr[j], fi[j]};
            fa[j] = cd{ar[j], ai[j]};
            fb[j] = cd{br[j], bi[j]};
        }
        return;
    }
    const bool neg = c0 < 0.;
    const double c0p = fabs(c0);
    const double x3 = c1 / 3.;
    const double c0max = 2. * x3 * sqrt(x3);
    const double th = acos(fmin(c0p / c0max, 1.)) / 3.;
    double sth, cth;
    sincos(th, &sth, &cth);
    const double u = sqrt(x3) * cth, w = sqrt(c1) * sth;
    double x0, x1;
    expq_xi(w, x0, x1);
    const double cw = cos(w), u2 = u * u, w2 = w * w;
    double s2u, c2u, su, cu;
    sincos(2. * u, &s2u, &c2u);
    sincos(u, &su, &cu);
    const cd e2{c2u, s2u}, em{cu, -su};
    // h_j = D f_j, D = 9u^2 - w^2; hu = dh/du; hw = (1/w) dh/dw
    cd h[3], hu[3], hw[3];
    h[0] = e2 * (u2 - w2) + em * cd{8. * u2 * cw, 2. * u * (3. * u2 + w2) * x0};
    h[1] = e2 * (2. * u) - em * cd{2. * u * cw, -(3. * u2 - w2) * x0};
    h[2] = e2 - em * cd{cw, 3. * u * x0};
    hu[0] = e2 * cd{2. * u, 2. * (u2 - w2)} +
            em * cd{16. * u * cw + 2. * u * (3. * u2 + w2) * x0, -8. * u2 * cw + 2. * (9. * u2 + w2) * x0};
    hw[0] = e2 * (-2.) + em * cd{-8. * u2 * x0, 2. * u * (2. * x0 + (3. * u2 + w2) * x1)};
    hu[1] = e2 * cd{2., 4. * u} + em * cd{(3. * u2 - w2) * x0 - 2. * cw, 2. * u * cw + 6. * u * x0};
    hw[1] = em * cd{2. * u * x0, -2. * x0 + (3. * u2 - w2) * x1};
    hu[2] = e2 * cd{0., 2.} + em * cd{-3. * u * x0, cw - 3. * x0};
    hw[2] = em * cd{x0, -3. * u * x1};
    const double D = 9. * u2 - w2;
    const double iD = 1. / D;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        f[j] = h[j] * iD;
        const cd fu = (hu[j] - f[j] * (18. * u)) * iD;
        const cd fw = (hw[j] + f[j] * 2.) * iD;
        fa[j] = (fu - fw * (3. * u)) * (0.5 * iD);
        fb[j] = fu * (u * iD) + fw * ((3. * u2 - w2) * 0.5 * iD);
    }
    if (neg) {
        // f_j(-c0) = (-1)^j conj f_j(c0); d/dc0 picks up one more sign
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double sj = (j & 1) ? -1. : 1.;
            f[j] = cd{sj * f[j].r, -sj * f[j].i};
            fa[j] = cd{-sj * fa[j].r, sj * fa[j].i};
            fb[j] = cd{sj * fb[j].r, -sj * fb[j].i};
        }
    }
}

// Q = Q^a t^a, t^a = lambda^a / 2 (Gell-Mann order)
IPG_HD M3 gen_matrix(const double (&Q)[8]) {
    const double s3 = 1. / (2. * sqrt(3.));
    M3 m;
    m.e[0] = cd{0.5 * Q[2] + s3 * Q[7], 0.};
    m.e[4] = cd{-0.5 * Q[2] + s3 * Q[7], 0.};
    m.e[8] = cd{-2. * s3 * Q[7], 0.};
    m.e[1] = cd{0.5 * Q[0], -0.5 * Q[1]};
    m.e[3] = cd{0.5 * Q[0], 0.5 * Q[1]};
    m.e[2] = cd{0.5 * Q[3], -0.5 * Q[4]};
    m.e[6] = cd{0.5 * Q[3], 0.5 * Q[4]};
    m.e[5] = cd{0.5 * Q[5], -0.5 * Q[6]};
    m.e[7] = cd{0.5 * Q[5], 0.5 * Q[6]};
    return m;
}
// Re tr(G t^a), a = 0..7
IPG_HD void gen_project(const M3& G, double (&out)[8]) {
    const double s3 = 1. / (2. * sqrt(3.));
    out[0] = 0.5 * (G.e[1].r + G.e[3].r);
    out[1] = 0.5 * (G.e[3].i - G.e[1].i);
    out[2] = 0.5 * (G.e[0].r - G.e[4].r);
    out[3] = 0.5 * (G.e[2].r + G.e[6].r);
    out[4] = 0.5 * (G.e[6].i - G.e[2].i);
    out[5] = 0.5 * (G.e[5].r + G.e[7].r);
    out[6] = 0.5 * (G.e[7].i - G.e[5].i);
    out[7] = s3 * (G.e[0].r + G.e[4].r - 2. * G.e[8].r);
}

// exp(i Q^a t^a) = f0 + f1 Q + f2 Q^2 and what its reverse needs: Q, Q^2,
// the coefficients f_j(c0, c1) and their derivatives (computed once when a
// caller wants both the exponential and its reverse)
struct ExpQ {
    M3 Qm, Q2;
    cd f[3], fa[3], fb[3];
};
IPG_HD void expq_prep(const double (&Q)[8], ExpQ& x) {
    x.Qm = gen_matrix(Q);
    x.Q2 = x.Qm * x.Qm;
    double c1 = 0.;
#pragma unroll
    for (int a = 0; a < 8; a++) c1 += Q[a] * Q[a];
    c1 *= 0.25;
    const double c0 = retrprod(x.Q2, x.Qm) / 3.;  // det Q = tr(Q^3) / 3
    expq_fcoef(c0, c1, x.f, x.fa, x.fb);
}
// the exponential from the coefficients (equal to exp_iQt to round-off)
IPG_HD M3 expq_value(const ExpQ& x) {
    M3 E = cscale(x.f[1], x.Qm) + cscale(x.f[2], x.Q2);
    E.e[0] = E.e[0] + x.f[0];
    E.e[4] = E.e[4] + x.f[0];
    E.e[8] = E.e[8] + x.f[0];
    return E;
}
// The setV slices' exponential exp(i Q^a t^a) without its reverse: the
// Cayley-Hamilton form f0 + f1 Q + f2 Q^2 with the f_j summed as the series
// of exp(i x) in the basis {1, Q, Q^2} (Q^3 = c0 + c1 Q) for c1 = tr(Q^2)/2
// below kExpqSeriesMax, stopped once the terms left are below 1.5e-19 (the
// bound of expq_fcoef); larger Q scaled into that range and the factor
// applied repeatedly (expq_apply).  No transcendental functions.  It agrees
// with exp_iQt's closed form (Matrix.cpp:310-426) to round-off except at
// (nearly) degenerate spectra, where the closed form loses digits or
// returns 1, and at |Q| below ~1e-8, where the closed form cancels
// catastrophically (its 9u^2 - w^2 denominator and a floor of 1e-15 on
// c0max) and its value, off the exponential by up to |Q|, depends on the
// last bits of acos and sincos.
constexpr double kExpqSeriesMax = 0.5;
// the f_j of the series: exp(i x) summed in the basis {1, Q, Q^2} with
// Q^3 = c0 + c1 Q (the powers Q^n = a_n + b_n Q + d_n Q^2)
IPG_HD void expq_series(double c0, double c1, cd (&f)[3]) {
    double a = 1., b = 0., d = 0.;
    double fr[3] = {0., 0., 0.}, fi[3] = {0., 0., 0.};
    double coef = 1.;  // 1/n!
    for (int n = 0; n < 26; n++) {
        const int m = n & 3;
        const double cr = m == 0 ? coef : (m == 2 ? -coef : 0.);
        const double ci = m == 1 ? coef : (m == 3 ? -coef : 0.);
        fr[0] += cr * a;
        fi[0] += ci * a;
        fr[1] += cr * b;
        fi[1] += ci * b;
        fr[2] += cr * d;
        fi[2] += ci * d;
        if (n >= 4 && coef * (fabs(a) + fabs(b) + fabs(d)) < 1e-19) break;
        const double na = c0 * d, nb = a + c1 * d, nd = b;
        a = na, b = nb, d = nd;
        coef /= (double)(n + 1);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) f[j] = cd{fr[j], fi[j]};
}
// det(Q^a t^a) from the components (gen_matrix's entries: real diagonal
// d_i, upper triangle q01, q02, q12):
// d0 d1 d2 - d0 |q12|^2 - d1 |q02|^2 - d2 |q01|^2 + 2 Re(q01 q12 conj q02)
IPG_HD double gen_det(const double (&Q)[8]) {
    const double s3 = 1. / (2. * sqrt(3.));
    const double d0 = 0.5 * Q[2] + s3 * Q[7], d1 = -0.5 * Q[2] + s3 * Q[7], d2 = -2. * s3 * Q[7];
    const cd q01{0.5 * Q[0], -0.5 * Q[1]}, q02{0.5 * Q[3], -0.5 * Q[4]}, q12{0.5 * Q[5], -0.5 * Q[6]};
    const double n01 = q01.r * q01.r + q01.i * q01.i, n02 = q02.r * q02.r + q02.i * q02.i,
                 n12 = q12.r * q12.r + q12.i * q12.i;
    return d0 * d1 * d2 - d0 * n12 - d1 * n02 - d2 * n01 + 2. * mulc(q01 * q12, q02).r;
}
// V <- (f0 + f1 Q + f2 Q^2) V column by column: v <- f0 v + f1 (Q v) +
// f2 Q (Q v), Q = Q^a t^a from its components.  Neither Q^2 nor the
// exponential is formed, so far fewer values are live than in E * V.
IPG_HD void expq_coef_apply(const double (&Q)[8], const cd (&f)[3], M3& V) {
    const double s3 = 1. / (2. * sqrt(3.));
    const double d0 = 0.5 * Q[2] + s3 * Q[7], d1 = -0.5 * Q[2] + s3 * Q[7], d2 = -2. * s3 * Q[7];
    const cd q01{0.5 * Q[0], -0.5 * Q[1]}, q02{0.5 * Q[3], -0.5 * Q[4]}, q12{0.5 * Q[5], -0.5 * Q[6]};
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const cd v0 = V.e[j], v1 = V.e[3 + j], v2 = V.e[6 + j];
        const cd w0 = v0 * d0 + q01 * v1 + q02 * v2;
        const cd w1 = cmul(q01, v0) + v1 * d1 + q12 * v2;
        const cd w2 = cmul(q02, v0) + cmul(q12, v1) + v2 * d2;
        const cd x0 = w0 * d0 + q01 * w1 + q02 * w2;
        const cd x1 = cmul(q01, w0) + w1 * d1 + q12 * w2;
        const cd x2 = cmul(q02, w0) + cmul(q12, w1) + w2 * d2;
        V.e[j] = f[0] * v0 + f[1] * w0 + f[2] * x0;
        V.e[3 + j] = f[0] * v1 + f[1] * w1 + f[2] * x1;
        V.e[6 + j] = f[0] * v2 + f[1] * w2 + f[2] * x2;
    }
}
// V <- exp(i Q^a t^a) V for any Q (the setV slices): Q scaled by 2^-s
// (exact) until c1 = tr(Q^2)/2 is below kExpqSeriesMax, the series
// coefficients of the scaled Q (expq_series, c0 = det Q), applied 2^s times
// by expq_coef_apply.  No transcendental functions; s is at most 12
// (c1 < 8.4e6), so that no input can make the loop long.
IPG_HD void expq_apply(const double (&Q)[8], M3& V) {
    double c1 = 0.;
#pragma unroll
    for (int a = 0; a < 8; a++) c1 += Q[a] * Q[a];
    c1 *= 0.25;
    double sc = 1.;
    int reps = 1;
    while (c1 >= kExpqSeriesMax && reps < 4096) {
        c1 *= 0.25;
        sc *= 0.5;
        reps *= 2;
    }
    double Qs[8];
#pragma unroll
    for (int a = 0; a < 8; a++) Qs[a] = Q[a] * sc;
    cd f[3];
    expq_series(gen_det(Qs), c1, f);
    for (int r = 0; r < reps; r++) expq_coef_apply(Qs, f, V);
}
// the slice exponential itself (expq_apply to the identity)
IPG_HD M3 expq_forward(const double (&Q)[8]) {
    M3 E = m_one();
    expq_apply(Q, E);
    return E;
}
// Qb^a = Re tr(Eb^+ dE/dQ^a)
IPG_HD void expq_back(const ExpQ& x, const M3& Eb, double (&Qb)[8]) {
    const M3 Ebd = dag(Eb);
    const cd t0 = tr(Ebd), t1 = trprod(Ebd, x.Qm), t2 = trprod(Ebd, x.Q2);
    const double b0 = (x.fa[0] * t0 + x.fa[1] * t1 + x.fa[2] * t2).r;
    const double b1 = (x.fb[0] * t0 + x.fb[1] * t1 + x.fb[2] * t2).r;
    const M3 G = (b0 * x.Q2 + b1 * x.Qm) + (cscale(x.f[1], Ebd) + cscale(x.f[2], x.Qm * Ebd + Ebd * x.Qm));
    gen_project(G, Qb);
}
IPG_HD void expq_adj(const double (&Q)[8], const M3& Eb, double (&Qb)[8]) {
    ExpQ x;
    expq_prep(Q, x);
    expq_back(x, Eb, Qb);
}

}  // namespace ipgad

