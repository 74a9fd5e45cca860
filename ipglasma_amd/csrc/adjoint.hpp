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
// Y = M1 M2 M3 M4: cotangents of the four factors
IPG_HD void adj_prod4(const M3& M1, const M3& M2, const M3& M3_, const M3& M4, const M3& Yb, M3& B1, M3& B2,
                      M3& B3, M3& B4) {
    const M3 P12 = M1 * M2, P34 = M3_ * M4;
    const M3 P234 = M2 * P34, P123 = P12 * M3_;
    B1 = mul_bd(Yb, P234);                   // Yb (M2 M3 M4)^+
    B2 = mul_bd(mul_ad(M1, Yb), P34);        // M1^+ Yb (M3 M4)^+
    B3 = mul_bd(mul_ad(P12, Yb), M4);        // (M1 M2)^+ Yb M4^+
    B4 = mul_ad(P123, Yb);                   // (M1 M2 M3)^+ Yb
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
IPG_HD void e_adj(const EIn& q, double tau, double dtau, double g, const M3& Y1b, const M3& Y2b, EBar& b) {
    const cd c1c{0., -tau * dtau / (2. * g * g)}, c2c{0., -dtau / tau};  // conjugates
    const M3 Z1 = traceless(Y1b), Z2 = traceless(Y2b);
    b.e1 = b.e1 + Z1;
    b.e2 = b.e2 + Z2;
    const M3 t1b = traceless(cscale(c1c, Z1)), t1pb = traceless(cscale(c1c, Z2));
    const M3 t2b = cscale(c2c, Z1), t2pb = cscale(c2c, Z2);
    // plaquettes: t1 = U12 + U1m2 - U12^+ - U1m2^+, t1' = U12^+ + U2m1 - U12 - U2m1^+
    const M3 U12b = (t1b - dag(t1b)) + (dag(t1pb) - t1pb);
    const M3 U1m2b = t1b - dag(t1b);
    const M3 U2m1b = t1pb - dag(t1pb);
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
    // commutators t2 = tx phi - phi tx (tx, phi Hermitian)
    const M3 tx = q.ux * mul_bd(q.phpx, q.ux), ty = q.uy * mul_bd(q.phpy, q.uy);
    M3 txb = t2b * q.ph - q.ph * t2b;
    b.ph = b.ph + (tx * t2b - t2b * tx);
    adj_conj(q.ux, q.phpx, txb, b.ux, b.phpx);
    M3 tyb = t2pb * q.ph - q.ph * t2pb;
    b.ph = b.ph + (ty * t2pb - t2pb * ty);
    adj_conj(q.uy, q.phpy, tyb, b.uy, b.phpy);
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

}  // namespace ipgad
