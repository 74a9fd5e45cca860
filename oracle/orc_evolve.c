/* orc_evolve.c -- oracle leapfrog, T^{mu nu}, Gauss law, dE/deta.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   orc_evolve_u     Evolution::evolveU   (Evolution.cpp:35-88):
 *                    U <- (1 + X + X^2/2) U,  X = i g^2 dtau/(tau+dtau/2) E
 *   orc_evolve_phi   Evolution::evolvePhi (Evolution.cpp:90-115)
 *   orc_evolve_pi    Evolution::evolvePi  (Evolution.cpp:117-178)
 *   orc_evolve_e     Evolution::evolveE   (Evolution.cpp:180-293)
 *   orc_gauss_law    Evolution::checkGaussLaw (Evolution.cpp:295-364)
 *   orc_run          Evolution::run schedule (Evolution.cpp:366-431): initial
 *                    phi/U half-kick at tau=0, itmax-1 full steps, T^{mu nu}
 *                    at it=itmax before the final half step of pi and E
 *   orc_tmunu        Evolution::Tmunu (Evolution.cpp:659-1246)
 *   orc_dedeta       tau a^2 hbarc sum T^tautau (the reference prints no
 *                    dE/deta of this form; it is the defined observable)
 * Boundaries are clamped (copy-edge), Lattice.cpp:19-28.
 */
#include <math.h>
#include <stdlib.h>

#include "ipg_oracle.h"
#include "orc_internal.h"

static const double HBARC = 0.1973269718;

#define LD(f, s) m_load(&(f)[18 * (size_t)(s)])
#define ST(f, s, m) m_store(&(f)[18 * (size_t)(s)], (m))

static inline int cl(int v, int N) { return v < 0 ? 0 : (v > N - 1 ? N - 1 : v); }
static inline int at(int i, int j, int N) { return cl(i, N) * N + cl(j, N); }

void orc_evolve_u(int N, double g, double dtau, double tau, const double *E1,
                  const double *E2, double *Ux, double *Uy) {
    const int S = N * N;
    const cplx c = CMPLX(0., g * g * dtau / (tau + dtau / 2.) * orc_mut[ORC_MUT_EVOLVE_U]);
#pragma omp parallel for
    for (int s = 0; s < S; s++) {
        M3 one = m_one();
        M3 X = m_cscale(c, LD(E1, s));
        M3 t2 = m_add(one, m_scale(1. / 2., X));
        M3 M1 = m_add(one, m_scale(1., m_mul(X, t2)));
        X = m_cscale(c, LD(E2, s));
        t2 = m_add(one, m_scale(1. / 2., X));
        M3 M2 = m_add(one, m_scale(1., m_mul(X, t2)));
        ST(Ux, s, m_mul(M1, LD(Ux, s)));
        ST(Uy, s, m_mul(M2, LD(Uy, s)));
    }
}

void orc_evolve_phi(int N, double dtau, double tau, const double *pi, double *phi) {
    const int S = N * N;
    const double c = (tau + dtau / 2.) * dtau * orc_mut[ORC_MUT_EVOLVE_PHI];
#pragma omp parallel for
    for (int s = 0; s < S; s++) ST(phi, s, m_add(LD(phi, s), m_scale(c, LD(pi, s))));
}

void orc_evolve_pi(int N, double dtau, double tau, const double *Ux, const double *Uy,
                   const double *phi, double *pi) {
#pragma omp parallel for
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) {
            int s = i * N + j;
            M3 ux = LD(Ux, s), uy = LD(Uy, s);
            M3 pX = m_mul(ux, m_mul_bd(LD(phi, at(i + 1, j, N)), ux));
            M3 pY = m_mul(uy, m_mul_bd(LD(phi, at(i, j + 1, N)), uy));
            M3 uxm = LD(Ux, at(i - 1, j, N)), uym = LD(Uy, at(i, j - 1, N));
            M3 pmX = m_mul(m_mul_ad(uxm, LD(phi, at(i - 1, j, N))), uxm);
            M3 pmY = m_mul(m_mul_ad(uym, LD(phi, at(i, j - 1, N))), uym);
            M3 br = m_sub(m_add(m_add(m_add(pX, pmX), pY), pmY), m_scale(4., LD(phi, s)));
            ST(pi, s, m_add(LD(pi, s), m_scale(dtau / (tau) * orc_mut[ORC_MUT_EVOLVE_PI], br)));
        }
    }
}

void orc_evolve_e(int N, double g, double dtau, double tau, const double *Ux,
                  const double *Uy, const double *phi, double *E1, double *E2) {
    const cplx c1 = CMPLX(0., 1.) * tau * dtau / (2. * g * g) * orc_mut[ORC_MUT_E_PLAQ];
    const cplx c2 = CMPLX(0., 1.) * dtau / tau * orc_mut[ORC_MUT_E_COMM];
#pragma omp parallel for
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) {
            int s = i * N + j;
            M3 ux = LD(Ux, s), uy = LD(Uy, s), ph = LD(phi, s);
            M3 phN = m_mul(ux, m_mul_bd(LD(phi, at(i + 1, j, N)), ux));
            M3 U12 = m_mul(m_mul(ux, LD(Uy, at(i + 1, j, N))),
                           m_mul_bd(m_dag(LD(Ux, at(i, j + 1, N))), uy));
            M3 U1m2 = m_mul(m_mul_bd(ux, LD(Uy, at(i + 1, j - 1, N))),
                            m_mul_ad(LD(Ux, at(i, j - 1, N)), LD(Uy, at(i, j - 1, N))));
            M3 U2m1 = m_mul(m_mul_bd(uy, LD(Ux, at(i - 1, j + 1, N))),
                            m_mul_ad(LD(Uy, at(i - 1, j, N)), LD(Ux, at(i - 1, j, N))));
            M3 U12d = m_dag(U12);
            M3 t1 = m_sub(m_sub(m_add(U12, U1m2), U12d), m_dag(U1m2));
            t1 = m_traceless(t1);
            M3 t2 = m_sub(m_mul(phN, ph), m_mul(ph, phN));
            M3 En = m_add(LD(E1, s), m_add(m_cscale(c1, t1), m_cscale(c2, t2)));
            ST(E1, s, m_traceless(En));

            t1 = m_sub(m_sub(m_add(U12d, U2m1), U12), m_dag(U2m1));
            t1 = m_traceless(t1);
            phN = m_mul(uy, m_mul_bd(LD(phi, at(i, j + 1, N)), uy));
            t2 = m_sub(m_mul(phN, ph), m_mul(ph, phN));
            En = m_add(LD(E2, s), m_add(m_cscale(c1, t1), m_cscale(c2, t2)));
            ST(E2, s, m_traceless(En));
        }
    }
}

double orc_gauss_law(int N, const double *Ux, const double *Uy, const double *E1,
                     const double *E2, const double *phi, const double *pi) {
    double largest = 0;
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) {
            int s = i * N + j, mx = at(i - 1, j, N), my = at(i, j - 1, N);
            M3 uxm = LD(Ux, mx), uym = LD(Uy, my);
            M3 ph = LD(phi, s), p = LD(pi, s);
            M3 G = m_mul(m_mul(m_dag(uxm), LD(E1, mx)), uxm);
            G = m_sub(G, LD(E1, s));
            G = m_add(G, m_mul(m_mul(m_dag(uym), LD(E2, my)), uym));
            G = m_sub(G, LD(E2, s));
            G = m_sub(G, m_cscale(CMPLX(0., orc_mut[ORC_MUT_GAUSS_COMM]), m_sub(m_mul(ph, p), m_mul(p, ph))));
            double sq = m_square(G);
            if (sq > largest) largest = sq;
        }
    }
    return largest;
}

/* (P - Q) minus its trace part, as at Evolution.cpp:1050-1060 */
static inline M3 tlpair(M3 P, M3 Q) {
    M3 d = m_sub(P, Q);
    M3 o = m_cscale(m_tr(d) / 3.0, m_one());
    return m_sub(d, o);
}
static inline M3 tlquad(M3 P1, M3 P2, M3 P3, M3 P4) {
    M3 d1 = m_sub(P1, P2);
    M3 d2 = m_sub(P3, P4);
    M3 r = m_sub(d1, m_cscale(m_tr(d1) / 3.0, m_one()));
    r = m_sub(m_add(r, P3), P4);
    return m_sub(r, m_cscale(m_tr(d2) / 3.0, m_one()));
}

void orc_tmunu(int N, double g, double a, double dtau, int it, const double *Ux,
               const double *Uy, const double *E1, const double *E2, const double *phi,
               const double *pi, double *T) {
    const int S = N * N;
    const double tau = it * dtau;
    const double Nc = 3.0;
    double *TT = T, *Txx = T + S, *Tyy = T + 2 * (size_t)S, *Tee = T + 3 * (size_t)S;
    double *Ttx = T + 4 * (size_t)S, *Tty = T + 5 * (size_t)S, *Tte = T + 6 * (size_t)S;
    double *Txy = T + 7 * (size_t)S, *Txe = T + 8 * (size_t)S, *Tye = T + 9 * (size_t)S;
#pragma omp parallel for
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) {
            int s = i * N + j;
            int pX = at(i + 1, j, N), pY = at(i, j + 1, N), pXY = at(i + 1, j + 1, N);
            int mX = at(i - 1, j, N), mY = at(i, j - 1, N);
            int mXpY = at(i - 1, j + 1, N), pXmY = at(i + 1, j - 1, N);
            int p2X = at(i + 2, j, N), p2Y = at(i, j + 2, N);
            int p2XY = at(i + 2, j + 1, N), pX2Y = at(i + 1, j + 2, N);
            /* electric and longitudinal (Evolution.cpp:754-814) */
            M3 e1 = LD(E1, s), e2 = LD(E2, s), e1p = LD(E1, pY), e2p = LD(E2, pX);
            M3 p = LD(pi, s), pXm = LD(pi, pX), pYm = LD(pi, pY), pXYm = LD(pi, pXY);
            cplx t11 = m_tr(m_mul(e1, e1)), t1p = m_tr(m_mul(e1p, e1p));
            cplx t22 = m_tr(m_mul(e2, e2)), t2p = m_tr(m_mul(e2p, e2p));
            double lng = (creal(m_tr(m_mul(p, p))) + creal(m_tr(m_mul(pXm, pXm))) +
                          creal(m_tr(m_mul(pYm, pYm))) + creal(m_tr(m_mul(pXYm, pXYm)))) / 4. *
                         orc_mut[ORC_MUT_T_LONG];
            double gt = g * g / tau / tau * orc_mut[ORC_MUT_T_ELEC];
            double tt = gt * creal(t11 + t1p + t22 + t2p) / 2. + lng;
            double xx = gt * creal(-1. * t11 - t1p + t22 + t2p) / 2. + lng;
            double yy = gt * creal(t11 + t1p - t22 - t2p) / 2. + lng;
            double ee = 1. / tau / tau * (gt * creal(t11 + t1p + t22 + t2p) / 2. - lng);
            /* magnetic and transverse phi gradients (Evolution.cpp:816-905) */
            M3 ux = LD(Ux, s), uy = LD(Uy, s);
            M3 plaq = (i == N - 1) ? m_one()
                                   : m_mul(ux, m_mul(LD(Uy, pX), m_mul(m_dag(LD(Ux, pY)), m_dag(uy))));
            M3 ph = LD(phi, s), phX = LD(phi, pX), phY = LD(phi, pY), phXY = LD(phi, pXY);
            M3 ptX = m_mul(m_mul(ux, phX), m_dag(ux));
            M3 ptY = m_mul(m_mul(uy, phY), m_dag(uy));
            M3 uxpY = LD(Ux, pY), uypX = LD(Uy, pX);
            M3 ptXY1 = m_mul(m_mul(uxpY, phXY), m_dag(uxpY));
            M3 ptXY2 = m_mul(m_mul(uypX, phXY), m_dag(uypX));
            M3 d;
            d = m_sub(ph, ptX);
            double A = creal(m_tr(m_mul(d, d)));
            d = m_sub(phY, ptXY1);
            double B = creal(m_tr(m_mul(d, d)));
            d = m_sub(ph, ptY);
            double C = creal(m_tr(m_mul(d, d)));
            d = m_sub(phX, ptXY2);
            double D = creal(m_tr(m_mul(d, d)));
            double mag = 2. / pow(g, 2.) * (Nc - creal(m_tr(plaq))) * orc_mut[ORC_MUT_T_MAG];
            double h = 0.5 / tau / tau * orc_mut[ORC_MUT_T_PHIGRAD];
            tt = (tt + mag) + h * (A + B + C + D);
            xx = (xx + mag) + h * (A + B - C - D);
            yy = (yy + mag) + h * (-A - B + C + D);
            ee = ee + 1. / tau / tau * (-mag + h * (A + B + C + D));
            TT[s] = tt * 1 / pow(a, 4.);
            Txx[s] = xx * 1 / pow(a, 4.);
            Tyy[s] = yy * 1 / pow(a, 4.);
            Tee[s] = ee * 1 / pow(a, 6.);

            /* flux components (Evolution.cpp:929-1241) */
            M3 Udx = m_dag(ux), Udy = m_dag(uy);
            M3 UxmX = LD(Ux, mX), UDxmX = m_dag(UxmX);
            M3 UxmXpY = LD(Ux, mXpY), UDxmXpY = m_dag(UxmXpY);
            M3 UxpX = LD(Ux, pX), UDxpX = m_dag(UxpX);
            M3 UxpY = uxpY, UDxpY = m_dag(UxpY);
            M3 UxpXpY = LD(Ux, pXY), UDxpXpY = m_dag(UxpXpY);
            M3 UymY = LD(Uy, mY), UDymY = m_dag(UymY);
            M3 UypXmY = LD(Uy, pXmY), UDypXmY = m_dag(UypXmY);
            M3 UypY = LD(Uy, pY), UDypY = m_dag(UypY);
            M3 UypX = uypX, UDypX = m_dag(UypX);
            M3 Uyp2X = LD(Uy, p2X), UDyp2X = m_dag(Uyp2X);
            M3 Uxp2Y = LD(Ux, p2Y), UDxp2Y = m_dag(Uxp2Y);
            M3 UypXpY = LD(Uy, pXY), UDypXpY = m_dag(UypXpY);
            M3 UymX = LD(Uy, mX), UDymX = m_dag(UymX);
            M3 UxmY = LD(Ux, mY), UDxmY = m_dag(UxmY);
            M3 phmX = LD(phi, mX), phmY = LD(phi, mY), phmXpY = LD(phi, mXpY);
            M3 phpXmY = LD(phi, pXmY), ph2X = LD(phi, p2X), ph2XY = LD(phi, p2XY);
            M3 ph2Y = LD(phi, p2Y), phX2Y = LD(phi, pX2Y);

            M3 P1 = m_mul4(uy, UxpY, UDypX, Udx), P2 = m_mul4(ux, UypX, UDxpY, Udy);
            M3 P3 = m_mul4(UDxmX, UymX, UxmXpY, Udy), P4 = m_mul4(uy, UDxmXpY, UDymX, UxmX);
            M3 P5 = m_mul4(UypX, UxpXpY, UDyp2X, UDxpX), P6 = m_mul4(UxpX, Uyp2X, UDxpXpY, UDypX);
            M3 P7 = m_mul4(Udx, uy, UxpY, UDypX), P8 = m_mul4(UypX, UDxpY, Udy, ux);
            M3 fx = m_add(m_mul(e2, tlquad(P1, P2, P3, P4)), m_mul(e2p, tlquad(P5, P6, P7, P8)));
            M3 kx = m_add(m_add(m_add(
                m_mul(p, m_sub(m_mul3(ux, phX, Udx), m_mul3(UDxmX, phmX, UxmX))),
                m_mul(pYm, m_sub(m_mul3(UxpY, phXY, UDxpY), m_mul3(UDxmXpY, phmXpY, UxmXpY)))),
                m_mul(pXm, m_sub(m_mul3(UxpX, ph2X, UDxpX), m_mul3(Udx, ph, ux)))),
                m_mul(pXYm, m_sub(m_mul3(UxpXpY, ph2XY, UDxpXpY), m_mul3(UDxpY, phY, UxpY))));
            double ttx = +2. / tau / 8. * cimag(m_tr(fx)) * orc_mut[ORC_MUT_TX_F] -
                         2. / 8. / tau * creal(m_tr(kx)) * orc_mut[ORC_MUT_TX_K];

            M3 Q1 = m_mul4(ux, UypX, UDxpY, Udy), Q2 = m_mul4(uy, UxpY, UDypX, Udx);
            M3 Q3 = m_mul4(UDymY, UxmY, UypXmY, Udx), Q4 = m_mul4(ux, UDypXmY, UDxmY, UymY);
            M3 Q5 = m_mul4(UxpY, UypXpY, UDxp2Y, UDypY), Q6 = m_mul4(UypY, Uxp2Y, UDypXpY, UDxpY);
            M3 Q7 = m_mul4(Udy, ux, UypX, UDxpY), Q8 = m_mul4(UxpY, UDypX, Udx, uy);
            M3 fy = m_add(m_mul(e1, tlquad(Q1, Q2, Q3, Q4)), m_mul(e1p, tlquad(Q5, Q6, Q7, Q8)));
            M3 ky = m_add(m_add(m_add(
                m_mul(p, m_sub(m_mul3(uy, phY, Udy), m_mul3(UDymY, phmY, UymY))),
                m_mul(pXm, m_sub(m_mul3(UypX, phXY, UDypX), m_mul3(UDypXmY, phpXmY, UypXmY)))),
                m_mul(pYm, m_sub(m_mul3(UypY, ph2Y, UDypY), m_mul3(Udy, ph, uy)))),
                m_mul(pXYm, m_sub(m_mul3(UypXpY, phX2Y, UDypXpY), m_mul3(UDypX, phX, UypX))));
            double tty = +2. / tau / 8. * cimag(m_tr(fy)) * orc_mut[ORC_MUT_TY_F] -
                         2. / 8. / tau * creal(m_tr(ky)) * orc_mut[ORC_MUT_TY_K];

            M3 gX = m_sub(m_mul3(ux, phX, Udx), ph);       /* Ux phiX Ux^+ - phi */
            M3 gXY1 = m_sub(m_mul3(UxpY, phXY, UDxpY), phY);
            M3 gY = m_sub(m_mul3(uy, phY, Udy), ph);
            M3 gXY2 = m_sub(m_mul3(UypX, phXY, UDypX), phX);
            M3 fe = m_add(m_add(m_add(m_mul(e1, gX), m_mul(e1p, gXY1)), m_mul(e2, gY)), m_mul(e2p, gXY2));
            double tte = g / tau / tau / tau * creal(m_tr(fe)) * orc_mut[ORC_MUT_TE];

            M3 el = m_mul(m_scale(-1. / 4. * g * g, m_add(e1, m_mul3(uy, e1p, Udy))),
                          m_add(e2, m_mul3(ux, e2p, Udx)));
            M3 b1 = m_mul(gX, gY);
            M3 b2 = m_mul(m_mul(m_mul(uy, gXY1), Udy), gY);
            M3 b3 = m_mul(m_mul(m_mul(gX, ux), gXY2), Udx);
            M3 b4 = m_mul(m_mul(m_mul(m_mul(m_mul(uy, gXY1), Udy), ux), gXY2), Udx);
            M3 fxy = m_add(m_scale(orc_mut[ORC_MUT_XY_E], el),
                           m_scale(1. / 4. * orc_mut[ORC_MUT_XY_B], m_add(m_add(m_add(b1, b2), b3), b4)));
            double txy = 2. / tau / tau * creal(m_tr(fxy));

            M3 R1 = m_mul4(ux, UypX, UDxpY, Udy), R2 = m_mul4(uy, UxpY, UDypX, Udx);
            M3 R3 = m_mul4(uy, UDxmXpY, UDymX, UxmX), R4 = m_mul4(UDxmX, UymX, UxmXpY, Udy);
            M3 R5 = m_mul4(UypX, UDxpY, Udy, ux), R6 = m_mul4(Udx, uy, UxpY, UDypX);
            M3 R7 = m_mul4(UxpX, Uyp2X, UDxpXpY, UDypX), R8 = m_mul4(UypX, UxpXpY, UDyp2X, UDxpX);
            double xe1 = 1. / 4. * g *
                         creal(m_tr(m_add(m_mul(e1, m_add(p, m_mul3(ux, pXm, Udx))),
                                          m_mul(e1p, m_add(pYm, m_mul3(UxpY, pXYm, UDxpY))))));
            double xe2 = 1. / 8. / g *
                         cimag(m_tr(m_add(m_mul(tlquad(R1, R2, R3, R4), gY),
                                          m_mul(tlquad(R5, R6, R7, R8), gXY2))));
            double txe = -2. / tau / tau * (xe1 * orc_mut[ORC_MUT_XE_E] + xe2 * orc_mut[ORC_MUT_XE_B]);

            M3 S1 = m_mul4(uy, UxpY, UDypX, Udx), S2 = m_mul4(ux, UypX, UDxpY, Udy);
            M3 S3 = m_mul4(ux, UDypXmY, UDxmY, UymY), S4 = m_mul4(UDymY, UxmY, UypXmY, Udx);
            M3 S5 = m_mul4(UxpY, UDypX, Udx, uy), S6 = m_mul4(Udy, ux, UypX, UDxpY);
            M3 S7 = m_mul4(UypY, Uxp2Y, UDypXpY, UDxpY), S8 = m_mul4(UxpY, UypXpY, UDxp2Y, UDypY);
            double ye1 = 1. / 4. * g *
                         creal(m_tr(m_add(m_mul(e2, m_add(p, m_mul3(uy, pYm, Udy))),
                                          m_mul(e2p, m_add(pXm, m_mul3(UypX, pXYm, UDypX))))));
            double ye2 = 1. / 8. / g *
                         cimag(m_tr(m_add(m_mul(tlquad(S1, S2, S3, S4), gX),
                                          m_mul(tlquad(S5, S6, S7, S8), gXY1))));
            double tye = -2. / tau / tau * (ye1 * orc_mut[ORC_MUT_YE_E] + ye2 * orc_mut[ORC_MUT_YE_B]);

            Ttx[s] = ttx * 1 / pow(a, 4.);
            Tty[s] = tty * 1 / pow(a, 4.);
            Tte[s] = tte * 1 / pow(a, 5.);
            Txy[s] = txy * 1 / pow(a, 4.);
            Txe[s] = txe * 1 / pow(a, 5.);
            Tye[s] = tye * 1 / pow(a, 5.);
        }
    }
    (void)tlpair;
}

double orc_dedeta(int N, double a, double tau_fm, const double *Ttt) {
    /* Neumaier-compensated sum in site order */
    double sum = 0., comp = 0.;
    for (size_t s = 0; s < (size_t)N * N; s++) {
        double v = Ttt[s], t = sum + v;
        if (fabs(sum) >= fabs(v))
            comp += (sum - t) + v;
        else
            comp += (v - t) + sum;
        sum = t;
    }
    return tau_fm * a * a * HBARC * (sum + comp);
}

void orc_time_steps(double L, int N, double maxtime, double *dtau, int *itmax) {
    double a = L / (double)N;
    int steps = (int)(10 * maxtime / a) + 1;
    *dtau = maxtime / (steps * a);
    *itmax = (int)(maxtime / (a * *dtau) + 0.1);
}

int orc_run(int N, double g, double L, double dtau, int itmax, double *Ux, double *Uy,
            double *E1, double *E2, double *phi, double *pi, double *T, double obs[3]) {
    const double a = L / N;
    orc_evolve_phi(N, dtau, 0., pi, phi);
    orc_evolve_u(N, g, dtau, 0., E1, E2, Ux, Uy);
    for (int it = 1; it <= itmax; it++) {
        if (it == itmax) orc_tmunu(N, g, a, dtau, it, Ux, Uy, E1, E2, phi, pi, T);
        if (it < itmax) {
            orc_evolve_pi(N, dtau, it * dtau, Ux, Uy, phi, pi);
            orc_evolve_e(N, g, dtau, it * dtau, Ux, Uy, phi, E1, E2);
            orc_evolve_phi(N, dtau, it * dtau, pi, phi);
            orc_evolve_u(N, g, dtau, it * dtau, E1, E2, Ux, Uy);
        } else {
            orc_evolve_pi(N, dtau / 2., it * dtau, Ux, Uy, phi, pi);
            orc_evolve_e(N, g, dtau / 2., it * dtau, Ux, Uy, phi, E1, E2);
        }
    }
    double tau_fm = itmax * dtau * a;
    obs[0] = itmax >= 1 ? orc_dedeta(N, a, tau_fm, T) : 0.;
    obs[1] = orc_gauss_law(N, Ux, Uy, E1, E2, phi, pi);
    obs[2] = tau_fm;
    return 0;
}
