"""The deuteron of the oracle (Glauber::SampleTARejection, Glauber.cpp:1188-1243,
with the Hulthen thickness of Glauber.cpp:513-576 and 847-1100, placed by
Init::sampleTA, Init.cpp:90-152):
- the tabulated, adaptively integrated and cubically interpolated thickness
  against scipy.integrate.quad of the same integrand (the reference's form:
  the density (e^{-a r} - e^{-w r})^2 / r^2 integrated in xi = e^{-a z});
- the proton-neutron separation against its target density r T(r) on
  [0, 15] fm by a Kolmogorov-Smirnov test (rotations keep the 3D distance);
- the two nucleons sit at +-1/2 of it (centre at the origin);
- at s = 0 the integral has a closed form: in the reference's variable
  xi = e^{-z} (z in units of 1/a, no 1/xi Jacobian, Glauber.cpp:874-903) it
  is int_0^inf e^{-z} (e^{-z} - e^{-k z})^2 / z^2 dz with k = w / a, a sum
  c_i e^{-l_i z} / z^2 with sum c_i = sum c_i l_i = 0, whose integral is
  sum c_i l_i ln l_i = 3 ln 3 - 2 (2+k) ln(2+k) + (1+2k) ln(1+2k); this pins
  the integrand's form (the exponents, the 1/r^2, the missing Jacobian) and
  the normalisation rho = A / f of CalcRho independently of the quadrature."""
import ctypes as C

import numpy as np
import pytest
from scipy import integrate, stats

import _oracle as O

A_HUL, W_HUL = 0.228, 1.18


def lib():
    L = O.lib()
    L.orc_deuteron_thickness.argtypes = [C.c_double] * 5
    L.orc_deuteron_thickness.restype = C.c_double
    return L


def thickness_quad(s, a, w, R0, sigma):
    f0 = 2.0 * (a - w) ** 2 * np.pi / a / w / (a + w)
    rho = 2.0 / ((f0 * R0) / R0)

    def f(xi):
        xi = min(max(xi, 1e-10), 1 - 1e-10)
        z = -np.log(xi)
        r = np.hypot(s * a, z)
        g = (np.exp(-r) - np.exp(-(w / a) * r)) / r
        return 2.0 * a * rho * sigma * g * g

    return integrate.quad(f, 0.0, 1.0, epsabs=0, epsrel=1e-12, limit=400)[0]


@pytest.mark.parametrize("a,R0", [(A_HUL, 1.0), (0.535, 6.38)])
def test_thickness_table_matches_quadrature(a, R0):
    """at the 101 table nodes (0.24 fm apart) the adaptive integral to 1e-6;
    between them the cubic through the four nodes around s (x between the
    second and third, the first or last four at the ends), which np.polyfit
    builds here from the quadrature's node values"""
    sigma = 4.2
    L = lib()
    h = 24.0 / 100
    nodes = np.arange(101) * h
    tq = np.array([thickness_quad(x, a, W_HUL, R0, sigma) for x in nodes])
    for k in range(0, 101, 5):
        got = L.orc_deuteron_thickness(nodes[k], a, W_HUL, R0, sigma)
        assert got == pytest.approx(tq[k], rel=1e-5, abs=1e-9 * tq[0]), nodes[k]
    for s in (0.05, 0.13, 0.5, 1.0, 1.7, 2.5, 4.0, 6.3, 9.0, 12.5, 23.9):
        k = min(max(int(100 * s / 24.0) - 1, 0), 97)
        c = np.polyfit(nodes[k:k + 4] - nodes[k], tq[k:k + 4], 3)
        want = max(np.polyval(c, s - nodes[k]), 0.0)
        got = L.orc_deuteron_thickness(s, a, W_HUL, R0, sigma)
        assert got == pytest.approx(want, rel=1e-5, abs=1e-9 * tq[0]), s
    # beyond the table (2 x 12 fm) the thickness is 0
    assert L.orc_deuteron_thickness(24.5, a, W_HUL, R0, sigma) == 0.0


@pytest.mark.parametrize("a,R0,sigma", [(A_HUL, 1.0, 4.2), (0.535, 6.38, 7.0)])
def test_thickness_at_zero_closed_form(a, R0, sigma):
    k = W_HUL / a
    I = 3 * np.log(3) - 2 * (2 + k) * np.log(2 + k) + (1 + 2 * k) * np.log(1 + 2 * k)
    f0 = 2.0 * (a - W_HUL) ** 2 * np.pi / a / W_HUL / (a + W_HUL)
    want = 2.0 * a * (2.0 / f0) * sigma * I
    assert lib().orc_deuteron_thickness(0.0, a, W_HUL, R0, sigma) == pytest.approx(want, rel=1e-7)


def test_separation_distribution_and_placement():
    import _cases as K
    L = lib()
    p = K.params("pPb", N=32, Projectile="d", Target="p", ws_override=False)
    cfg = K.glauber_cfg(p)
    assert cfg.A1 == 2 and cfg.Z1 == 1
    st = O.MT()
    L.orc_mt_init(C.byref(st), 99)
    nA = (O.Nucleon * 2)()
    nB = (O.Nucleon * 1)()
    info = O.EventInfo()
    r = np.empty(20000)
    for k in range(r.size):
        assert L.orc_sample_event(C.byref(st), C.byref(cfg), nA, nB, C.byref(info)) == 0
        a, b = K.nucleon_array(nA)[:, :3], K.nucleon_array(nA)[:, :3]
        # centred before the -b/2 shift: the pair's midpoint is the shift
        mid = (a[0] + a[1]) / 2
        assert abs(mid[2]) < 1e-12
        r[k] = np.linalg.norm(a[0] - a[1])
        assert nA[0].proton == 1 and nA[1].proton == 1
    sigma = 0.1 * p.SigmaNN
    grid = np.linspace(0.0, 15.0, 3001)
    dens = grid * np.array([L.orc_deuteron_thickness(x, A_HUL, W_HUL, 1.0, sigma) for x in grid])
    cdf = np.concatenate([[0.0], np.cumsum((dens[1:] + dens[:-1]) / 2 * np.diff(grid))])
    cdf /= cdf[-1]
    ks = stats.kstest(r, lambda x: np.interp(x, grid, cdf))
    assert ks.pvalue > 1e-3, ks
    assert r.max() <= 15.0 + 1e-9


def test_deuteron_in_target_second_nucleon_is_a_neutron():
    import _cases as K
    L = lib()
    p = K.params("pPb", N=32, Projectile="p", Target="d", ws_override=False)
    cfg = K.glauber_cfg(p)
    st = O.MT()
    L.orc_mt_init(C.byref(st), 7)
    nA = (O.Nucleon * 1)()
    nB = (O.Nucleon * 2)()
    info = O.EventInfo()
    assert L.orc_sample_event(C.byref(st), C.byref(cfg), nA, nB, C.byref(info)) == 0
    assert nB[0].proton == 1 and nB[1].proton == 0
