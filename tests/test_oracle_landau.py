"""Pins of the oracle's Landau frame (MyEigen.cpp:25-321) and eccentricities
(Evolution.cpp:1280-1810).

The reference calls GSL's gsl_eigen_nonsymmv, which is not in this image.
The oracle's 4x4 solver is pinned against LAPACK (numpy.linalg.eig) and by
the defining equation; the Landau matching by closed forms of boosted
perfect and anisotropic fluids; the eccentricities by Gaussian profiles whose
lattice moments equal the continuum ones to roundoff, by point sources and by
translation invariance.
"""
import numpy as np
import pytest

import _oracle as O

def metric(tau):
    return np.diag([1., -1., -1., -tau * tau])


def pack_T(Tmn):
    """symmetric T^{mu nu} [..., 4, 4] (tau, x, y, eta) -> the 10 components
    in the order tautau, xx, yy, etaeta, taux, tauy, taueta, xy, xeta, yeta"""
    idx = [(0, 0), (1, 1), (2, 2), (3, 3), (0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    return np.stack([Tmn[..., i, j] for i, j in idx], 0)


def random_frame(rng, tau, speed=0.8):
    """a unit timelike u and a g-orthonormal spacelike triad (X, Y, Z)"""
    g = metric(tau)
    v = rng.uniform(-1, 1, 3)
    v *= speed * rng.uniform() / np.linalg.norm(v)
    gam = 1. / np.sqrt(1. - v @ v)
    # local orthonormal frame (t, x, y, z) -> Milne components: eta scaled 1/tau
    lam = np.eye(4)
    lam[0, 0] = gam
    lam[0, 1:] = lam[1:, 0] = gam * v
    lam[1:, 1:] = np.eye(3) + (gam - 1.) * np.outer(v, v) / (v @ v)
    R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    rot = np.eye(4)
    rot[1:, 1:] = R
    E = lam @ rot                       # columns: boosted, rotated frame
    E[3, :] /= tau                      # Milne eta component
    for k in range(4):
        n = E[:, k] @ g @ E[:, k]
        assert abs(abs(n) - 1.) < 1e-12
    return E[:, 0], E[:, 1], E[:, 2], E[:, 3]


def fluid_T(eps, P, u, tau):
    """T^{mu nu} = eps u u + sum_i P_i e_i e_i  for frame (u, e_1, e_2, e_3)"""
    U, X, Y, Z = u
    return eps * np.outer(U, U) + P[0] * np.outer(X, X) + P[1] * np.outer(Y, Y) + P[2] * np.outer(Z, Z)


def lattice_of(Tsite, N=16):
    """one T per site, [S, 4, 4] -> T[10, S]"""
    return np.ascontiguousarray(pack_T(Tsite))


def run_landau(T10, N, tau, L=8.0):
    """tau [fm] = it * dtau * a with it = 1, a = L/N"""
    a = L / N
    dtau = tau / a
    out, avg, nf = O.landau(N, L, dtau, 1, T10)
    return out, avg, nf


# ---- the 4x4 eigenproblem ---------------------------------------------------

def test_eig4_against_lapack_and_definition():
    rng = np.random.default_rng(5)
    kinds = 0
    for t in range(3000):
        A = rng.normal(size=(4, 4)) * 10. ** rng.uniform(-3, 3)
        if t % 5 == 1:
            A = A + A.T                                  # real spectrum
        elif t % 5 == 2:
            A = np.triu(A)                               # already triangular
        elif t % 5 == 3:
            A[2:, :2] = 0.                               # block structure
        rc, w, V = O.eig4(A)
        assert rc == 0
        scale = np.abs(A).max()
        assert np.abs(A @ V - V * w[None, :]).max() <= 1e-12 * scale
        assert np.allclose(np.linalg.norm(V, axis=0), 1., atol=1e-13)
        we = np.linalg.eigvals(A)
        # same spectrum (greedy matching)
        left = list(we)
        for x in w:
            k = int(np.argmin([abs(x - y) for y in left]))
            assert abs(x - left[k]) <= 1e-9 * scale
            left.pop(k)
        # complex pairs: positive imaginary part first, conjugate vectors
        for k in range(3):
            if w[k].imag > 0:
                assert w[k + 1] == np.conj(w[k])
                assert np.array_equal(V[:, k + 1], np.conj(V[:, k]))
                kinds += 1
    assert kinds > 100


def test_eig4_special_matrices():
    for A in (np.zeros((4, 4)), np.eye(4), np.diag([3., -1., -1., 2.])):
        rc, w, V = O.eig4(A)
        assert rc == 0
        assert np.array_equal(V, np.eye(4))     # no rotation of a diagonal matrix
        assert np.array_equal(w.real, np.diag(A))
    # a standard 2x2 rotation block: eigenvector (1, -i)/sqrt2 for +i b
    A = np.zeros((4, 4))
    A[0, 0] = A[1, 1] = 2.
    A[0, 1], A[1, 0] = -3., 3.
    A[2, 2], A[3, 3] = -1., -5.
    rc, w, V = O.eig4(A)
    assert rc == 0
    assert abs(w[0] - (2 + 3j)) < 1e-15 and w[1] == np.conj(w[0])
    assert np.allclose(V[:, 0], np.array([1, -1j, 0, 0]) / np.sqrt(2), atol=1e-15)


# ---- Landau matching ----------------------------------------------------------

@pytest.mark.parametrize("tau", [0.2, 1.0])
def test_landau_boosted_fluids(tau):
    rng = np.random.default_rng(11)
    N = 16
    S = N * N
    Ts, truth = [], []
    for s in range(S):
        frame = random_frame(rng, tau)
        eps = rng.uniform(1., 50.)
        # distinct principal pressures (an exactly degenerate spacelike
        # eigenvalue splits into a roundoff-sized complex pair whose vectors
        # the bilinear test z.z may call timelike, with GSL as here)
        w = rng.uniform(0.2, 1.0, 3)
        if s % 2:
            P = eps * w / w.sum()                          # traceless (conformal)
        else:
            P = rng.uniform(0.05, 0.6, 3) * eps            # anisotropic, P_i < eps
        Ts.append(fluid_T(eps, P, frame, tau))
        truth.append((eps, frame, P))
    out, avg, nf = run_landau(lattice_of(np.array(Ts)), N, tau)
    assert nf == 0
    g = metric(tau)
    for s in range(S):
        eps, (U, X, Y, Z), P = truth[s]
        assert abs(out[0, s] - eps) <= 1e-12 * eps
        u = out[1:5, s]
        assert np.allclose(u, U, atol=1e-11), (s, u, U)
        assert abs(u @ g @ u - 1.) < 1e-12 and u[0] > 0
        # pi^{mu nu} = sum_i (P_i - eps/3) e_i e_i
        pim = sum((P[k] - eps / 3.) * np.outer(e, e) for k, e in enumerate((X, Y, Z)))
        assert np.allclose(out[5:, s], pack_T(pim), atol=1e-10 * eps), s
    # averages: sum u_x^2 eps etc. over the timelike eigenvectors (one per site)
    e0 = np.array([t[0] for t in truth])
    Us = np.array([t[1][0] for t in truth])
    assert np.isclose(avg[3], e0.sum(), rtol=1e-13)
    assert np.isclose(avg[0], (Us[:, 1] ** 2 * e0).sum(), rtol=1e-10)
    assert np.isclose(avg[2], (Us[:, 3] ** 2 * e0 * tau * tau).sum(), rtol=1e-10)


def test_landau_small_energy_density_keeps_rest_frame():
    """eps <= 0.1: u = (1,0,0,0), eps = the eigenvalue, pi = 0 (MyEigen.cpp:158-180)"""
    rng = np.random.default_rng(3)
    tau, N = 0.5, 16
    Ts = [fluid_T(0.05, np.array([0.01, 0.015, 0.02]), random_frame(rng, tau), tau) for _ in range(N * N)]
    out, _, _ = run_landau(lattice_of(np.array(Ts)), N, tau)
    assert np.allclose(out[0], 0.05, rtol=1e-12)
    assert np.array_equal(out[1], np.ones(N * N))
    assert np.array_equal(out[2:5], np.zeros((3, N * N)))
    assert np.array_equal(out[5:], np.zeros((10, N * N)))


def test_landau_site_without_timelike_vector_uses_carried_state():
    """a tau-x rotation block has complex eigenvalues eps0 +- i b and
    eigenvectors with z.z = 0: no timelike vector is accepted, the cell keeps
    eps = T^tautau and u = (1,0,0,0), and pi^{mu nu} is formed with the eps of
    this site's test and the u^mu of the previous site (MyEigen.cpp:103-121,
    267-309)"""
    tau, N = 0.5, 32
    rng = np.random.default_rng(7)
    Ts = []
    frames = []
    for s in range(N * N):
        fr = random_frame(rng, tau, speed=0.5)
        frames.append(fr)
        Ts.append(fluid_T(10., np.array([2., 3., 4.]), fr, tau))
    bad = [0 * N + 7, 15 * N + 15]          # an edge site and an interior site
    for s in bad:
        T = np.zeros((4, 4))
        T[0, 0] = 6.                         # M[0][0] = 6
        T[1, 1] = -6.                        # M[1][1] = 6
        T[0, 1] = T[1, 0] = 4.               # M = [[6, -4], [4, 6]], z = (1, -i)
        T[2, 2] = 1.                         # (exactly: sqrt(4) sqrt(4) = 4)
        T[3, 3] = 2. / tau ** 2
        Ts[s] = T
    out, _, _ = run_landau(lattice_of(np.array(Ts)), N, tau)
    for s in bad:
        assert out[0, s] == 6. and out[1, s] == 1. and not out[2:5, s].any()
        prev = frames[s - 1][0]
        assert np.allclose(out[1:5, s - 1], prev, atol=1e-11)
        # eps of this site's test: Re(lambda) = 6 at the edge, T^tautau = 6
        # in the interior (|Im lambda| > 0.001); both 6 here
        uu = np.outer(out[1:5, s - 1], out[1:5, s - 1])
        ginv = np.diag([1., -1., -1., -1. / tau ** 2])
        pim = Ts[s] - 4. / 3. * 6. * uu + 6. / 3. * ginv
        assert np.allclose(out[5:, s], pack_T(pim), atol=1e-10), s


# ---- eccentricities -------------------------------------------------------------

def gauss_profile(N, L, sx, sy, theta=0., x0=0., y0=0.):
    a = L / N
    x = -L / 2. + a * np.arange(N)
    X, Y = np.meshgrid(x, x, indexing="ij")
    c, s = np.cos(theta), np.sin(theta)
    xr = c * (X - x0) + s * (Y - y0)
    yr = -s * (X - x0) + c * (Y - y0)
    return np.exp(-xr ** 2 / (2 * sx ** 2) - yr ** 2 / (2 * sy ** 2)).reshape(-1)


def test_eccentricity_elliptic_gaussian_closed_form():
    N, L, sx, sy = 128, 40., 2.0, 1.4      # tails at the edge ~ exp(-50)
    e = 40. * gauss_profile(N, L, sx, sy)
    r = O.eccentricity(N, L, e, np.ones(N * N))
    assert abs(r[1] - (sx ** 2 - sy ** 2) / (sx ** 2 + sy ** 2)) < 1e-13
    assert abs(r[7] - np.pi / 2.) < 1e-13               # Psi_2 = (atan(0) + pi)/2
    assert abs(r[13] - sx * sy / np.sqrt(sx ** 2 + sy ** 2)) < 1e-12   # Rbar
    assert abs(r[16] - np.sqrt(sx ** 2 + sy ** 2)) < 1e-12             # sqrt<r^2>
    assert abs(r[17]) < 1e-13 and abs(r[18]) < 1e-13
    for k in (0, 2, 4):                                   # odd harmonics vanish
        assert r[k] < 1e-12
    assert abs(r[12] - L * L) < 1e-9                      # eps >= 0 everywhere


def test_eccentricity_rotated_gaussian_angle():
    N, L, sx, sy, th = 128, 40., 2.2, 1.3, 0.3
    e = gauss_profile(N, L, sx, sy, theta=th, x0=0.7, y0=-0.4)
    r = O.eccentricity(N, L, e, np.ones(N * N))
    assert abs(r[1] - (sx ** 2 - sy ** 2) / (sx ** 2 + sy ** 2)) < 1e-12
    # Psi_2 = (atan(tan 2theta) + pi) / 2
    assert abs(r[7] - (np.arctan(np.tan(2 * th)) + np.pi) / 2.) < 1e-12
    assert abs(r[17] - 0.7) < 1e-12 and abs(r[18] + 0.4) < 1e-12


def test_eccentricity_point_sources_and_translation():
    N, L = 64, 16.
    e = np.zeros(N * N)
    e[30 * N + 32] = e[38 * N + 32] = 5.                  # two sites on a line in x
    r = O.eccentricity(N, L, e, np.ones(N * N))
    assert abs(r[1] - 1.) < 1e-14 and abs(r[3] - 1.) < 1e-14   # eps_2 = eps_4 = 1
    assert r[2] < 1e-14 and r[0] < 1e-14                  # eps_3 = eps_1 = 0
    # a random profile is invariant under whole-site translations
    rng = np.random.default_rng(2)
    f = np.zeros((N, N))
    f[20:40, 22:38] = rng.uniform(0, 3, (20, 16))
    u = 1. + rng.uniform(0, 0.1, (N, N))
    r0 = O.eccentricity(N, L, f.reshape(-1), u.reshape(-1))
    r1 = O.eccentricity(N, L, np.roll(f, (5, -7), (0, 1)).reshape(-1),
                        np.roll(u, (5, -7), (0, 1)).reshape(-1))
    assert np.allclose(r0[:12], r1[:12], rtol=1e-11, atol=1e-13)
    assert np.allclose(r0[12:17], r1[12:17], rtol=1e-12)
    # the weight is eps u^tau; negative eps is dropped, and so is its area
    g = f.copy().reshape(-1)
    g[0] = -1.
    r2 = O.eccentricity(N, L, g, u.reshape(-1))
    assert np.isclose(r2[12], r0[12] - (L / N) ** 2, rtol=1e-14)
    assert np.allclose(r2[:12], r0[:12], rtol=1e-14)
