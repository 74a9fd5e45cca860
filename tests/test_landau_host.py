"""The Landau solver that the GPU runs (ipglasma_amd/csrc/landau.hpp), built
for the host, against the oracle: the same operations in the same order
without contraction, so the results must be bit-identical.  Covers random
matrices of every block structure, lattices of boosted fluids, arbitrary
symmetric tensors and sites without a timelike eigenvector (whose pi^{mu nu}
comes from the carried state)."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import _oracle as O
from test_oracle_landau import fluid_T, pack_T, random_frame

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host", "landau_host.cpp")
OUT = os.path.join(HERE, "_build", "liblandau_host.so")
HIPCC = "/opt/rocm/bin/hipcc"

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
            pytest.skip("hipcc not available")
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-x", "hip", "--offload-host-only", "-O2",
               "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", OUT, SRC]
        subprocess.run(cmd, check=True, capture_output=True)
        L = C.CDLL(OUT)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C")
        L.ipgl_eig4_host.argtypes = [dp] * 5
        L.ipgl_eig4_host.restype = C.c_int
        L.ipgl_landau_host.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, dp, dp, dp,
                                       C.POINTER(C.c_int)]
        L.ipgl_landau_cell_host.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ipgl_landau_adj_host.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                           C.c_double, C.c_double, dp]
        L.ipgl_ecc2_adj_host.argtypes = [C.c_int, C.c_double, C.c_double, dp, dp, dp, dp]
        _lib = L
    return _lib


def test_eig4_device_code_bit_identical():
    L = lib()
    rng = np.random.default_rng(0)
    for t in range(20000):
        A = rng.normal(size=(4, 4)) * 10. ** rng.uniform(-3, 3)
        if t % 5 == 1:
            A = A + A.T
        elif t % 5 == 2:
            A = np.triu(A)
        elif t % 5 == 3:
            A[2:, :2] = 0.
        if t % 7 == 4:
            A[rng.integers(0, 4), rng.integers(0, 4)] = 0.
            A[1, 0] = 0.
        A = np.ascontiguousarray(A.reshape(16))
        o = [np.zeros(4), np.zeros(4), np.zeros(16), np.zeros(16)]
        h = [np.zeros(4), np.zeros(4), np.zeros(16), np.zeros(16)]
        assert O.lib().orc_eig4(A, *o) == L.ipgl_eig4_host(A, *h)
        for x, y in zip(o, h):
            assert np.array_equal(x, y), t


@pytest.mark.parametrize("tau", [0.1, 0.4, 2.0])
def test_landau_device_code_bit_identical(tau):
    L = lib()
    rng = np.random.default_rng(int(tau * 10))
    N, Lbox = 32, 8.
    S = N * N
    a = Lbox / N
    dtau = tau / a
    Ts = []
    for s in range(S):
        fr = random_frame(rng, tau, 0.7)
        e = rng.uniform(0.01, 30.)
        w = rng.uniform(0.2, 1., 3)
        T = fluid_T(e, e * w / w.sum(), fr, tau)
        if s % 17 == 3:
            T = rng.normal(size=(4, 4))
            T = T + T.T
        if s % 29 == 5:
            T = np.zeros((4, 4))
            T[0, 0], T[1, 1] = 6., -6.
            T[0, 1] = T[1, 0] = 4.
            T[2, 2], T[3, 3] = 1., 2. / tau ** 2
        if s % 31 == 7:
            T = np.zeros((4, 4))
        Ts.append(T)
    T10 = np.ascontiguousarray(pack_T(np.array(Ts)).reshape(-1))
    o1, avg1, nf1 = O.landau(N, Lbox, dtau, 1, T10)
    o2 = np.zeros(15 * S)
    avg2 = np.zeros(4)
    nf2 = C.c_int()
    L.ipgl_landau_host(N, Lbox, dtau, 1, T10, o2, avg2, C.byref(nf2))
    assert nf1 == nf2.value == 0
    assert np.array_equal(o1, o2.reshape(15, S))
    np.testing.assert_allclose(avg2, avg1, rtol=1e-13)


def _fluid(rng, tau, eps, P, boost):
    """T^{mu nu} (tautau, xx, yy, etaeta, taux, tauy, taueta, xy, xeta, yeta) of
    an anisotropic fluid: eps u u + P_x X X + P_y Y Y + P_e Z Z with an
    orthonormal frame (u, X, Y, Z) of g = diag(1, -1, -1, -tau^2)"""
    g = np.diag([1.0, -1.0, -1.0, -tau * tau])
    # boosted frame: Lorentz boost of the rest frame
    v = np.array(boost)
    b2 = v @ v
    gam = 1.0 / np.sqrt(1.0 - b2)
    Lm = np.eye(4)
    Lm[0, 0] = gam
    Lm[0, 1:] = Lm[1:, 0] = gam * v
    Lm[1:, 1:] += (gam - 1.0) * np.outer(v, v) / b2
    # rest-frame basis in the (t, x, y, tau*eta) orthonormal frame, then eta
    # components divided by tau
    E = Lm @ np.eye(4)
    E[3, :] /= tau
    pres = [eps, P[0], P[1], P[2]]
    T = sum(pres[k] * np.outer(E[:, k], E[:, k]) for k in range(4))
    T += 1e-3 * eps * (lambda M: M + M.T)(rng.normal(size=(4, 4))) * np.outer([1, 1, 1, 1 / tau], [1, 1, 1, 1 / tau])
    idx = [(0, 0), (1, 1), (2, 2), (3, 3), (0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    return np.array([T[i, j] for i, j in idx])


@pytest.mark.parametrize("case", range(6))
def test_landau_cell_reverse_mode_matches_finite_differences(case):
    """d(eps_b eps + ut_b u^tau)/dT of the cell values against central
    differences of landau_site in random symmetric directions: boosted
    anisotropic fluids (eigenvector branch), a cold cell (eps <= 0.1: rest
    frame, only eps varies) and an edge site"""
    rng = np.random.default_rng(100 + case)
    N, a, dtau, it = 64, 0.1, 0.1, 20
    tau = it * dtau * a
    eps = [30.0, 12.0, 5.0, 50.0, 0.05, 8.0][case]
    boost = [(0.3, 0.1, 0.05), (0.5, -0.2, 0.1), (0.1, 0.4, -0.2), (0.6, 0.3, 0.2), (0.2, 0.1, 0.0),
             (0.4, 0.2, 0.1)][case]
    P = np.array([0.3, 0.25, 0.2]) * eps * (1 + 0.3 * rng.random(3))
    T = _fluid(rng, tau, eps, P, boost)
    si, sj = (30, 30) if case < 5 else (0, 7)
    eb, ub = rng.normal(), rng.normal()
    Tb = np.zeros(10)
    lib().ipgl_landau_adj_host(T, N, a, dtau, it, si, sj, eb, ub, Tb)

    def loss(Tx):
        e, u = C.c_double(), C.c_double()
        lib().ipgl_landau_cell_host(np.ascontiguousarray(Tx), N, a, dtau, it, si, sj, C.byref(e), C.byref(u))
        return eb * e.value + ub * u.value

    for trial in range(4):
        d = rng.normal(size=10) * np.abs(T).max()
        h = 1e-6
        fd = (loss(T + h * d) - loss(T - h * d)) / (2 * h)
        pred = float(Tb @ d)
        assert abs(pred - fd) <= 1e-6 * max(1.0, abs(fd)), (case, trial, pred, fd)


def test_eps2_reverse_mode_matches_oracle_finite_differences():
    """d eps_2 / d(eps, u^tau) (centroid included, cutoff at 0 drops the
    negative-eps sites) against central differences of the oracle's
    orc_eccentricity on an off-centre elliptic profile with noise"""
    import _oracle as O
    rng = np.random.default_rng(5)
    N, L = 24, 12.0
    a = L / N
    x = -L / 2 + a * np.arange(N)
    X, Y = np.meshgrid(x, x, indexing="ij")
    eps = 20 * np.exp(-((X - 0.7) ** 2 / 6 + (Y + 0.4) ** 2 / 3)) + 0.5 * rng.normal(size=(N, N))
    ut = 1 + 0.1 * rng.random((N, N))
    eps, ut = eps.reshape(-1), ut.reshape(-1)
    assert (eps < 0).any()
    eb, ub = np.zeros(N * N), np.zeros(N * N)
    lib().ipgl_ecc2_adj_host(N, L, 0.0, np.ascontiguousarray(eps), np.ascontiguousarray(ut), eb, ub)
    z = np.zeros(N * N)

    def e2(e, u):
        return O.eccentricity(N, L, e, u, z, z, 0.8, 0.8, 0.0)[1]

    for trial in range(4):
        de = rng.normal(size=N * N) * (eps > 0.05)  # keep every site on its side of the cutoff
        du = rng.normal(size=N * N) * 0.1
        h = 1e-6
        fd = (e2(eps + h * de, ut + h * du) - e2(eps - h * de, ut - h * du)) / (2 * h)
        pred = float(eb @ de + ub @ du)
        assert abs(pred - fd) <= 1e-7 * max(abs(fd), 1e-3), (trial, pred, fd)


def test_landau_cell_reverse_complex_eigenvalue_at_the_edge():
    """a site on the lattice edge whose timelike eigenvalue is complex keeps
    eps = Re(lambda) (no fallback there): its derivative against central
    differences"""
    rng = np.random.default_rng(77)
    N, a, dtau, it = 64, 0.25, 0.094, 17
    tau = it * dtau * a
    ok = 0
    for trial in range(40):
        T = np.array([10.0, 4.0, 3.0, 2.0 / tau ** 2, 8.0, 1.0, 0.2 / tau, 0.5, 0.1 / tau, 0.1 / tau])
        T = T * (1 + 0.3 * rng.normal(size=10))
        wr, wi = np.zeros(4), np.zeros(4)
        vr, vi = np.zeros(16), np.zeros(16)
        g = np.diag([1.0, -1.0, -1.0, -tau * tau])
        idx = [(0, 0), (1, 1), (2, 2), (3, 3), (0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
        M = np.zeros((4, 4))
        for k, (i, j) in enumerate(idx):
            M[i, j] = M[j, i] = T[k]
        A = np.ascontiguousarray((M @ g).reshape(-1))
        lib().ipgl_eig4_host(A, wr, wi, vr, vi)
        if not (np.abs(wi) > 1e-3).any():
            continue
        e0, u0 = C.c_double(), C.c_double()
        lib().ipgl_landau_cell_host(T, N, a, dtau, it, 0, 7, C.byref(e0), C.byref(u0))
        if not any(abs(e0.value - wr[k]) == 0 and wi[k] != 0 for k in range(4)):
            continue
        ok += 1
        Tb = np.zeros(10)
        lib().ipgl_landau_adj_host(T, N, a, dtau, it, 0, 7, 1.0, 0.0, Tb)

        def eps(Tx):
            e, u = C.c_double(), C.c_double()
            lib().ipgl_landau_cell_host(np.ascontiguousarray(Tx), N, a, dtau, it, 0, 7, C.byref(e), C.byref(u))
            return e.value

        d = rng.normal(size=10) * np.abs(T)
        h = 1e-7
        fd = (eps(T + h * d) - eps(T - h * d)) / (2 * h)
        assert abs(float(Tb @ d) - fd) <= 1e-6 * max(1.0, abs(fd)), (trial, float(Tb @ d), fd)
    assert ok >= 3
