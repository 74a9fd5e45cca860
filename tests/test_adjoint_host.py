"""The per-site reverse-mode functions of the leapfrog step and of T^tautau
(ipglasma_amd/csrc/adjoint.hpp), built for the host, against central finite
differences of their own forward functions: for random inputs x, output
cotangent yb and direction d, Re<xb, d> = d/dh Re tr(yb^+ f(x + h d))."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host", "adjoint_host.cpp")
OUT = os.path.join(HERE, "_build", "libadjoint_host.so")
HIPCC = "/opt/rocm/bin/hipcc"
_lib = None
dp = np.ctypeslib.ndpointer(np.float64, flags="C")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
            pytest.skip("hipcc not available")
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        subprocess.run([HIPCC if os.path.exists(HIPCC) else "hipcc", "-x", "hip", "--offload-host-only", "-O2",
                        "-std=c++17", "-fPIC", "-shared", "-o", OUT, SRC], check=True, capture_output=True)
        L = C.CDLL(OUT)
        L.adjh_pi.argtypes = [dp, C.c_double, dp, dp, dp]
        L.adjh_e.argtypes = [dp, C.c_double, C.c_double, C.c_double, dp, dp, dp]
        L.adjh_u.argtypes = [dp, C.c_double, dp, dp, dp]
        L.adjh_tt.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_double, dp]
        L.adjh_tt.restype = C.c_double
        L.adjh_newton.argtypes = [dp] * 6
        L.adjh_acc.argtypes = [dp] * 4
        L.adjh_expq.argtypes = [dp] * 4
        L.adjh_expq_value.argtypes = [dp] * 2
        L.adjh_expq_forward.argtypes = [dp] * 2
        L.adjh_expq_apply.argtypes = [dp] * 3
        L.adjh_gen_det.argtypes = [dp]
        L.adjh_gen_det.restype = C.c_double
        _lib = L
    return _lib


def cm(a):
    return np.ascontiguousarray(np.stack([a.real, a.imag], -1).reshape(-1))


def rand_link(rng, n):
    return rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3)) * 0.5 + 2 * np.eye(3)


def rand_herm(rng, n, s=1.0):
    a = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    return s * (a + a.conj().transpose(0, 2, 1)) / 2


def dot(xb, d):
    """Re sum tr(xb^+ d) for stacked complex matrices as flat arrays"""
    return float(np.dot(xb, d))


def fd_check(f, x, herm_mask, nout, rng, h=1e-6, trials=6):
    """f(x_flat) -> y_flat (nout matrices); compares the reverse pass"""
    for _ in range(trials):
        yb = cm(rng.normal(size=(nout, 3, 3)) + 1j * rng.normal(size=(nout, 3, 3)))
        y, xb = f(x, yb)
        nin = len(herm_mask)
        d = []
        for k in range(nin):
            dd = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
            if herm_mask[k]:
                dd = (dd + dd.conj().T) / 2
            d.append(dd)
        d = cm(np.array(d))
        yp, _ = f(x + h * d, yb)
        ym, _ = f(x - h * d, yb)
        fd = (np.dot(yb, yp) - np.dot(yb, ym)) / (2 * h)
        # the Hermitian inputs take the Hermitian projection of their cotangent
        assert abs(dot(xb, d) - fd) <= 1e-6 * max(1.0, abs(fd)), (dot(xb, d), fd)


def test_pi_update_adjoint():
    rng = np.random.default_rng(1)
    L = lib()
    links = rand_link(rng, 4)
    hs = rand_herm(rng, 6)
    x = cm(np.concatenate([links, hs]))
    mask = [0, 0, 0, 0, 1, 1, 1, 1, 1, 1]

    def f(x, yb):
        y = np.zeros(18)
        xb = np.zeros(180)
        L.adjh_pi(np.ascontiguousarray(x), 0.37, np.ascontiguousarray(yb), y, xb)
        return y, xb
    fd_check(f, x, mask, 1, rng)


def test_e_update_adjoint():
    rng = np.random.default_rng(2)
    L = lib()
    links = rand_link(rng, 10)
    hs = rand_herm(rng, 5)
    x = cm(np.concatenate([links, hs]))
    mask = [0] * 10 + [1] * 5

    def f(x, yb):
        y = np.zeros(36)
        xb = np.zeros(270)
        L.adjh_e(np.ascontiguousarray(x), 1.7, 0.1, 2.0, np.ascontiguousarray(yb), y, xb)
        return y, xb
    fd_check(f, x, mask, 2, rng)


def test_link_update_adjoint():
    rng = np.random.default_rng(3)
    L = lib()
    x = cm(np.concatenate([rand_link(rng, 1), rand_herm(rng, 1, 0.3)]))

    def f(x, yb):
        y = np.zeros(18)
        xb = np.zeros(36)
        L.adjh_u(np.ascontiguousarray(x), 0.8, np.ascontiguousarray(yb), y, xb)
        return y, xb
    fd_check(f, x, [0, 1], 1, rng)


@pytest.mark.parametrize("plaq", [0, 1])
def test_tautau_adjoint(plaq):
    rng = np.random.default_rng(4 + plaq)
    L = lib()
    # e1 e1py e2 e2px p ppx ppy ppxy (Hermitian), ux uy uy_px ux_py (links), 4 phi
    x = cm(np.concatenate([rand_herm(rng, 8), rand_link(rng, 4), rand_herm(rng, 4)]))
    mask = [1] * 8 + [0] * 4 + [1] * 4
    for _ in range(6):
        yb = rng.normal()
        xb = np.zeros(16 * 18)
        L.adjh_tt(np.ascontiguousarray(x), plaq, 1.3, 2.0, yb, xb)
        d = []
        for k in range(16):
            dd = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
            if mask[k]:
                dd = (dd + dd.conj().T) / 2
            d.append(dd)
        d = cm(np.array(d))
        h = 1e-6
        scratch = np.zeros(16 * 18)
        fp = L.adjh_tt(np.ascontiguousarray(x + h * d), plaq, 1.3, 2.0, yb, scratch)
        fm = L.adjh_tt(np.ascontiguousarray(x - h * d), plaq, 1.3, 2.0, yb, scratch)
        fd = yb * (fp - fm) / (2 * h)
        assert abs(np.dot(xb, d) - fd) <= 1e-6 * max(1.0, abs(fd)), (np.dot(xb, d), fd)


def test_pi_assembly_adjoint():
    rng = np.random.default_rng(7)
    L = lib()
    x = cm(np.concatenate([rand_link(rng, 4)]))

    def f(x, yb):
        y = np.zeros(18)
        xb = np.zeros(72)
        L.adjh_acc(np.ascontiguousarray(x), np.ascontiguousarray(yb), y, xb)
        return y, xb
    fd_check(f, x, [0, 0, 0, 0], 1, rng)


def test_root_finder_implicit_adjoint():
    """against central differences of the oracle's root finder (an
    independent forward implementation), at unitary U1, U2"""
    import scipy.linalg as sl
    import _oracle as O
    rng = np.random.default_rng(8)
    L = lib()

    def su3(scale):
        a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        h = scale * (a + a.conj().T) / 2
        return sl.expm(1j * (h - np.trace(h) / 3 * np.eye(3)))

    for trial in range(5):
        U1, U2 = su3(0.6), su3(0.6)
        U, st = O.newton_link(U1, U2)
        assert st.converged
        Ub = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        b1 = np.zeros(18)
        b2 = np.zeros(18)
        L.adjh_newton(cm(U1), cm(U2), cm(U), cm(Ub), b1, b2)
        d1 = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        d2 = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        h = 1e-6

        def loss(a, b):
            Us, s2 = O.newton_link(a, b)
            assert s2.converged
            return float(np.real(np.trace(Ub.conj().T @ Us)))
        fd = (loss(U1 + h * d1, U2 + h * d2) - loss(U1 - h * d1, U2 - h * d2)) / (2 * h)
        pred = float(np.dot(b1, cm(d1)) + np.dot(b2, cm(d2)))
        assert abs(pred - fd) <= 1e-5 * max(1.0, abs(fd)), (trial, pred, fd)


def test_su3_exponential_adjoint_against_frechet():
    """Qb_a = Re tr(Eb^+ dE/dQ_a) against scipy's Frechet derivative of expm"""
    import scipy.linalg as sl
    rng = np.random.default_rng(9)
    L = lib()
    lam = np.zeros((8, 3, 3), complex)
    lam[0][0, 1] = lam[0][1, 0] = 0.5
    lam[1][0, 1], lam[1][1, 0] = -0.5j, 0.5j
    lam[2][0, 0], lam[2][1, 1] = 0.5, -0.5
    lam[3][0, 2] = lam[3][2, 0] = 0.5
    lam[4][0, 2], lam[4][2, 0] = -0.5j, 0.5j
    lam[5][1, 2] = lam[5][2, 1] = 0.5
    lam[6][1, 2], lam[6][2, 1] = -0.5j, 0.5j
    lam[7] = np.diag([1.0, 1.0, -2.0]) / (2 * np.sqrt(3))
    # random charges across both branches of the coefficient functions (the
    # series below c1 = tr(Q^2)/2 = 0.5, the closed form above), then
    # degenerate spectra: Q along t^8 (c0 = -c0max and +c0max), along t^3
    # (c0 = 0), nearly degenerate, and tiny
    cases = [rng.normal(size=8) * sc for sc in (1e-4, 0.05, 0.3, 0.45, 0.55, 0.7, 1.5, 3.0, 6.0)]
    exact_degenerate = []
    for q8 in (-2.0, 0.9, 4.0):
        exact_degenerate.append(len(cases))
        cases.append(np.eye(8)[7] * q8)
    cases.append(np.eye(8)[2] * 1.7)
    near_degenerate = [len(cases), len(cases) + 1]
    cases.append(np.eye(8)[7] * 2.5 + rng.normal(size=8) * 1e-7)
    cases.append(np.eye(8)[7] * -1.3 + rng.normal(size=8) * 1e-4)
    for n, Q in enumerate(cases):
        scale = float(np.abs(Q).max())
        Eb = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        e = np.zeros(18)
        qb = np.zeros(8)
        L.adjh_expq(Q, cm(Eb), e, qb)
        H = 1j * np.einsum("a,aij->ij", Q, lam)
        E = sl.expm(H)
        # the reference's closed form returns the identity at exactly
        # degenerate spectra (w = 0 or |c0| > c0max by round-off gives NaN
        # coefficients, zeroed, Matrix.cpp:338-345, Init.cpp:3487-3489); the
        # reverse mode differentiates the exponential itself
        if n not in exact_degenerate:
            assert np.allclose(e[0::2] + 1j * e[1::2], E.reshape(-1), atol=1e-13), (n, scale)
        ref = np.array([np.real(np.trace(Eb.conj().T @ sl.expm_frechet(H, 1j * lam[a], compute_expm=False)))
                        for a in range(8)])
        assert np.allclose(qb, ref, rtol=1e-10, atol=1e-11 * np.abs(ref).max()), (scale, qb, ref)
        # the exponential rebuilt from the reverse's coefficients (what the
        # reverse setV pass uses to walk the Wilson lines back): the matrix
        # exponential to round-off everywhere, and the forward's exp_iQt to
        # round-off away from (nearly) degenerate spectra, where the closed
        # form exp_iQt follows loses digits (1.7e-9 at a 1e-7 splitting of
        # 2.5 t^8; such a Q among Gaussian colour charges has probability
        # ~ splitting^3)
        ev = np.zeros(18)
        L.adjh_expq_value(Q, ev)
        ev = ev[0::2] + 1j * ev[1::2]
        assert np.abs(ev - E.reshape(-1)).max() <= 2e-15 * max(1.0, scale), (n, scale)
        if n not in exact_degenerate + near_degenerate:
            assert np.abs(ev - (e[0::2] + 1j * e[1::2])).max() <= 1e-13 * max(1.0, scale), (n, scale)
        # the forward setV exponential: the series below c1 = 0.5, above it
        # the series of Q / 2^s applied 2^s times; the matrix exponential to
        # round-off everywhere, degenerate spectra included (where exp_iQt
        # returns the identity or loses digits)
        ef = np.zeros(18)
        L.adjh_expq_forward(Q, ef)
        ef = ef[0::2] + 1j * ef[1::2]
        c1 = 0.25 * float(np.sum(Q * Q))
        assert np.abs(ef - E.reshape(-1)).max() <= 1e-14 * max(1.0, scale), (n, scale, c1)
        # applied to a Wilson line: E V to round-off; det Q from the
        # components against the determinant of the matrix
        V = sl.expm(1j * np.einsum("a,aij->ij", rng.normal(size=8), lam))
        va = np.zeros(18)
        L.adjh_expq_apply(Q, cm(V), va)
        assert np.abs(va[0::2] + 1j * va[1::2] - (E @ V).reshape(-1)).max() <= 1e-14 * max(1.0, scale), (n, c1)
        Hq = np.einsum("a,aij->ij", Q, lam)
        assert abs(L.adjh_gen_det(Q) - np.linalg.det(Hq).real) <= 1e-15 * max(1.0, scale) ** 3, n
    # tiny charges: the series keeps the exponential (1 + iQ - Q^2/2 to
    # round-off) where the reference's closed form, in this host build as in
    # the oracle, is off by up to |Q| (catastrophic cancellation)
    for sc in (1e-8, 1e-9, 1e-11):
        Q = rng.normal(size=8) * sc
        H = 1j * np.einsum("a,aij->ij", Q, lam)
        ef = np.zeros(18)
        L.adjh_expq_forward(Q, ef)
        ef = (ef[0::2] + 1j * ef[1::2]).reshape(3, 3)
        assert np.abs(ef - (np.eye(3) + H + H @ H / 2)).max() <= 1e-17, sc
        e = np.zeros(18)
        qb = np.zeros(8)
        L.adjh_expq(Q, cm(np.eye(3) + 0j), e, qb)
        u = np.zeros(18)
        O.lib().orc_u_from_exponent(np.ascontiguousarray(Q), u)
        assert np.abs(e - u).max() <= 1e-20, sc  # the two closed forms agree with each other
