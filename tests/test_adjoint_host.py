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
