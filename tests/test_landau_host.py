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
