"""The per-site gauge-fixing code that the GPU runs
(ipglasma_amd/csrc/gaugefix.hpp), built for the host, against the oracle's
restatement of Matrix::expm and Matrix::reu (Matrix.cpp:431-560,
Matrix.h:67-90); and the library's formulation of the spectra (unshifted
transforms of the nine real planes of a Hermitian field packed into five
complex planes, sum_m w_m (|Z_m(k)|^2 + |Z_m(-k)|^2) / 2 with w = 1, 1, 2, 2,
2) written in numpy against the oracle's shifted full-matrix form."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import _oracle as O
from test_oracle_gaugefix import T, alg

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host", "gaugefix_host.cpp")
OUT = os.path.join(HERE, "_build", "libgaugefix_host.so")
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
        L.gfh_expm.argtypes = [dp, dp]
        L.gfh_expm.restype = C.c_int
        L.gfh_reunitarize.argtypes = [dp, dp]
        L.gfh_rotation.argtypes = [dp, dp]
        _lib = L
    return _lib


def test_expm_and_reunitarize_match_oracle():
    rng = np.random.default_rng(11)
    for scale in (1e-4, 0.3, 0.7, 2.0, 9.0):
        for _ in range(30):
            A = 1j * alg(rng.normal(size=8)) * scale
            out = np.zeros(18)
            assert lib().gfh_expm(O.to_c(A), out) == 0
            np.testing.assert_allclose(O.from_c(out), O.expm_pade(A), rtol=0, atol=1e-14 * max(1, scale))
            M = O.from_c(out) + 0.01 * rng.normal(size=(3, 3))
            r = np.zeros(18)
            lib().gfh_reunitarize(O.to_c(M), r)
            np.testing.assert_allclose(O.from_c(r), O.reunitarize(M), rtol=0, atol=2e-15)


def test_rotation_is_reunitarized_exponential():
    rng = np.random.default_rng(12)
    for scale in (1e-6, 0.05, 1.0, 4.0):
        for _ in range(30):
            chi = rng.normal(size=8) * scale
            out = np.zeros(18)
            lib().gfh_rotation(np.ascontiguousarray(chi), out)
            ref = O.reunitarize(O.expm_pade(1j * alg(chi)))
            np.testing.assert_allclose(O.from_c(out), ref, rtol=0, atol=1e-14)
    out = np.zeros(18)
    lib().gfh_rotation(np.full(8, np.nan), out)
    assert np.array_equal(O.from_c(out), np.eye(3))


def packed_spectrum(F, N):
    """the library's per-mode Re tr(F(k) F(-k)) from the five packed planes"""
    M = O.from_c(F).reshape(N, N, 3, 3)
    planes = [M[..., 0, 0].real + 1j * M[..., 1, 1].real, M[..., 2, 2].real + 0j,
              M[..., 0, 1].real + 1j * M[..., 0, 1].imag, M[..., 0, 2].real + 1j * M[..., 0, 2].imag,
              M[..., 1, 2].real + 1j * M[..., 1, 2].imag]
    w = (1., 1., 2., 2., 2.)
    v = np.zeros((N, N))
    for wm, z in zip(w, planes):
        Z = np.fft.fft2(z)
        Zm = np.roll(Z[::-1, ::-1], 1, axis=(0, 1))  # Z(-k)
        v += wm * 0.5 * (np.abs(Z) ** 2 + np.abs(Zm) ** 2)
    return v


def test_packed_unshifted_spectrum_equals_shifted_matrix_form():
    N, L, g, dtau, it = 16, 3.0, 2.0, 0.1, 2
    rng = np.random.default_rng(13)
    F = O.to_c(alg(rng.normal(size=(N * N, 8))) + np.eye(3) * rng.normal(size=(N * N, 1, 1)))
    c = g * g / ((it - 0.5) * dtau)
    v = packed_spectrum(F, N)
    q = np.arange(N)
    s2 = 4 * np.sin(np.pi * q / N) ** 2
    kt2 = s2[:, None] + s2[None, :]
    keep = (q[:, None] != N // 2) & (q[None, :] != N // 2) & (kt2 > 0)
    nkt = np.where(keep, 2 / np.sqrt(np.where(keep, kt2, 1)) / (N * N) * c * v, 0.)
    z = np.zeros_like(F)
    ref = O.multiplicity(N, L, g, dtau, it, F, z, z)
    assert nkt.sum() == pytest.approx(ref[0], rel=1e-12)
    a = L / N
    assert (nkt * np.sqrt(kt2) * 0.1973269718 / a).sum() == pytest.approx(ref[1], rel=1e-12)
