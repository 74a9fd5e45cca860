"""The LDS transform of the colour-charge solve (ipglasma_amd/csrc/fft_lds.hpp),
built for the host, against numpy.fft: every supported length, both signs,
three rows in the padded LDS layout (every slot distinct, inside the
allocation, for the row counts the kernels use), and the stage count of the
plan."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host", "fft_host.cpp")
OUT = os.path.join(HERE, "_build", "libfft_host.so")
HIPCC = "/opt/rocm/bin/hipcc"
_lib = None

SIZES = [16, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024]
STAGES = {16: 2, 32: 2, 48: 3, 64: 2, 96: 3, 128: 3, 192: 3, 256: 3, 384: 4, 512: 3, 768: 4, 1024: 4}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
            pytest.skip("hipcc not available")
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        subprocess.run([HIPCC if os.path.exists(HIPCC) else "hipcc", "-x", "hip", "--offload-host-only", "-O2",
                        "-std=c++17", "-fPIC", "-shared", "-o", OUT, SRC], check=True, capture_output=True)
        L = C.CDLL(OUT)
        L.ipgfft_host.argtypes = [C.c_int, np.ctypeslib.ndpointer(np.float64, flags="C"), C.c_int, C.c_int]
        L.ipgfft_host.restype = C.c_int
        _lib = L
    return _lib


@pytest.mark.parametrize("N", SIZES)
def test_lds_fft_matches_numpy(N):
    rng = np.random.default_rng(N)
    x = rng.normal(size=(3, N)) + 1j * rng.normal(size=(3, N))
    for inverse in (0, 1):
        buf = np.ascontiguousarray(np.stack([x.real, x.imag], -1).reshape(-1))
        ns = lib().ipgfft_host(N, buf, 3, inverse)
        assert ns == STAGES[N]
        y = (buf[0::2] + 1j * buf[1::2]).reshape(3, N)
        ref = np.fft.fft(x, axis=1) if not inverse else np.fft.ifft(x, axis=1) * N
        assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max(), (N, inverse)
