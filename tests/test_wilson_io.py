"""Wilson-line files (the writeInitialWilsonLines output of setV,
Init.cpp:1857-1980, and Init::readV, Init.cpp:1988-2301) through the
library's host entry points, against a numpy restatement of the reference's
loops: the byte layout of the binary file, the text layout, the row shift by
-+b/2 on reading, and the error cases.  No GPU: these entry points are host
code."""
import math
import struct

import numpy as np
import pytest

import ipglasma_amd as I


def _random_v(N, seed):
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(N * N, 18))
    V[::7] *= -1e-300  # tiny and negative values keep their digits
    V[3, 5] = -0.0
    return V


def test_binary_layout_matches_reference_loops(tmp_path):
    """Header int N, int Nc, double L, double a, double rapidity; then for
    ix, iy (iy fastest) the cell N*iy + ix, entries row by row as (re, im)
    (Init.cpp:1913-1961)."""
    N, L, rap = 6, 3.0, 0.25
    V = _random_v(N, 1)
    f = tmp_path / "V-2"
    I.write_wilson_line_file(f, V, L, rap, fmt=2)
    want = struct.pack("<iiddd", N, 3, L, L / N, rap)
    for ix in range(N):
        for iy in range(N):
            want += V[N * iy + ix].astype("<f8").tobytes()
    assert f.read_bytes() == want


def test_text_layout_matches_reference_loops(tmp_path):
    """"ix iy" and Matrix::MatrixToString (Matrix.cpp:606-627: column by
    column, precision 15) per site in (ix, iy) order, an empty line after
    each ix (Init.cpp:1883-1893)."""
    N = 3
    V = _random_v(N, 2)
    f = tmp_path / "V-2.txt"
    I.write_wilson_line_file(f, V, 1.5, 0.0, fmt=1)
    lines = []
    for ix in range(N):
        for iy in range(N):
            m = V[ix * N + iy]
            nums = []
            for c in range(3):
                for r in range(3):
                    k = 3 * r + c
                    nums += ["%.15g" % m[2 * k], "%.15g" % m[2 * k + 1]]
            lines.append(f"{ix} {iy} " + " ".join(nums))
        lines.append("")
    assert f.read_text() == "\n".join(lines) + "\n"


@pytest.mark.parametrize("nucleus", [0, 1])
def test_binary_round_trip_is_exact(tmp_path, nucleus):
    N, L = 8, 4.0
    V = _random_v(N, 3)
    f = tmp_path / "V-1"
    I.write_wilson_line_file(f, V, L, 0.0, fmt=2)
    W = I.read_wilson_line_file(f, N, L, b=0.0, nucleus=nucleus, fmt=2)
    assert np.array_equal(W.view(np.uint64), V.view(np.uint64))


def test_text_round_trip_transposes(tmp_path):
    """The text writer goes column by column, the reader row by row
    (Init.cpp:2058-2071), so a round trip gives V^T, as in the reference."""
    N, L = 5, 2.5
    V = _random_v(N, 4)
    f = tmp_path / "V-1.txt"
    I.write_wilson_line_file(f, V, L, 0.0, fmt=1)
    W = I.read_wilson_line_file(f, N, L, b=0.0, nucleus=0, fmt=1)
    cV = (V[:, 0::2] + 1j * V[:, 1::2]).reshape(-1, 3, 3)
    cW = (W[:, 0::2] + 1j * W[:, 1::2]).reshape(-1, 3, 3)
    np.testing.assert_allclose(cW, cV.transpose(0, 2, 1), rtol=1e-14, atol=1e-300)


def _c_round(x):
    return math.copysign(math.floor(abs(x) + 0.5), x)


def _identity(N):
    V = np.zeros((N * N, 18))
    V[:, [0, 8, 16]] = 1.0
    return V


@pytest.mark.parametrize("b", [1.3, -0.7, 0.0])
@pytest.mark.parametrize("nucleus", [0, 1])
def test_text_reader_row_shift(tmp_path, b, nucleus):
    """Line (i, j) lands on row int((a i -+ b/2) / a), truncated toward zero;
    rows outside the lattice are dropped (Init.cpp:2073-2080, 2117-2124)."""
    N, L = 8, 4.0
    V = _random_v(N, 5)
    f = tmp_path / "V.txt"
    I.write_wilson_line_file(f, V, L, 0.0, fmt=1)
    W = I.read_wilson_line_file(f, N, L, b=b, nucleus=nucleus, fmt=1)
    a = L / float(N)
    want = _identity(N)
    for i in range(N):
        for j in range(N):
            m = V[i * N + j].reshape(9, 2)
            # the numbers of line (i, j) in file order, taken row by row
            col = np.array([m[3 * r + c] for c in range(3) for r in range(3)])
            x = a * i - b / 2.0 if nucleus == 0 else a * i + b / 2.0
            ix = int(x / a)
            if ix < 0 or ix >= N:
                continue
            want[ix * N + j] = col.reshape(-1)
    np.testing.assert_allclose(W, want, rtol=1e-14, atol=1e-300)


@pytest.mark.parametrize("b", [1.3, -0.7])
@pytest.mark.parametrize("nucleus", [0, 1])
def test_binary_reader_row_shift(tmp_path, b, nucleus):
    """File pair t: position t // 9 = slow * N + fast, row
    round((a fast -+ b/2) / a) with a from the file's L, site N * row + slow,
    dropped outside [0, N^2) (Init.cpp:2167-2215, 2250-2290)."""
    N, L = 8, 4.0
    V = _random_v(N, 6)
    f = tmp_path / "V"
    I.write_wilson_line_file(f, V, L, 0.0, fmt=2)
    W = I.read_wilson_line_file(f, N, L, b=b, nucleus=nucleus, fmt=2)
    raw = np.frombuffer(f.read_bytes()[32:], dtype="<f8").reshape(-1, 2)
    a = L / float(N)
    want = _identity(N)
    for t in range(raw.shape[0]):
        pos, e = divmod(t, 9)
        slow, fast = divmod(pos, N)
        x = a * fast - b / 2.0 if nucleus == 0 else a * fast + b / 2.0
        ix = int(_c_round(x / a))
        site = N * ix + slow
        if site < 0 or site >= N * N:
            continue
        want[site, 2 * e:2 * e + 2] = raw[t]
    assert np.array_equal(W, want)


def test_reader_errors(tmp_path):
    N, L = 4, 2.0
    V = _random_v(N, 7)
    f = tmp_path / "V"
    I.write_wilson_line_file(f, V, L, 0.0, fmt=2)
    with pytest.raises(I.IPGError, match="wrong lattice size"):
        I.read_wilson_line_file(f, 8, L, fmt=2)
    with pytest.raises(I.IPGError, match="grid length"):
        I.read_wilson_line_file(f, N, L + 1e-3, fmt=2)
    I.read_wilson_line_file(f, N, L + 1e-6, fmt=2)  # within the 1e-5 the reference allows
    with pytest.raises(I.IPGError, match="does not exist|not found"):
        I.read_wilson_line_file(tmp_path / "missing", N, L, fmt=2)
    with pytest.raises(I.IPGError, match="not found"):
        I.read_wilson_line_file(tmp_path / "missing.txt", N, L, fmt=1)
    with pytest.raises(I.IPGError, match="unknown Wilson-line format"):
        I.read_wilson_line_file(f, N, L, fmt=3)
    with pytest.raises(I.IPGError, match="unknown Wilson-line format"):
        I.write_wilson_line_file(tmp_path / "x", V, L, fmt=0)


def test_reference_file_names():
    """V-<eventId + 2 seed mpiSize> and V-<eventId + (1 + 2 seed) mpiSize>,
    .txt for text (Init.cpp:1863-1877)."""
    assert I.wilson_line_file_names(5, event_id=2, mpi_size=3, fmt=1) == ("V-32.txt", "V-35.txt")
    assert I.wilson_line_file_names(1) == ("V-2", "V-3")
