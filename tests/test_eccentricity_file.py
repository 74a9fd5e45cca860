"""The maxX, maxY columns of eccentricities<id>.dat (Evolution::eccentricity,
Evolution.cpp:1492-1640) through the library's host entry point, against a
numpy restatement of the reference's loop on synthetic eps, u^tau fields:
the row and column through the centroid shifted by floor(<x>/a + 1e-11),
the interior 2 <= ix, iy < N-2, the last site with positive weight, and the
zero weight where eps < cutoff.  No GPU."""
import math

import numpy as np
import pytest

import ipglasma_amd as I


def _extents(eps, utau, N, L, avx, avy, cutoff):
    a = L / N
    xshift = int(math.floor(avx / a + 0.00000000001))
    yshift = int(math.floor(avy / a + 0.00000000001))
    mx = my = 0.0
    for ix in range(2, N - 2):
        x = -L / 2.0 + a * ix - avx
        for iy in range(2, N - 2):
            y = -L / 2.0 + a * iy - avy
            pos = ix * N + iy
            w = 0.0 if eps[pos] < cutoff else eps[pos] * utau[pos]
            if w > cutoff and iy == N // 2 + yshift:
                mx = x
            if w > cutoff and ix == N // 2 + xshift:
                my = y
    return mx, my


@pytest.mark.parametrize("N,L,cx,cy,sx,sy,cutoff", [
    (32, 10.0, 0.7, -1.1, 2.0, 1.2, 0.0),
    (48, 12.0, -2.3, 0.4, 1.0, 3.0, 0.0),
    (32, 10.0, 0.0, 0.0, 1.5, 1.5, 0.05),
    (64, 20.0, 3.1, 2.9, 2.5, 0.8, 0.0),
])
def test_extents_match_reference_loop(N, L, cx, cy, sx, sy, cutoff):
    a = L / N
    x = -L / 2 + a * np.arange(N)
    X, Y = np.meshgrid(x, x, indexing="ij")
    # an elliptic blob cut off at 2.5 sigma, with negative eps outside
    r2 = ((X - cx) / sx) ** 2 + ((Y - cy) / sy) ** 2
    eps = np.where(r2 < 6.25, np.exp(-r2 / 2), -1e-3).reshape(-1)
    utau = (1.0 + 0.1 * np.sqrt(r2)).reshape(-1)
    w = np.where(eps < 0, 0.0, eps * utau)
    avx, avy = float((X.reshape(-1) * w).sum() / w.sum()), float((Y.reshape(-1) * w).sum() / w.sum())
    got = I.eccentricity_extents(eps, utau, L, avx, avy, cutoff)
    want = _extents(eps, utau, N, L, avx, avy, cutoff)
    assert got == want
    assert got[0] > 0 and got[1] > 0
