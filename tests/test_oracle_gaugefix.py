"""Pins of the oracle's Coulomb gauge fixing and gluon spectra
(GaugeFix.cpp:17-210, Evolution.cpp:2284-3188, Matrix.cpp:431-560,
Matrix.h:67-90, FFT.cpp:260-372) by what the mathematics fixes:

  - the Pade exponential against scipy.linalg.expm;
  - reunitarisation against a QR factorisation of the first two rows, with
    det = 1 and exact SU(3) input returned unchanged;
  - the quadrant-swapped transform against numpy's
    fftshift(fft2(ifftshift(.)));
  - one gauge-fixing step on a weak field divides div A by 4 (the residual
    by 16): with U = exp(i A), chi = (div A)/2 and the -1.5/k^2 filter give
    div A' = div A - k^2 (1.5/k^2) (div A)/2 = div A / 4, so a wrong filter
    constant, sign or normalisation moves the ratio;
  - gauge-invariant traces kept site by site, a Coulomb residual below 1e-9
    and unitary links after the fix;
  - single-mode E and pi fields against the closed-form multiplicity
    dN/dy = c A^2 N^2 / (2 omega) of a cosine wave A cos(k.x) t^3, its energy
    and its k_T bin; modes on the excluded line k = pi carry nothing.
"""
import numpy as np
import pytest
import scipy.linalg

import _oracle as O

HBARC = 0.1973269718

T = [np.zeros((3, 3), complex) for _ in range(8)]
T[0][0, 1] = T[0][1, 0] = 0.5
T[1][0, 1], T[1][1, 0] = -0.5j, 0.5j
T[2][0, 0], T[2][1, 1] = 0.5, -0.5
T[3][0, 2] = T[3][2, 0] = 0.5
T[4][0, 2], T[4][2, 0] = -0.5j, 0.5j
T[5][1, 2] = T[5][2, 1] = 0.5
T[6][1, 2], T[6][2, 1] = -0.5j, 0.5j
T[7][0, 0] = T[7][1, 1] = 1 / (2 * np.sqrt(3))
T[7][2, 2] = -1 / np.sqrt(3)
T = np.array(T)


def alg(c):
    """sum_a c^a t^a for c [..., 8]"""
    return np.einsum("...a,aij->...ij", c, T)


def test_expm_pade_against_scipy():
    rng = np.random.default_rng(5)
    worst = 0.0
    for scale in (1e-3, 0.1, 0.4, 0.6, 1.0, 3.0, 10.0, 25.0):
        for _ in range(20):
            H = alg(rng.normal(size=8)) * scale
            for A in (1j * H, (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))) * scale / 3):
                ref = scipy.linalg.expm(A)
                got = O.expm_pade(A)
                worst = max(worst, np.abs(got - ref).max() / max(1.0, np.abs(ref).max()))
    assert worst < 1e-8, worst
    assert np.array_equal(O.expm_pade(np.zeros((3, 3))), np.eye(3))


def test_reunitarize_against_qr():
    rng = np.random.default_rng(6)
    for _ in range(50):
        U = scipy.linalg.expm(1j * alg(rng.normal(size=8)))
        M = U + 0.05 * (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
        R = O.reunitarize(M)
        q, r = np.linalg.qr(M[:2].T)
        ph = np.diag(r) / np.abs(np.diag(r))
        assert np.allclose(R[0], q[:, 0] * ph[0], atol=1e-13)
        assert np.allclose(R[1], q[:, 1] * ph[1], atol=1e-13)
        assert np.allclose(R @ R.conj().T, np.eye(3), atol=1e-13)
        assert abs(np.linalg.det(R) - 1) < 1e-13
        assert np.allclose(O.reunitarize(U), U, atol=1e-14)


def test_fftn_shifted_against_numpy():
    rng = np.random.default_rng(7)
    for N in (8, 12, 16):
        f = rng.normal(size=(N * N, 18))
        z = (f[:, 0::2] + 1j * f[:, 1::2]).reshape(N, N, 9)
        for sign in (1, -1):
            g = O.fftn_shifted(f, sign)
            got = (g[:, 0::2] + 1j * g[:, 1::2]).reshape(N, N, 9)
            op = np.fft.fft2 if sign > 0 else np.fft.ifft2
            ref = np.fft.fftshift(op(np.fft.ifftshift(z, axes=(0, 1)), axes=(0, 1)), axes=(0, 1))
            assert np.allclose(got, ref, atol=1e-12 * np.abs(ref).max())


def links_from(A):
    """A [N, N, 2, 8] -> Ux, Uy [S, 18] = exp(i A^a t^a)"""
    N = A.shape[0]
    U = np.array([scipy.linalg.expm(1j * alg(a)) for a in A.reshape(-1, 8)]).reshape(N * N, 2, 3, 3)
    return O.to_c(U[:, 0]), O.to_c(U[:, 1])


def state_with(Ux, Uy, N, rng=None):
    st = O.new_state(N)
    st["Ux"][:] = Ux
    st["Uy"][:] = Uy
    if rng is not None:
        for k in ("E1", "E2", "phi", "pi"):
            st[k][:] = O.to_c(alg(rng.normal(size=(N * N, 8))))
    return st


def test_gauge_fix_linear_step_divides_divergence_by_four():
    N = 16
    rng = np.random.default_rng(8)
    st = state_with(*links_from(1e-3 * rng.normal(size=(N, N, 2, 8))), N)
    it, res, h = O.gauge_fix(N, st, max_iter=3, history=True)
    assert len(h) >= 2 and h[0] > 1e-9
    ratio = h[1] / h[0]
    assert ratio == pytest.approx(1 / 16, rel=2e-2), ratio


def plaquettes(st, N):
    Ux = O.from_c(st["Ux"]).reshape(N, N, 3, 3)
    Uy = O.from_c(st["Uy"]).reshape(N, N, 3, 3)
    P = Ux @ np.roll(Uy, -1, 0) @ np.roll(Ux, -1, 1).conj().swapaxes(-1, -2) @ Uy.conj().swapaxes(-1, -2)
    return np.trace(P, axis1=-2, axis2=-1)


def test_gauge_fix_keeps_invariants_and_reaches_coulomb_gauge():
    N = 16
    rng = np.random.default_rng(9)
    # a smooth field in a random (moderate) gauge
    x = np.arange(N)
    smooth = 0.3 * np.sin(2 * np.pi * x / N)[:, None, None, None] * rng.normal(size=(1, 1, 2, 8))
    Ux, Uy = links_from(smooth + 0.02 * rng.normal(size=(N, N, 2, 8)))
    g = np.array([scipy.linalg.expm(1j * alg(c)) for c in 0.5 * rng.normal(size=(N * N, 8))]).reshape(N, N, 3, 3)
    ux, uy = O.from_c(Ux).reshape(N, N, 3, 3), O.from_c(Uy).reshape(N, N, 3, 3)
    gd = lambda m: m.conj().swapaxes(-1, -2)
    ux = g @ ux @ gd(np.roll(g, -1, 0))
    uy = g @ uy @ gd(np.roll(g, -1, 1))
    st = state_with(O.to_c(ux.reshape(-1, 3, 3)), O.to_c(uy.reshape(-1, 3, 3)), N, rng)
    inv0 = {k: np.einsum("sij,sji->s", O.from_c(st[k]), O.from_c(st[k])) for k in ("E1", "E2", "phi", "pi")}
    mix0 = np.einsum("sij,sji->s", O.from_c(st["E1"]), O.from_c(st["phi"]))
    pl0 = plaquettes(st, N)
    it, res = O.gauge_fix(N, st)
    assert 0 < it < 4000 and res < 1e-9, (it, res)
    for k, v in inv0.items():
        assert np.allclose(np.einsum("sij,sji->s", O.from_c(st[k]), O.from_c(st[k])), v, atol=1e-12)
    assert np.allclose(np.einsum("sij,sji->s", O.from_c(st["E1"]), O.from_c(st["phi"])), mix0, atol=1e-12)
    assert np.allclose(plaquettes(st, N), pl0, atol=1e-12)
    U = O.from_c(st["Ux"])
    assert np.allclose(U @ U.conj().swapaxes(-1, -2), np.eye(3), atol=1e-12)
    # the Coulomb condition itself: chi^a = Im tr(div A t^a) small everywhere
    ux, uy = O.from_c(st["Ux"]).reshape(N, N, 3, 3), O.from_c(st["Uy"]).reshape(N, N, 3, 3)
    div = ux - np.roll(ux, 1, 0) + uy - np.roll(uy, 1, 1)
    chi = np.einsum("xyij,aji->xya", div, T).imag
    assert np.sqrt((chi ** 2).sum(-1) / 6).max() < 1e-3


def wave(N, q, amp):
    """A cos(2 pi q.x / N) t^3 as a [S, 18] Hermitian field"""
    i, j = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    c = amp * np.cos(2 * np.pi * (q[0] * i + q[1] * j) / N).reshape(-1)
    return O.to_c(c[:, None, None] * T[2])


def omega(N, q):
    return np.sqrt(4 * np.sin(np.pi * q[0] / N) ** 2 + 4 * np.sin(np.pi * q[1] / N) ** 2)


def bin_count(N, w, ik, dkt):
    qx, qy = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    om = np.sqrt(4 * np.sin(np.pi * qx / N) ** 2 + 4 * np.sin(np.pi * qy / N) ** 2)
    keep = (qx != N // 2) & (qy != N // 2)
    return int(np.sum(keep & (om > ik * dkt) & (om <= (ik + 1) * dkt)))


@pytest.mark.parametrize("field", ["E1", "E2", "pi"])
def test_multiplicity_single_mode(field):
    N, L, g, dtau, it = 16, 2.0, 2.0, 0.1, 3
    a = L / N
    q = (3, 5)
    A = 0.7
    F = {k: np.zeros((N * N, 18)) for k in ("E1", "E2", "pi")}
    F[field] = wave(N, q, A)
    out = O.multiplicity(N, L, g, dtau, it, F["E1"], F["E2"], F["pi"])
    c = g * g / ((it - 0.5) * dtau) if field != "pi" else (it - 0.5) * dtau
    w = omega(N, q)
    dN = c * A * A * N * N / (2 * w)
    assert out[0] == pytest.approx(dN, rel=1e-12)
    assert out[1] == pytest.approx(dN * w * HBARC / a, rel=1e-12)
    dkt = 2.83 / 100
    ik = int(np.ceil(w / dkt)) - 1
    cnt = bin_count(N, w, ik, dkt)
    assert out[208 + ik] == cnt
    n_bin = dN * N * N / (4 * np.pi ** 2) / cnt
    assert out[8 + ik] == pytest.approx(n_bin, rel=1e-12)
    assert out[108 + ik] == pytest.approx(n_bin * w * HBARC / a, rel=1e-12)
    assert np.count_nonzero(np.abs(out[8:108]) > 1e-20 * n_bin) == 1  # FFT round-off elsewhere
    dN2 = n_bin * (ik + 0.5) * dkt * dkt * 2 * np.pi
    assert out[2] == pytest.approx(dN2, rel=1e-12)
    assert out[3] == pytest.approx(n_bin * w * HBARC / a * (ik + 0.5) * dkt * dkt * 2 * np.pi, rel=1e-12)
    # k_T > 3 GeV and > 6 GeV cuts on the bin's lower edge
    tiny = 1e-20 * dN2
    assert out[4] == (pytest.approx(dN2, rel=1e-12) if ik * dkt / a * HBARC > 3 else pytest.approx(0, abs=tiny))
    assert out[6] == (pytest.approx(dN2, rel=1e-12) if ik * dkt / a * HBARC > 6 else pytest.approx(0, abs=tiny))
    assert ik * dkt / a * HBARC > 3  # the wave sits above the first cut


def test_multiplicity_excluded_line_and_sum_of_passes():
    N, L, g, dtau, it = 16, 2.0, 2.0, 0.1, 1
    z = np.zeros((N * N, 18))
    out = O.multiplicity(N, L, g, dtau, it, wave(N, (N // 2, 3), 1.0), z, wave(N, (5, N // 2), 1.0))
    assert abs(out[0]) < 1e-25 and abs(out[1]) < 1e-25
    a1 = O.multiplicity(N, L, g, dtau, it, wave(N, (1, 2), 0.5), z, z)
    a2 = O.multiplicity(N, L, g, dtau, it, z, z, wave(N, (4, 1), 0.9))
    both = O.multiplicity(N, L, g, dtau, it, wave(N, (1, 2), 0.5), z, wave(N, (4, 1), 0.9))
    assert both[0] == pytest.approx(a1[0] + a2[0], rel=1e-12)
    assert both[1] == pytest.approx(a1[1] + a2[1], rel=1e-12)
    # counts come from one pass only: every included nonzero momentum once
    qx, qy = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    om = np.sqrt(4 * np.sin(np.pi * qx / N) ** 2 + 4 * np.sin(np.pi * qy / N) ** 2)
    keep = (qx != N // 2) & (qy != N // 2) & (om > 0) & (om <= 2.83)
    assert both[208:308].sum() == keep.sum()


def test_multiplicity_parseval_white_noise():
    """for a random Hermitian field, sum_k n(k) omega N^2 / 2 / c over the
    included modes equals sum_k |E(k)|^2 over them (Parseval check of the
    normalisation on a dense spectrum)"""
    N, L, g, dtau, it = 12, 3.0, 1.5, 0.2, 2
    rng = np.random.default_rng(10)
    E = O.to_c(alg(rng.normal(size=(N * N, 8))))
    z = np.zeros_like(E)
    out = O.multiplicity(N, L, g, dtau, it, E, z, z)
    c = g * g / ((it - 0.5) * dtau)
    Ek = np.fft.fft2(O.from_c(E).reshape(N, N, 3, 3), axes=(0, 1))
    qx, qy = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    om = np.sqrt(4 * np.sin(np.pi * qx / N) ** 2 + 4 * np.sin(np.pi * qy / N) ** 2)
    keep = (qx != N // 2) & (qy != N // 2) & (om > 0)
    pw = np.einsum("xyij,xyij->xy", Ek, Ek.conj()).real
    ref = (2 / (N * N) * c * pw[keep] / om[keep]).sum()
    assert out[0] == pytest.approx(ref, rel=1e-12)
