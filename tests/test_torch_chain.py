"""The float64 torch restatement of the chain (tests/torch_chain.py), which the
GPU reverse mode is checked against at N = 64, validated on the CPU before
it is trusted: its forward dE/deta against the C oracle's chain (setV from
the same normals, the root finder polished to round-off, Evolution::run),
and its autograd gradient (with the root differentiated implicitly) against
central differences of itself.

The coset equation of the tau = 0 links can have more than one solution; a
Newton iteration from U1 U2 with the exact Jacobian can end on another one
than the reference's iteration (from alpha = 0 with its difference Jacobian
and step guard).  The torch chain is therefore started from the oracle's
roots here (and from the device's on the GPU); it iterates its own residual
to 1e-14 from there."""
import ctypes as C

import numpy as np
import pytest

import _cases as K
import _oracle as O

torch = pytest.importorskip("torch")
import torch_chain as TC  # noqa: E402


def _event(N=16, Ny=2, seed=3):
    p = K.params("pPb", N=N, Ny=Ny, L=6.0, maxtime=0.3)
    import ipglasma_amd as I
    tab = K.load_table(I.QS_TABLE)
    for s in range(seed, seed + 50):
        st, nA, nB, info = K.oracle_event(p, s)
        TpA, TpB, gA, gB = K.oracle_color_charge(p, nA, nB, tab)
        if gA.max() > 0 and gB.max() > 0:
            break
    rng = np.random.default_rng(seed)
    z = rng.normal(size=(2, Ny, N * N, 8))
    return p, z, np.stack([gA, gB])


def _oracle_chain(p, z, g2):
    N, S = p.size, p.size ** 2
    L = O.lib()
    L.orc_set_newton_polish.argtypes = [C.c_int]
    m_lat = p.m * (p.L / N) / TC.HBARC
    V = np.zeros((2, S, 18))
    L.orc_setv(N, p.Ny, p.g, m_lat, 0.0, g2[0].copy(), g2[1].copy(), z.ctypes.data, None, 0, V.reshape(-1))
    st = {k: np.zeros((S, 18)) for k in O.FIELDS}
    stats = np.zeros(5)
    L.orc_set_newton_polish(4)
    try:
        L.orc_init_fields(N, p.g, V[0].copy(), V[1].copy(), 5, *[st[k] for k in O.FIELDS], stats)
    finally:
        L.orc_set_newton_polish(0)
    assert stats[3] == 0  # every link converged
    links = [O.from_c(st[k]).reshape(N, N, 3, 3).copy() for k in ("Ux", "Uy")]
    T = np.zeros((10, S))
    obs = np.zeros(3)
    L.orc_run(N, p.g, p.L, p.dtau, p.itmax, *[st[k] for k in O.FIELDS], T.reshape(-1), obs)
    return obs[0], links


def test_torch_chain_forward_matches_oracle():
    p, z, g2 = _event()
    N = p.size
    ref, links = _oracle_chain(p, np.ascontiguousarray(z), g2)
    m_lat = p.m * (p.L / N) / TC.HBARC
    zt = torch.tensor(z.reshape(2, p.Ny, N, N, 8))
    with torch.no_grad():
        Lt = TC.dedeta(zt, torch.tensor(g2.reshape(2, N, N)), p.g, m_lat, p.Ny, p.L, p.dtau, p.itmax,
                       start=links)
    assert ref > 0
    assert Lt.item() == pytest.approx(ref, rel=1e-10)


def test_torch_chain_gradient_matches_its_own_differences():
    p, z, g2 = _event(seed=9)
    N = p.size
    m_lat = p.m * (p.L / N) / TC.HBARC
    gt = torch.tensor(g2.reshape(2, N, N))
    _, links = _oracle_chain(p, np.ascontiguousarray(z), g2)

    def f(zz, grad=False):
        zt = torch.tensor(zz.reshape(2, p.Ny, N, N, 8), requires_grad=grad)
        L = TC.dedeta(zt, gt, p.g, m_lat, p.Ny, p.L, p.dtau, p.itmax, start=links)
        if grad:
            L.backward()
            return L.item(), zt.grad.numpy().reshape(z.shape)
        return L.item()

    L0, g = f(z, grad=True)
    rng = np.random.default_rng(4)
    h = 1e-5
    for _ in range(3):
        d = rng.normal(size=z.shape)
        fd = (f(z + h * d) - f(z - h * d)) / (2 * h)
        assert float(np.sum(g * d)) == pytest.approx(fd, rel=1e-6, abs=1e-8 * abs(L0))
