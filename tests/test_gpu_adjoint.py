"""GPU reverse mode of dE/deta through the leapfrog (ipg_adjoint_dedeta)
against central finite differences of the GPU forward path: for random
directions d in the tau = 0 fields (complex for the links, Hermitian for pi),
Re sum tr(grad^+ d) = (L(x + h d) - L(x - h d)) / 2h to 1e-6 relative.  The
per-site reverse functions are checked on the CPU (test_adjoint_host.py)."""
import numpy as np
import pytest

import _cases as K
import _oracle as O

pytestmark = pytest.mark.gpu


def _loss_at(p, fields):
    import ipglasma_amd as I
    ev = I.Event(p, seed=1)
    for k, v in fields.items():
        ev.set_field(k, v)
    ev.run()
    v = ev.observables()["dEdeta"]
    ev.close()
    return v


@pytest.mark.parametrize("N,maxtime", [(16, 0.6), (32, 0.4)])
def test_adjoint_dedeta_matches_finite_differences(N, maxtime):
    import ipglasma_amd as I
    p = K.params("AuAu", N=N, Ny=3, maxtime=maxtime, L=8.0)
    ev = I.Event(p, seed=21)
    assert ev.init()
    x0 = {k: ev.field(k).copy() for k in O.FIELDS}
    L0, g = ev.adjoint_dedeta()
    ev.close()
    assert L0 > 0 and np.all(np.isfinite(g["Ux"]))
    # the value agrees with the ordinary forward run
    assert L0 == pytest.approx(_loss_at(p, x0), rel=1e-12)
    rng = np.random.default_rng(N)
    S = N * N
    for trial in range(3):
        d = {}
        for k in ("Ux", "Uy"):
            d[k] = rng.normal(size=(S, 18))
        a = rng.normal(size=(S, 3, 3)) + 1j * rng.normal(size=(S, 3, 3))
        d["pi"] = O.to_c((a + a.conj().transpose(0, 2, 1)) / 2)
        pred = sum(float(np.sum(g[k] * d[k])) for k in d)
        h = 1e-5 * max(1.0, np.abs(x0["pi"]).max())
        xp = {k: v.copy() for k, v in x0.items()}
        xm = {k: v.copy() for k, v in x0.items()}
        for k in d:
            xp[k] = xp[k] + h * d[k]
            xm[k] = xm[k] - h * d[k]
        fd = (_loss_at(p, xp) - _loss_at(p, xm)) / (2 * h)
        assert abs(pred - fd) <= 1e-6 * abs(fd), (trial, pred, fd)
