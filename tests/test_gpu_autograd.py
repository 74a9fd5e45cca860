"""The reverse mode wrapped in a torch.autograd.Function (a thin test
harness only: the library takes and returns host arrays through the C ABI).
The colour-charge normals z enter a torch graph; backward() must return the
library's d loss / d rho chained through rho = g sqrt(g2mu2/Ny) z, and a
directional central difference of the composed function must agree."""
import numpy as np
import pytest

import _cases as K

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


class _GlasmaLoss(torch.autograd.Function):
    """z [2, Ny, S, 8] (float64, CPU) -> dE/deta or eps_2 at tau_max"""

    @staticmethod
    def forward(ctx, z, p, seed, loss):
        import ipglasma_amd as I
        ev = I.Event(p, seed=seed)
        ev.set_rho_normals(z.detach().numpy())
        assert ev.init()
        gA, gB = ev.field("g2mu2A"), ev.field("g2mu2B")
        value, grho = ev.adjoint(loss=loss, wrt="rho")
        ev.close()
        amp = np.stack([p.g * np.sqrt(gA / p.Ny), p.g * np.sqrt(gB / p.Ny)])
        ctx.save_for_backward(torch.from_numpy(grho * amp[:, None, :, None]))
        return torch.tensor(value, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        (gz,) = ctx.saved_tensors
        return gout * gz, None, None, None


def _value(p, seed, loss, z):
    import ipglasma_amd as I
    ev = I.Event(p, seed=seed)
    ev.set_rho_normals(z)
    assert ev.init()
    ev.run()
    o = ev.observables()
    ev.close()
    return o["dEdeta"] if loss == "dEdeta" else o["eccentricity"][1]


@pytest.mark.parametrize("loss", ["dEdeta", "eps2"])
def test_autograd_function_backward(loss):
    import ipglasma_amd as I
    N, Ny, seed = 32, 3, 21
    p = K.params("AuAu", N=N, Ny=Ny, maxtime=0.4, L=8.0)
    p.set(bmin=3.0, bmax=3.0)
    probe = I.Event(p, seed=seed)
    probe.init_nucleons()
    assert probe.init_color_charge()
    z0 = np.stack([np.stack([probe.rho_normals(n, k) for k in range(Ny)]) for n in range(2)])
    probe.close()
    z = torch.tensor(z0, dtype=torch.float64, requires_grad=True)
    # a torch graph around the library call: 3 L(z) + |z|^2 / 2
    out = 3.0 * _GlasmaLoss.apply(z, p, seed, loss) + 0.5 * (z * z).sum()
    out.backward()
    assert out.item() == pytest.approx(3.0 * _value(p, seed, loss, z0) + 0.5 * float(np.sum(z0 * z0)), rel=1e-12)
    rng = np.random.default_rng(3)
    d = rng.normal(size=z0.shape)
    h = 2e-4 if loss == "dEdeta" else 2e-5
    fd = 3.0 * (_value(p, seed, loss, z0 + h * d) - _value(p, seed, loss, z0 - h * d)) / (2 * h)
    pred = float(np.sum((z.grad.numpy() - z0) * d))
    assert abs(pred - fd) <= 1e-3 * abs(fd), (loss, pred, fd)
