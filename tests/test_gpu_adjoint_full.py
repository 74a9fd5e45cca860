"""The reverse mode to the colour charges at the adjoint workload's geometry
(Au+Au b = 5 fm, L = 30 fm, Ny = 50, tau_max = 0.4 fm/c), checked against
two independent references:
  - at N = 64 (a = 0.47 fm, 9 steps): the gradient of a float64 torch
    restatement of the whole chain, differentiated by torch autograd
    (tests/torch_chain.py: torch.fft, torch.linalg.matrix_exp, an autograd
    Newton root with implicit differentiation);
  - at N = 256 (35 steps): central finite differences of the GPU forward path
    along 32 seeded random directions in the normals z (rho = g sqrt(g2mu2/Ny)
    z), in four groups of eight, with h = 1e-6.
The root finder runs with newton_polish = 4 (four Newton iterations after the
reference's stopping rule Sum|F| <= 1e-6): the links are then solved to
round-off, so the forward map is smooth at the scale of the differences and
the implicit-function gradient is exact.

Error measure.  dE/deta carries round-off of about 1e-13 of its value from
the depth of the chain (50 slices of exponentials, the root finder, 35
steps), so a central difference carries about 1e-13 |dE/deta| / h: ~1e-4
absolute at h = 1e-6 (dE/deta ~ 700 GeV), which exceeds 1e-6 of a
directional derivative whenever that is below ~100.  The directional
derivatives of random directions are ~50 on average with some near 0, so
the error is measured against the size of the gradient along the
directions: |adjoint - difference| <= 5e-6 rms(differences) for every
direction, and the root-mean-square error <= 2e-6 rms(differences) or, when
the differences' own round-off is larger, twice the root-mean-square
distance between the differences at h and at 2h (the same round-off, half
as large at 2h; the truncation error is O(h^2)).  A
larger step does not help: at h = 1e-5 some directions move the root finder
onto another root of the coset equation for a few links (a jump of the
forward map), and their differences are meaningless.  The N = 64
comparison with the torch gradient has no such floor and holds to 1e-6 of
the gradient's largest component, pointwise."""
import sys

import numpy as np
import pytest

import _cases as K

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[adjoint-full] {msg}", file=sys.stderr, flush=True)


def normals(p, seed):
    import ipglasma_amd as I
    probe = I.Event(p, seed=seed)
    probe.init_nucleons()
    assert probe.init_color_charge()
    z = np.stack([np.stack([probe.rho_normals(n, k) for k in range(p.Ny)]) for n in range(2)])
    probe.close()
    return z


def gpu_gradient(p, seed, z, loss="dEdeta"):
    """loss and d loss / dz [2, Ny, S, 8] from the reverse mode; g2mu2 of
    both nuclei; the tau = 0 links Ux, Uy [N, N, 3, 3]"""
    import ipglasma_amd as I
    import _oracle as O
    ev = I.Event(p, seed=seed)
    ev.set_rho_normals(z)
    assert ev.init()
    gA, gB = ev.field("g2mu2A"), ev.field("g2mu2B")
    N = p.size
    links = [O.from_c(ev.field(k)).reshape(N, N, 3, 3) for k in ("Ux", "Uy")]
    L0, grho = ev.adjoint(loss=loss, wrt="rho")
    ev.close()
    amp = np.stack([p.g * np.sqrt(gA / p.Ny), p.g * np.sqrt(gB / p.Ny)])
    return L0, grho * amp[:, None, :, None], np.stack([gA, gB]), links


def forward(p, seed, z):
    import ipglasma_amd as I
    ev = I.Event(p, seed=seed)
    ev.set_rho_normals(z)
    assert ev.init()
    ev.run()
    o = ev.observables()
    ev.close()
    return o["dEdeta"], o["eccentricity"][1]


def test_rho_gradient_against_torch_autograd_n64():
    torch = pytest.importorskip("torch")
    import torch_chain as TC
    p = K.params("AuAu", N=64, Ny=50, seed=21, newton_polish=4)
    assert p.itmax == 9
    z = normals(p, 21)
    L0, gz, g2, links = gpu_gradient(p, 21, z)
    N = p.size
    m_lat = p.m * (p.L / N) / TC.HBARC
    zt = torch.tensor(z.reshape(2, p.Ny, N, N, 8), requires_grad=True)
    # the torch root finder starts from the device's roots (so that it picks
    # the same root where the coset equation has several) and iterates
    # torch's own residual to 1e-14; its gradient comes from torch autograd
    # and its own implicit differentiation
    Lt = TC.dedeta(zt, torch.tensor(g2.reshape(2, N, N)), p.g, m_lat, p.Ny, p.L, p.dtau, p.itmax, start=links)
    Lt.backward()
    gt = zt.grad.numpy().reshape(gz.shape)
    rel_L = abs(Lt.item() - L0) / abs(L0)
    err = np.abs(gt - gz).max() / np.abs(gt).max()
    rng = np.random.default_rng(64)
    dirs = []
    for _ in range(8):
        d = rng.normal(size=z.shape)
        a, b = float(np.sum(gz * d)), float(np.sum(gt * d))
        dirs.append(abs(a - b) / abs(b))
    say(f"N=64 Ny=50: dE/deta gpu {L0:.12g} torch {Lt.item():.12g} rel {rel_L:.1e}; gradient max diff / max "
        f"{err:.2e}; directional rel max {max(dirs):.2e}")
    assert rel_L < 1e-10
    assert err < 1e-6
    assert max(dirs) < 1e-7


@pytest.fixture(scope="module")
def c4():
    p = K.params("AuAu", N=256, Ny=50, seed=21, newton_polish=4)
    assert p.itmax == 35
    z = normals(p, 21)
    L0, gz, _, _ = gpu_gradient(p, 21, z)
    assert L0 == pytest.approx(forward(p, 21, z)[0], rel=1e-12)
    return p, z, gz


RESULTS = {}


@pytest.mark.parametrize("group", range(4))
def test_rho_gradient_finite_differences_c4(c4, group):
    """32 seeded directions (8 per group), central differences with
    h = 1e-6 against the reverse mode"""
    p, z, gz = c4
    h = 1e-6
    for k in range(8):
        rng = np.random.default_rng(1000 + 8 * group + k)
        d = rng.normal(size=z.shape)
        pred = float(np.sum(gz * d))
        fd = (forward(p, 21, z + h * d)[0] - forward(p, 21, z - h * d)[0]) / (2 * h)
        fd2 = (forward(p, 21, z + 2 * h * d)[0] - forward(p, 21, z - 2 * h * d)[0]) / (4 * h)
        RESULTS[8 * group + k] = (pred, fd, fd2)
        say(f"C4 direction {8 * group + k}: adjoint {pred:.12g} central difference {fd:.12g} (2h: {fd2:.12g}) "
            f"rel {abs(pred - fd) / abs(fd):.2e}")
    rows = np.array([RESULTS[k] for k in range(8 * group, 8 * group + 8)])
    scale = np.sqrt(np.mean(rows[:, 1] ** 2))
    err = np.abs(rows[:, 0] - rows[:, 1]) / scale
    rms = np.sqrt(np.mean(err ** 2))
    noise = np.sqrt(np.mean(((rows[:, 1] - rows[:, 2]) / scale) ** 2))
    say(f"C4 group {group}: |adjoint - difference| / rms(difference): max {err.max():.2e}, rms {rms:.2e}; "
        f"rms |difference(h) - difference(2h)| / rms(difference) {noise:.2e}")
    assert err.max() < 5e-6, err
    assert rms < max(2e-6, 2 * noise), (rms, noise)
