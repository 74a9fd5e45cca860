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


def _loss_z(p, seed, z):
    import ipglasma_amd as I
    ev = I.Event(p, seed=seed)
    ev.set_rho_normals(z)
    assert ev.init()
    ev.run()
    v = ev.observables()["dEdeta"]
    ev.close()
    return v


@pytest.mark.parametrize("N,Ny", [(32, 3), (64, 2)])
def test_adjoint_rho_matches_finite_differences(N, Ny):
    """d(dE/deta)/d rho through evolution, pi, root finder, links and setV,
    against central differences in the colour-charge normals z
    (rho = g sqrt(g2mu2/Ny) z).  The forward root finder stops at
    sum |F| <= 1e-6 with a forward-difference Jacobian (linear convergence),
    so its link solutions carry ~1e-9 errors that move with every
    perturbation: the step is 2e-4 (noise ~5e-6, truncation ~1e-7) and the
    bound 1e-4 relative.  Coarser lattices (N=16 at L=8 fm, a=0.5 fm) are not
    used: there the root finder restarts and changes branch under such steps,
    so the forward map is not smooth at any usable step."""
    import ipglasma_amd as I
    p = K.params("AuAu", N=N, Ny=Ny, maxtime=0.4, L=8.0)
    seed = 21
    probe = I.Event(p, seed=seed)
    probe.init_nucleons()
    ok = probe.init_color_charge()
    assert ok
    z = np.stack([np.stack([probe.rho_normals(n, k) for k in range(Ny)]) for n in range(2)])
    probe.close()
    ev = I.Event(p, seed=seed)
    ev.set_rho_normals(z)
    assert ev.init()
    gA, gB = ev.field("g2mu2A"), ev.field("g2mu2B")
    L0, grho = ev.adjoint_rho()
    ev.close()
    assert L0 == pytest.approx(_loss_z(p, seed, z), rel=1e-12)
    amp = np.stack([p.g * np.sqrt(gA / Ny), p.g * np.sqrt(gB / Ny)])  # [2, S]
    gz = grho * amp[:, None, :, None]
    rng = np.random.default_rng(N + Ny)
    for trial in range(3):
        d = rng.normal(size=z.shape)
        h = 2e-4
        fd = (_loss_z(p, seed, z + h * d) - _loss_z(p, seed, z - h * d)) / (2 * h)
        pred = float(np.sum(gz * d))
        print(f"rho gradient N={N} trial {trial}: adjoint {pred:.10g} central difference {fd:.10g} "
              f"rel {abs(pred - fd) / abs(fd):.2e}")
        assert abs(pred - fd) <= 1e-4 * abs(fd), (trial, pred, fd)


def _loss_generic(p, fields, loss, w):
    import ipglasma_amd as I
    ev = I.Event(p, seed=1)
    for k, v in fields.items():
        ev.set_field(k, v)
    ev.run()
    if loss == "tmunu":
        v = float(np.sum(w[:, None] * ev.tmunu()))
    else:
        v = ev.observables()["eccentricity"][1]
    ev.close()
    return v


def _fd_fields(p, x0, g, loss, w, rng, trials, h, tol):
    S = p.size * p.size
    for trial in range(trials):
        d = {}
        for k in ("Ux", "Uy"):
            d[k] = rng.normal(size=(S, 18))
        a = rng.normal(size=(S, 3, 3)) + 1j * rng.normal(size=(S, 3, 3))
        d["pi"] = O.to_c((a + a.conj().transpose(0, 2, 1)) / 2)
        pred = sum(float(np.sum(g[k] * d[k])) for k in d)
        xp = {k: v.copy() for k, v in x0.items()}
        xm = {k: v.copy() for k, v in x0.items()}
        for k in d:
            xp[k] = xp[k] + h * d[k]
            xm[k] = xm[k] - h * d[k]
        fd = (_loss_generic(p, xp, loss, w) - _loss_generic(p, xm, loss, w)) / (2 * h)
        print(f"{loss} gradient trial {trial}: adjoint {pred:.10g} central difference {fd:.10g} "
              f"rel {abs(pred - fd) / max(abs(fd), 1e-300):.2e}")
        assert abs(pred - fd) <= tol * abs(fd), (loss, trial, pred, fd)


@pytest.mark.parametrize("comp", list(range(10)) + ["all"])
def test_adjoint_tmunu_components_match_finite_differences(comp):
    """the reverse mode of each of the ten T^{mu nu} components at itmax
    (loss = sum over sites of T_k, or a random mix of all ten), through the
    leapfrog to the tau = 0 fields, against central differences"""
    import ipglasma_amd as I
    p = K.params("AuAu", N=16, Ny=3, maxtime=0.6, L=8.0)
    ev = I.Event(p, seed=21)
    assert ev.init()
    x0 = {k: ev.field(k).copy() for k in O.FIELDS}
    rng = np.random.default_rng(7 if comp == "all" else 100 + comp)
    w = rng.normal(size=10) if comp == "all" else np.eye(10)[comp]
    L0, g = ev.adjoint(loss="tmunu", weights=w)
    ev.close()
    assert L0 == pytest.approx(_loss_generic(p, x0, "tmunu", w), rel=1e-11, abs=1e-12)
    h = 1e-5 * max(1.0, np.abs(x0["pi"]).max())
    _fd_fields(p, x0, g, "tmunu", w, rng, 2, h, 1e-6)


@pytest.mark.parametrize("N,maxtime", [(32, 0.4)])
def test_adjoint_eps2_fields_match_finite_differences(N, maxtime):
    """d eps_2 / d(tau = 0 fields): eps_2 of the Landau-frame weights
    eps u^tau at tau_max (the reverse of the eccentricity sums and centroid,
    of the 4x4 eigenproblem per site and of all ten T^{mu nu} components)"""
    import ipglasma_amd as I
    p = K.params("AuAu", N=N, Ny=3, maxtime=maxtime, L=8.0)
    p.set(bmin=3.0, bmax=3.0)
    ev = I.Event(p, seed=21)
    assert ev.init()
    x0 = {k: ev.field(k).copy() for k in O.FIELDS}
    L0, g = ev.adjoint(loss="eps2")
    ev.close()
    assert 0 < L0 < 1
    assert L0 == pytest.approx(_loss_generic(p, x0, "eps2", None), rel=1e-11)
    rng = np.random.default_rng(33)
    # the Landau frame of strongly boosted edge cells (u^tau up to ~20 here)
    # curves eps_2 sharply: central differences settle for steps of 1e-6 to
    # 1e-8 of the field scale (at 1e-5 the second direction is still 4e-4 off)
    h = 1e-7 * max(1.0, np.abs(x0["pi"]).max())
    _fd_fields(p, x0, g, "eps2", None, rng, 3, h, 1e-5)
