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


def _rho_setup(p, seed):
    import ipglasma_amd as I
    probe = I.Event(p, seed=seed)
    probe.init_nucleons()
    assert probe.init_color_charge()
    z = np.stack([np.stack([probe.rho_normals(n, k) for k in range(p.Ny)]) for n in range(2)])
    probe.close()
    return z


def test_full_tape_gradient_matches_checkpointed():
    """the reverse mode with every forward state stored (checkpoint interval
    1, nothing recomputed) against sqrt(T) checkpoints (interval 3 for the 8
    steps) at N=64:
    the recomputed states are bit-identical, so the gradients agree to the
    round-off of the atomic sums (the loss, T^{mu nu}, pi and link reverse
    kernels add in a run-dependent order): 1e-12 of the largest entry
    (dE/deta with respect to rho, eps_2 with respect to the tau = 0 fields)"""
    import ipglasma_amd as I
    p = K.params("AuAu", N=64, Ny=2, maxtime=0.4, L=8.0)
    p.set(bmin=3.0, bmax=3.0)
    z = _rho_setup(p, 21)
    out = {}
    for k in (3, 1):
        ev = I.Event(p, seed=21)
        ev.set_rho_normals(z)
        assert ev.init()
        ev.set_checkpoint_interval(k)
        out[("rho", k)] = ev.adjoint(loss="dEdeta", wrt="rho")
        assert ev.checkpoint_interval() == k
        ev.close()
        ev = I.Event(p, seed=21)
        ev.set_rho_normals(z)
        assert ev.init()
        ev.set_checkpoint_interval(k)
        out[("eps2", k)] = ev.adjoint(loss="eps2", wrt="fields")
        ev.close()
    assert out[("rho", 3)][0] == out[("rho", 1)][0]
    a, b = out[("rho", 3)][1], out[("rho", 1)][1]
    print("rho gradient, full tape vs checkpoints: max diff / max", np.abs(a - b).max() / np.abs(a).max())
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    assert out[("eps2", 3)][0] == out[("eps2", 1)][0]
    scale = max(np.abs(v).max() for v in out[("eps2", 3)][1].values())
    for f in out[("eps2", 3)][1]:
        a, b = out[("eps2", 3)][1][f], out[("eps2", 1)][1][f]
        assert np.abs(a - b).max() <= 1e-12 * scale, f


def test_adjoint_eps2_rho_matches_finite_differences():
    """d eps_2 / d rho (normals z) through the whole chain, against central
    differences.  The root finder's stopping noise (~1e-9 in the links) and
    the curvature of the boosted edge cells bound the usable steps from
    both sides; the step 2e-5 and the bound 1e-3 sit between them."""
    import ipglasma_amd as I
    N, Ny, seed = 32, 3, 21
    p = K.params("AuAu", N=N, Ny=Ny, maxtime=0.4, L=8.0)
    p.set(bmin=3.0, bmax=3.0)
    z = _rho_setup(p, seed)
    ev = I.Event(p, seed=seed)
    ev.set_rho_normals(z)
    assert ev.init()
    gA, gB = ev.field("g2mu2A"), ev.field("g2mu2B")
    L0, grho = ev.adjoint(loss="eps2", wrt="rho")
    ev.close()

    def e2(zz):
        e = I.Event(p, seed=seed)
        e.set_rho_normals(zz)
        assert e.init()
        e.run()
        v = e.observables()["eccentricity"][1]
        e.close()
        return v

    assert L0 == pytest.approx(e2(z), rel=1e-11)
    amp = np.stack([p.g * np.sqrt(gA / Ny), p.g * np.sqrt(gB / Ny)])
    gz = grho * amp[:, None, :, None]
    rng = np.random.default_rng(5)
    for trial in range(3):
        d = rng.normal(size=z.shape)
        h = 2e-5
        fd = (e2(z + h * d) - e2(z - h * d)) / (2 * h)
        pred = float(np.sum(gz * d))
        print(f"eps2 rho gradient trial {trial}: adjoint {pred:.10g} central difference {fd:.10g} "
              f"rel {abs(pred - fd) / abs(fd):.2e}")
        assert abs(pred - fd) <= 1e-3 * abs(fd), (trial, pred, fd)


def test_batched_adjoint_matches_single_event_contexts():
    """ipg_batch_adjoint: initial conditions of three events of two lattice
    sizes in one launch sequence, then each event's reverse mode through one
    shared workspace; the losses and rho gradients equal those of one
    context per event (to the round-off of the reverse kernels' atomic
    sums), for dE/deta and for eps_2"""
    import ipglasma_amd as I
    ps = [K.params("AuAu", N=32, Ny=3, L=8.0), K.params("pPb", N=16, Ny=3), K.params("AuAu", N=32, Ny=3, L=8.0)]
    seeds = [21, 5, 22]
    for loss in ("dEdeta", "eps2"):
        b = I.Batch(ps, seeds)
        vals = b.adjoint(loss=loss)
        peak, kint = b.adjoint_info()
        assert peak > 0 and kint >= 1
        for e, (p, s) in enumerate(zip(ps, seeds)):
            ev = I.Event(p, seed=s)
            tries = 0
            while not ev.init():
                tries += 1
                assert tries < 20
            L0, g = ev.adjoint(loss=loss, wrt="rho")
            ev.close()
            assert vals[e] == pytest.approx(L0, rel=1e-13), (loss, e)
            gb = b.rho_grad(e, p.Ny, p.size ** 2)
            assert np.abs(gb - g).max() <= 1e-12 * np.abs(g).max(), (loss, e)
        b.close()


def test_reverse_step_kernels_agree():
    """the reverse leapfrog step as three kernels (tuning off) and, while the
    tuner measures them, alternating with the fused forms, then
    with the tuner's choice: the same arithmetic per site, so the gradients
    agree to the round-off of the loss's atomic sums"""
    import os

    import ipglasma_amd as I
    p = K.params("AuAu", N=96, Ny=3, maxtime=0.4, L=8.0)
    key = "adj_step_kernels/N=96"
    assert key not in I.autotune_report()
    runs = []
    for tune in ("0", "1", "1"):
        os.environ["IPG_AUTOTUNE"] = tune
        try:
            ev = I.Event(p, seed=21)
            assert ev.init()
            runs.append(ev.adjoint_dedeta())
            ev.close()
        finally:
            os.environ.pop("IPG_AUTOTUNE")
    rep = I.autotune_report()
    assert key in rep and rep[key]["chosen"] is not None, rep
    L0, g0 = runs[0]
    for L1, g1 in runs[1:]:
        assert L1 == pytest.approx(L0, rel=1e-14)
        for k in O.FIELDS:
            assert np.abs(g1[k] - g0[k]).max() <= 1e-12 * np.abs(g0[k]).max(), k


def test_kept_slices_give_the_same_rho_gradient():
    """ipg_set_keep_slices: the reverse setV takes the forward's filtered
    colour fields (Ny = 11: two chunks of 8 slices, the second short)
    instead of recomputing them; the gradient equals the recomputed one.
    After a field is set (here pi, to its own values) the kept fields are
    dropped and the reverse pass recomputes them."""
    import ipglasma_amd as I
    p = K.params("AuAu", N=32, Ny=11, maxtime=0.4, L=8.0)

    def grad(keep, touch=False):
        ev = I.Event(p, seed=21)
        if keep:
            ev.keep_slices(True)
        assert ev.init()
        if touch:
            ev.set_field("pi", ev.field("pi"))
        L, g = ev.adjoint(loss="dEdeta", wrt="rho")
        ev.close()
        return L, np.asarray(g)

    L0, g0 = grad(False)
    for kw in ({"keep": True}, {"keep": True, "touch": True}):
        L1, g1 = grad(**kw)
        assert L1 == pytest.approx(L0, rel=1e-14), kw
        assert np.abs(g1 - g0).max() <= 1e-12 * np.abs(g0).max(), kw
