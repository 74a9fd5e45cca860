"""GPU parity of the Landau frame (MyEigen.cpp:25-321) and the eccentricities
(Evolution.cpp:1280-1810) through the C ABI against the oracle.

The per-site Landau solver is built without contraction and uses only
correctly rounded operations, so the device reproduces the oracle's bit
pattern for the same T^{mu nu}; the tests require agreement to 1e-13 of the
component scale and report the exactly equal fraction.  Eccentricities are
sums in a different order (and with z^n in place of atan/cos/sin): 1e-12
relative.  End to end (device chain vs oracle chain from the same seeds)
eps_2 is held to 1e-8 relative: the tau=0 root finder stops at
sum |F_a| <= 1e-6, so the two chains start from links that agree to ~1e-8
(the root-finder test allows 1e-7), and eps_2 is a ratio of moments with
cancellations; dE/deta, a sum of positive terms, stays within 1e-9.
"""
import numpy as np
import pytest

import _cases as K
import _oracle as O
from test_oracle_landau import fluid_T, pack_T, random_frame

pytestmark = pytest.mark.gpu


def _compare_landau(Lg, Lo, what):
    exact = np.mean(Lg == Lo)
    for k in range(15):
        scale = max(np.abs(Lo[k]).max(), 1e-300)
        err = np.abs(Lg[k] - Lo[k]).max() / scale
        assert err <= 1e-13, (what, k, err)
    print(f"{what}: fraction bit-exact {exact:.6f}")
    return exact


def _ecc_check(o, r, rel=1e-12):
    for n in range(6):
        assert o["eccentricity"][n] == pytest.approx(r[n], rel=rel, abs=1e-15), n
        assert o["Psi"][n] == pytest.approx(r[6 + n], rel=rel, abs=1e-15), n
    for key, k in (("area", 12), ("Rbar", 13), ("avg_eps", 14), ("avg_qs2A_qs2B", 15), ("rms_r", 16)):
        assert o[key] == pytest.approx(r[k], rel=rel), key
    assert o["centroid"][0] == pytest.approx(r[17], rel=1e-10, abs=1e-13)
    assert o["centroid"][1] == pytest.approx(r[18], rel=1e-10, abs=1e-13)


def test_landau_synthetic_fluids_and_eccentricity():
    import ipglasma_amd as I
    N = 32
    p = K.params("AuAu", N=N, L=8.0)
    ev = I.Event(p, seed=1)
    S = N * N
    a = p.L / N
    it = 7
    tau = it * p.dtau * a
    rng = np.random.default_rng(4)
    Ts = []
    for s in range(S):
        fr = random_frame(rng, tau, 0.7)
        e = rng.uniform(0.01, 30.)
        w = rng.uniform(0.2, 1., 3)
        T = fluid_T(e, e * w / w.sum(), fr, tau)
        if s % 17 == 3:                      # arbitrary symmetric tensors
            T = rng.normal(size=(4, 4))
            T = T + T.T
        if s % 29 == 5:                      # no timelike eigenvector
            T = np.zeros((4, 4))
            T[0, 0], T[1, 1] = 6., -6.
            T[0, 1] = T[1, 0] = 4.
            T[2, 2], T[3, 3] = 1., 2. / tau ** 2
        Ts.append(T)
    T10 = np.ascontiguousarray(pack_T(np.array(Ts)))
    g2A = rng.uniform(0, 1e-3, S)
    g2B = rng.uniform(0, 1e-3, S)
    ev.set_field("g2mu2A", g2A)
    ev.set_field("g2mu2B", g2B)
    ev.landau_from(T10, it)
    Lg = ev.landau()
    Lo, avg, nf = O.landau(N, p.L, p.dtau, it, T10)
    assert nf == 0
    _compare_landau(Lg, Lo, "synthetic")
    o = ev.observables()
    assert o["landau_failures"] == 0
    assert o["landau_no_timelike"] >= sum(1 for s in range(S) if s % 29 == 5)
    r = O.eccentricity(N, p.L, Lo[0], Lo[1], g2A, g2B, p.QsmuRatio, p.QsmuRatio)
    _ecc_check(o, r)
    for k in range(3):
        assert o["u_rms"][k] == pytest.approx(np.sqrt(avg[k] / avg[3]), rel=1e-12)


def test_landau_after_evolution_end_to_end():
    """Init + evolution on the GPU; the Landau frame of the device T^{mu nu}
    against the oracle on that same T, and eps_2 against the oracle chain"""
    import ipglasma_amd as I
    p = K.params("AuAu", N=32, Ny=4)
    ev = I.Event(p, seed=21)
    assert ev.init()
    ev.run()
    o = ev.observables()
    N = p.size
    S = N * N
    Tg = ev.tmunu()
    Lo, avg, nf = O.landau(N, p.L, p.dtau, p.itmax, Tg)
    Lg = ev.landau()
    _compare_landau(Lg, Lo, "evolved")
    g2A = ev.field("g2mu2A")
    g2B = ev.field("g2mu2B")
    r = O.eccentricity(N, p.L, Lo[0], Lo[1], g2A, g2B, p.QsmuRatio, p.QsmuRatio)
    _ecc_check(o, r)
    assert 0. < o["eccentricity"][1] < 1.
    # the oracle chain from the device's initial conditions
    st, nA, nB, info = K.oracle_event(p, 21)
    TpA, TpB, gA, gB = K.oracle_color_charge(p, nA, nB, K.load_table(I.QS_TABLE))
    V = np.zeros((2, S, 18))
    m_lat = p.m * (p.L / N) / 0.1973269718
    O.lib().orc_setv(N, p.Ny, p.g, m_lat, 0.0, gA, gB, None, None, o["rho_seed"], V.reshape(-1))
    ref = {k: np.zeros((S, 18)) for k in O.FIELDS}
    stats = np.zeros(5)
    O.lib().orc_init_fields(N, p.g, V[0].copy(), V[1].copy(), o["newton_seed"], *[ref[k] for k in O.FIELDS],
                            stats)
    T = np.zeros((10, S))
    obs = np.zeros(3)
    O.lib().orc_run(N, p.g, p.L, p.dtau, p.itmax, *[ref[k] for k in O.FIELDS], T.reshape(-1), obs)
    Lref, _, _ = O.landau(N, p.L, p.dtau, p.itmax, T)
    rr = O.eccentricity(N, p.L, Lref[0], Lref[1], gA, gB, p.QsmuRatio, p.QsmuRatio)
    assert abs(o["eccentricity"][1] - rr[1]) <= 1e-8 * rr[1], (o["eccentricity"][1], rr[1])
    assert abs(o["dEdeta"] - obs[0]) <= 1e-9 * obs[0], (o["dEdeta"], obs[0])


def test_batch_landau_matches_plain():
    import ipglasma_amd as I
    ps = [K.params("AuAu", N=32, Ny=3), K.params("pPb", N=16, Ny=3)]
    seeds = [5, 2]
    plain = []
    for p, s in zip(ps, seeds):
        ev = I.Event(p, seed=s)
        ok = ev.init()
        tries = 0
        while not ok and tries < 20:
            ok = ev.init()
            tries += 1
        ev.run()
        plain.append(ev.observables())
        ev.close()
    b = I.Batch(ps, seeds)
    b.run(use_graph=True)
    for e in range(len(ps)):
        o = b.observables(e)
        for n in range(6):
            assert o["eccentricity"][n] == pytest.approx(plain[e]["eccentricity"][n], rel=1e-12)
        assert o["area"] == plain[e]["area"]
    b.close()


def test_eccentricity_file_line(tmp_path):
    """eccentricities<id>.dat (Evolution.cpp:1797-1809): one appended line
    per call, the 23 columns printed as an ostream prints them (%g), maxX
    and maxY from the device's Landau frame by the reference's loop."""
    import ipglasma_amd as I
    p = K.params("AuAu", N=32, Ny=4, bmin=3.0, bmax=3.0)
    ev = I.Event(p, seed=13)
    assert ev.init()
    ev.run()
    o = ev.observables()
    f = tmp_path / "eccentricities0.dat"
    ev.write_eccentricities(f)
    ev.write_eccentricities(f)
    lines = f.read_text().splitlines()
    assert len(lines) == 2 and lines[0] == lines[1]
    land = ev.landau()
    S = p.size * p.size
    mx, my = I.eccentricity_extents(land[0][:S], land[1][:S], p.L, o["centroid"][0], o["centroid"][1])
    a = p.L / p.size
    cols = [p.itmax * a * p.dtau]
    for n in range(6):
        cols += [o["eccentricity"][n], o["Psi"][n]]
    cols += [0.0, o["rms_r"], mx, my, o["b"], o["Tpp"], o["area"], o["Rbar"], o["avg_eps"], o["avg_qs2A_qs2B"]]
    assert lines[0] == " ".join("%g" % v for v in cols)
    assert mx > 0 and my > 0


def test_hydro_file_of_an_event(tmp_path):
    """ipg_write_hydro on an evolved event writes the file the host writer
    makes from the device's Landau frame (MyEigen.cpp:326-800): header,
    size_out^2 lines per eta slice plus an empty one, u^tau >= 1, eps >= 0
    and the vacuum lines (eps = 0) at rest.  The weights are divided by the
    hydro spacing, so a hydro grid finer than the lattice extrapolates, as in
    the reference: eps is not bounded by the lattice's."""
    import ipglasma_amd as I
    p = K.params("AuAu", N=32, Ny=4)
    ev = I.Event(p, seed=17)
    assert ev.init()
    ev.run()
    f1, f2 = tmp_path / "a.dat", tmp_path / "b.dat"
    e1 = ev.write_hydro(f1, 40, 2, 0.9 * p.L, 0.5)
    land = ev.landau()
    e2 = I.write_hydro_file(f2, land, p.L, p.itmax, p.dtau, 40, 2, 0.9 * p.L, 0.5)
    assert f1.read_text() == f2.read_text() and e1 == e2 and e1 > 0
    rows = [r.split() for r in f1.read_text().splitlines()[1:] if r]
    assert len(rows) == 2 * 40 * 40
    vals = np.array(rows, dtype=float)
    assert np.all(vals[:, 4] >= 1.0)
    assert np.all(vals[:, 3] >= 0.0) and vals[:, 3].max() > 0
    vac = vals[:, 3] == 0
    assert np.all(vals[vac, 4] == 1.0) and np.all(vals[vac, 5:] == 0.0)
    # Tmunu-t (MyEigen.cpp:801-1046): the device path against the host path
    # on the event's T^{mu nu}
    t1, t2 = tmp_path / "t1.dat", tmp_path / "t2.dat"
    ev.write_tmunu(t1, 40, 1, 0.9 * p.L, 0.5)
    I.write_tmunu_file(t2, ev.tmunu(), p.L, p.itmax, p.dtau, 40, 1, 0.9 * p.L, 0.5)
    assert t1.read_text() == t2.read_text()
    trows = np.array([r.split() for r in t1.read_text().splitlines()[1:] if r], dtype=float)
    assert trows.shape == (40 * 40, 12) and trows[:, 2].max() > 1e-16


def test_jazma_file_of_an_event(tmp_path):
    """ipg_write_jazma (MyEigen.cpp:1047-1293) on an evolved event: the same
    text as the host writer on the event's g2mu2 fields, which the evolution
    leaves as the colour charge set them; the eps column integrates to Etot."""
    import ipglasma_amd as I
    p = K.params("AuAu", N=32, Ny=4)
    ev = I.Event(p, seed=19)
    assert ev.init()
    gA, gB = ev.field("g2mu2A"), ev.field("g2mu2B")
    ev.run()
    assert np.array_equal(ev.field("g2mu2A"), gA) and np.array_equal(ev.field("g2mu2B"), gB)
    etot = ev.write_hydro(tmp_path / "h.dat", 40, 1, 0.9 * p.L, 0.5)
    f1, f2 = tmp_path / "j1.dat", tmp_path / "j2.dat"
    ev.write_jazma(f1, 40, 1, 0.9 * p.L, 0.5, etot)
    I.write_jazma_file(f2, gA, gB, p.L, p.itmax, p.dtau, 40, 1, 0.9 * p.L, 0.5, etot)
    assert f1.read_text() == f2.read_text()
    vals = np.array([r.split() for r in f1.read_text().splitlines()[1:]], dtype=float)
    ha = 0.9 * p.L / 40
    tau = p.itmax * p.dtau * p.L / p.size
    assert vals[:, 3].sum() * ha * ha * tau == pytest.approx(etot, rel=1e-4)
