"""Pins of the oracle's leapfrog coefficients, T^{mu nu} components, pi at
tau = 0 and the Q_s statistics by what the field equations and closed forms
fix, each with a planted mistake (oracle/ipg_oracle.h orc_set_mutation) that
the same check must catch.

Boost-invariant Yang-Mills conserves T^{mu nu} in Milne coordinates
(tau, x, y, eta):
    d_tau(tau T^tt) + tau d_i T^it + tau^2 T^ee = 0
    d_tau(tau T^tx) + tau d_i T^ix             = 0    (and y)
    d_tau(tau^3 T^te) + tau^3 d_i T^ie         = 0
Integrated over the transverse plane (fields that vanish at the edges) with
weights 1, x, y these give moment identities,
    M1  d/dtau [tau sum T^tt]     = -tau^2 sum T^ee
    M2  d/dtau [tau sum x T^tt]   =  tau sum T^tx - tau^2 sum x T^ee  (M3: y)
    M5  d/dtau [tau sum y T^tx]   =  tau sum T^xy                     (and x T^ty)
    M6  d/dtau [tau sum x T^tx]   =  tau sum T^xx                     (and y T^ty)
    M7  d/dtau [tau^3 sum x T^te] =  tau^3 sum T^xe                   (M8: y)
that tie every component to the dynamics: M1 to the update coefficients of
Evolution.cpp:35-293 (a wrong coefficient breaks energy balance), M2/M3 to
the energy flux T^tx, T^ty (Evolution.cpp:929-1046), M5/M6 to T^xy, T^xx,
M7/M8 to T^xe, T^ye relative to T^te.  The oracle's step is run on a smooth
blob (Gaussian envelopes, random colour directions, E = 0 and pi = 0.7 phi
at the start so that Gauss's law holds); T^{mu nu} is taken at integer times
with E and pi averaged over the two half steps.  Lattice positions: T^tt of
site (i, j) is centred at (i + 1/2, j + 1/2) (links, plaquette and the four
pi and phi terms it averages).  The absolute size of T^te, T^xe and T^ye,
which the eta identities leave free, is fixed by constant abelian fields,
where T^{mu nu} is Maxwell's: (tau T^te)^2 = 4 T_E T_B for a transverse E
and a perpendicular phi gradient, (tau T^xe)^2 = 4 T_E T_L for E^x and pi.
Not pinned: a common sign of the three eta components (boost invariance
decouples them from the rest).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sl

import _oracle as O

GEN = np.zeros((8, 3, 3), complex)
GEN[0][0, 1] = GEN[0][1, 0] = 0.5
GEN[1][0, 1], GEN[1][1, 0] = -0.5j, 0.5j
GEN[2][0, 0], GEN[2][1, 1] = 0.5, -0.5
GEN[3][0, 2] = GEN[3][2, 0] = 0.5
GEN[4][0, 2], GEN[4][2, 0] = -0.5j, 0.5j
GEN[5][1, 2] = GEN[5][2, 1] = 0.5
GEN[6][1, 2], GEN[6][2, 1] = -0.5j, 0.5j
GEN[7] = np.diag([1.0, 1.0, -2.0]) / (2 * np.sqrt(3))

MUT = {"evolve_u": 0, "evolve_phi": 1, "evolve_pi": 2, "e_plaq": 3, "e_comm": 4, "t_elec": 5, "t_long": 6,
       "t_mag": 7, "t_phigrad": 8, "tx_f": 9, "tx_k": 10, "ty_f": 11, "ty_k": 12, "te": 13, "xy_e": 14,
       "xy_b": 15, "xe_e": 16, "xe_b": 17, "ye_e": 18, "ye_b": 19, "pi_init": 20, "gauss_comm": 21}
TT, XX, YY, EE, TX, TY, TE, XY, XE, YE = range(10)


def lib():
    L = O.lib()
    L.orc_set_mutation.argtypes = [C.c_int, C.c_double]
    return L


@pytest.fixture(autouse=True)
def _clear_mutations():
    lib().orc_clear_mutations()
    yield
    lib().orc_clear_mutations()


def mutate(name, factor):
    lib().orc_set_mutation(MUT[name], factor)


def smooth_state(N, sigma, amp, seed):
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")

    def field():
        out = np.zeros((N, N, 3, 3), complex)
        for _ in range(3):
            cx, cy = N / 2 + rng.normal() * sigma * 0.6, N / 2 + rng.normal() * sigma * 0.6
            kx, ky = rng.normal(size=2) / sigma
            env = np.exp(-((i - cx) ** 2 + (j - cy) ** 2) / (2 * sigma ** 2)) * np.cos(kx * i + ky * j +
                                                                                       rng.uniform(0, 6))
            out += amp * env[..., None, None] * np.einsum("a,aij->ij", rng.normal(size=8), GEN)
        return out.reshape(-1, 3, 3)

    Ax, Ay, ph = field(), field(), field()
    return {"Ux": O.to_c(np.array([sl.expm(1j * m) for m in Ax])),
            "Uy": O.to_c(np.array([sl.expm(1j * m) for m in Ay])),
            "E1": np.zeros((N * N, 18)), "E2": np.zeros((N * N, 18)),
            "phi": O.to_c(ph), "pi": O.to_c(0.7 * ph)}


def tmunu(N, g, dtau, it, st):
    T = np.zeros((10, N * N))
    lib().orc_tmunu(N, g, 1.0, dtau, it, *[st[k] for k in O.FIELDS], T.reshape(-1))
    return T


def evolve(N, sigma, amp, dtau=0.02, tau0=2.0, burn=1.0, nmeas=3, g=2.0, seed=1):
    """oracle steps from tau0 (lattice units); returns [(tau, T)] at nmeas
    consecutive integer times after a burn-in, and the largest Gauss-law
    violation seen at those times"""
    L = lib()
    st = smooth_state(N, sigma, amp, seed)
    it0 = int(round(tau0 / dtau))
    nburn = int(round(burn / dtau))
    hist, gauss = [], 0.0
    for k in range(nburn + nmeas):
        it = it0 + k
        tau = it * dtau
        old = {f: st[f].copy() for f in ("E1", "E2", "pi")}
        L.orc_evolve_pi(N, dtau, tau, st["Ux"], st["Uy"], st["phi"], st["pi"])
        L.orc_evolve_e(N, g, dtau, tau, st["Ux"], st["Uy"], st["phi"], st["E1"], st["E2"])
        if k >= nburn:
            mid = dict(st)
            for f in ("E1", "E2", "pi"):
                mid[f] = (old[f] + st[f]) / 2
            hist.append((tau, tmunu(N, g, dtau, it, mid)))
            gauss = max(gauss, L.orc_gauss_law(N, *[st[f] for f in O.FIELDS]))
        L.orc_evolve_phi(N, dtau, tau, st["pi"], st["phi"])
        L.orc_evolve_u(N, g, dtau, tau, st["E1"], st["E2"], st["Ux"], st["Uy"])
    return hist, gauss


def moments(N, hist):
    """relative residuals of the moment identities at the middle time"""
    i, j = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    x = (i + 0.5).reshape(-1) - N / 2
    y = (j + 0.5).reshape(-1) - N / 2
    (tm, Tm), (t0, T0), (tp, Tp) = hist[0], hist[1], hist[2]

    def d(f):
        return (f(tp, Tp) - f(tm, Tm)) / (tp - tm)

    rows = {
        "M1": (d(lambda t, T: t * T[TT].sum()), -t0 ** 2 * T0[EE].sum()),
        "M2": (d(lambda t, T: t * (x * T[TT]).sum()), t0 * T0[TX].sum() - t0 ** 2 * (x * T0[EE]).sum()),
        "M3": (d(lambda t, T: t * (y * T[TT]).sum()), t0 * T0[TY].sum() - t0 ** 2 * (y * T0[EE]).sum()),
        "M5x": (d(lambda t, T: t * (y * T[TX]).sum()), t0 * T0[XY].sum()),
        "M5y": (d(lambda t, T: t * (x * T[TY]).sum()), t0 * T0[XY].sum()),
        "M6x": (d(lambda t, T: t * (x * T[TX]).sum()), t0 * T0[XX].sum()),
        "M6y": (d(lambda t, T: t * (y * T[TY]).sum()), t0 * T0[YY].sum()),
        "M7": (d(lambda t, T: t ** 3 * (x * T[TE]).sum()), t0 ** 3 * T0[XE].sum()),
        "M8": (d(lambda t, T: t ** 3 * (y * T[TE]).sum()), t0 ** 3 * T0[YE].sum()),
    }
    return {k: abs(l - r) / max(abs(l), abs(r)) for k, (l, r) in rows.items()}


# a non-abelian blob (amplitude 0.4) and a nearly abelian one (0.02).  The
# residuals come from the time step (E and pi averaged over the half steps)
# and the lattice spacing; they grow with the amplitude, and the stress
# moments M5/M6, which go through d/dtau of momentum moments, only hold on
# the weak blob (0.25 at amplitude 0.4, 4e-4 at 0.02: the lattice conserves
# momentum up to terms of relative size ~amplitude).  Tolerances are about 3x
# the measured residuals.
CONFIGS = {
    "strong": (dict(N=48, sigma=5, amp=0.4), {"M1": 2e-3, "M2": 1e-3, "M3": 1e-3, "M7": 5e-4, "M8": 5e-4}),
    "weak": (dict(N=48, sigma=5, amp=0.02), {"M1": 3e-5, "M2": 5e-5, "M3": 1e-5, "M5x": 2e-3, "M5y": 2e-3,
                                             "M6x": 4e-2, "M6y": 4e-2, "M7": 2e-4, "M8": 2e-4}),
}


def _worst_ratio():
    """largest residual / tolerance over both blobs and all identities"""
    worst = (0.0, None)
    for name, (cfg, tol) in CONFIGS.items():
        hist, _ = evolve(**cfg)
        r = moments(cfg["N"], hist)
        for k, t in tol.items():
            if r[k] / t > worst[0]:
                worst = (r[k] / t, f"{name}:{k}={r[k]:.1e}")
    return worst


def test_moment_identities_hold():
    for name, (cfg, tol) in CONFIGS.items():
        hist, gauss = evolve(**cfg)
        r = moments(cfg["N"], hist)
        print(name, {k: f"{v:.1e}" for k, v in r.items()})
        for k, t in tol.items():
            assert r[k] < t, (name, k, r[k])
        # Gauss's law, zero at the start, stays at round-off through the steps
        assert gauss < 1e-9


# each planted mistake must push some identity to at least 3x its tolerance
PLANTED = [(n, 1.1) for n in MUT if n not in ("pi_init", "gauss_comm")] + \
    [(n, -1.0) for n in ("tx_f", "tx_k", "ty_f", "ty_k", "te", "xy_e", "xy_b", "xe_e", "xe_b", "ye_e", "ye_b")]


@pytest.mark.parametrize("name,factor", PLANTED, ids=[f"{n}x{f}" for n, f in PLANTED])
def test_planted_mistake_breaks_a_moment_identity(name, factor):
    mutate(name, factor)
    ratio, where = _worst_ratio()
    assert ratio > 3, (name, factor, ratio, where)


def test_gauss_law_pins_commutator_against_pi_update():
    """Gauss's law G = sum_i (U_i^+ E_i U_i - E_i)(x - i) - i [phi, pi] stays
    at round-off only when the [phi~, phi] coefficient of the E update
    matches the pi update (Evolution.cpp:117-293); the plaquette terms drop
    out of it identically and the link update commutes with E"""
    strong = CONFIGS["strong"][0]
    _, g0 = evolve(**strong)
    assert g0 < 1e-9
    for name in ("e_comm", "evolve_pi", "gauss_comm"):
        lib().orc_clear_mutations()
        mutate(name, 1.1)
        _, g = evolve(**strong)
        assert g > 1e3 * max(g0, 1e-15), (name, g, g0)


def _const_state(N, e1=0.0, pi0=0.0, c=0.0):
    """abelian constant fields along t^3: U = 1, E1 = e1 t3, pi = pi0 t3,
    phi = c x t3 (a constant phi gradient along x)"""
    S = N * N
    t3 = GEN[2]
    i = np.repeat(np.arange(N), N)
    st = {"Ux": O.to_c(np.broadcast_to(np.eye(3), (S, 3, 3))), "Uy": O.to_c(np.broadcast_to(np.eye(3), (S, 3, 3))),
          "E1": O.to_c(np.broadcast_to(e1 * t3, (S, 3, 3))), "E2": np.zeros((S, 18)),
          "phi": O.to_c(c * i[:, None, None] * t3), "pi": O.to_c(np.broadcast_to(pi0 * t3, (S, 3, 3)))}
    return st


def _interior(T, N):
    return T.reshape(10, N, N)[:, 4, 4]


def test_eta_components_maxwell_closed_forms():
    """constant abelian fields, where T^{mu nu} is Maxwell's stress tensor:
    the momentum density along eta of a transverse E and a perpendicular
    phi gradient, (tau T^te)^2 = 4 T_E T_B, and the shear of E^x and the
    longitudinal field pi, (tau T^xe)^2 = 4 T_E T_L, with T_E, T_B, T_L the
    electric, phi-gradient and pi parts of T^tt"""
    N, g, dtau, it = 12, 2.0, 0.05, 40
    tau = it * dtau
    e, c, p = 0.3, 0.2, 0.25

    def parts():
        TE_ = _interior(tmunu(N, g, dtau, it, _const_state(N, e1=e)), N)[TT]
        TB_ = _interior(tmunu(N, g, dtau, it, _const_state(N, c=c)), N)[TT]
        TL_ = _interior(tmunu(N, g, dtau, it, _const_state(N, pi0=p)), N)[TT]
        Teb = _interior(tmunu(N, g, dtau, it, _const_state(N, e1=e, c=c)), N)
        Tel = _interior(tmunu(N, g, dtau, it, _const_state(N, e1=e, pi0=p)), N)
        return TE_, TB_, TL_, Teb, Tel

    TE_, TB_, TL_, Teb, Tel = parts()
    assert TE_ > 0 and TB_ > 0 and TL_ > 0
    assert (tau * Teb[TE]) ** 2 == pytest.approx(4 * TE_ * TB_, rel=1e-12)
    assert (tau * Tel[XE]) ** 2 == pytest.approx(4 * TE_ * TL_, rel=1e-12)
    # T^tt of the pair is the sum of its parts; no eta flux without both
    assert Teb[TT] == pytest.approx(TE_ + TB_, rel=1e-12)
    assert Tel[TE] == 0.0
    for name in ("te", "xe_e"):
        lib().orc_clear_mutations()
        mutate(name, 1.1)
        TE_, TB_, TL_, Teb, Tel = parts()
        ok_te = (tau * Teb[TE]) ** 2 == pytest.approx(4 * TE_ * TB_, rel=1e-6)
        ok_xe = (tau * Tel[XE]) ** 2 == pytest.approx(4 * TE_ * TL_, rel=1e-6)
        assert not (ok_te and ok_xe), name


# ---- pi at tau = 0 (Init.cpp:2519-2690) -----------------------------------------

def _wilson_lines_for_uniform_links(N, WA, WB):
    """V(i, j) = W^{+i}: the x links V(s) V(s + x)^+ are W at every i < N-1
    (the last is clamped to 1), the y links are 1"""
    VA = np.zeros((N * N, 3, 3), complex)
    VB = np.zeros((N * N, 3, 3), complex)
    for i in range(N):
        a = np.linalg.matrix_power(WA.conj().T, i)
        b = np.linalg.matrix_power(WB.conj().T, i)
        VA[i * N:(i + 1) * N] = a
        VB[i * N:(i + 1) * N] = b
    return O.to_c(VA), O.to_c(VB)


def _pi_init(N, g, VA, VB):
    S = N * N
    st = {k: np.zeros((S, 18)) for k in O.FIELDS}
    stats = np.zeros(5)
    lib().orc_init_fields(N, g, VA.copy(), VB.copy(), 3, *[st[k] for k in O.FIELDS], stats)
    assert stats[3] == 0
    return O.from_c(st["pi"]).reshape(N, N, 3, 3)


def test_pi_init_closed_forms():
    """pi = -(2i/g) A^eta vanishes when either nucleus is the vacuum (the
    root finder then returns the other nucleus' link) and for commuting
    (abelian) links; in the weak-field limit U^(1) = e^{i eps A},
    U^(2) = e^{i eps B} on every x link it is -(i/g) eps^2 [A, B]
    (+ O(eps^3)): the glasma's initial longitudinal field, proportional to the
    commutator of the two nuclei's fields.  Swapping the nuclei flips it."""
    N, g = 10, 2.0
    rng = np.random.default_rng(4)
    A = np.einsum("a,aij->ij", rng.normal(size=8), GEN)
    B = np.einsum("a,aij->ij", rng.normal(size=8), GEN)
    one = np.eye(3, dtype=complex)
    W = sl.expm(0.7j * A)
    for WA, WB in ((W, one), (one, W)):
        assert np.abs(_pi_init(N, g, *_wilson_lines_for_uniform_links(N, WA, WB))).max() < 1e-14
    D1, D2 = np.diag(rng.normal(size=3)).astype(complex), np.diag(rng.normal(size=3)).astype(complex)
    D1 -= np.trace(D1) / 3 * one
    D2 -= np.trace(D2) / 3 * one
    p = _pi_init(N, g, *_wilson_lines_for_uniform_links(N, sl.expm(0.6j * D1), sl.expm(0.5j * D2)))
    assert np.abs(p).max() < 1e-7  # root-finder tolerance
    comm = A @ B - B @ A
    lead = {}
    for eps in (0.02, 0.01):
        pi_ab = _pi_init(N, g, *_wilson_lines_for_uniform_links(N, sl.expm(1j * eps * A), sl.expm(1j * eps * B)))
        pi_ba = _pi_init(N, g, *_wilson_lines_for_uniform_links(N, sl.expm(1j * eps * B), sl.expm(1j * eps * A)))
        site = pi_ab[4, 4]
        lead[eps] = site / eps ** 2
        np.testing.assert_allclose(pi_ba[4, 4], -site, rtol=0, atol=1e-9 * np.abs(site).max())
    expect = -1j / g * comm
    # the O(eps) remainder halves with eps; Richardson removes it
    extrap = 2 * lead[0.01] - lead[0.02]
    err = np.abs(extrap - expect).max() / np.abs(expect).max()
    assert err < 1e-3, err
    mutate("pi_init", 1.1)
    pi_m = _pi_init(N, g, *_wilson_lines_for_uniform_links(N, sl.expm(0.01j * A), sl.expm(0.01j * B)))[4, 4]
    assert np.abs(pi_m / 0.01 ** 2 - lead[0.01]).max() > 0.05 * np.abs(expect).max()


# ---- Q_s statistics and success (Init.cpp:1423-1618) --------------------------

def test_stats_uniform_closed_forms():
    """uniform g^2 mu^2 and one collided nucleon per nucleus at the origin:
    Q_s^2 = g2mu2 QsmuRatio^2 g^2 hbarc^2 / a^2 (the inverse of the g^2 mu^2
    definition, Init.cpp:1347-1352) averaged over the sites within
    sqrt(sigma_NN / (10 pi)) of both nucleons (counted here geometrically),
    the larger of the two nuclei for <Q_s>, the smaller for Q_s,min, their
    mean for Q_s,avg; Q_s^2 S_T over all sites but the first and last;
    T_pp = a^2 sum T_A T_B / hbarc^4; and the success test"""
    HBARC = 0.1973269718
    N, Lbox, g, ratio = 32, 8.0, 2.0, 0.8
    a = Lbox / N
    S = N * N
    qc = O.QsCfg()
    qc.N, qc.L, qc.g, qc.QsmuRatio, qc.QsmuRatioB = N, Lbox, g, ratio, ratio
    sigma = 42.0
    rcut = np.sqrt(0.1 * sigma / np.pi)
    nA = (O.Nucleon * 1)()
    nB = (O.Nucleon * 1)()
    for n in (nA, nB):
        n[0].x = n[0].y = 0.0
        n[0].collided = 1
    i, j = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    x, y = -Lbox / 2 + a * i, -Lbox / 2 + a * j
    count = int(np.count_nonzero(np.sqrt(x * x + y * y) < rcut))
    for cA, cB in ((0.02, 0.05), (0.05, 0.02)):
        gA, gB = np.full(S, cA), np.full(S, cB)
        TA, TB = np.full(S, 0.3), np.full(S, 0.7)
        out = np.zeros(8)
        lib().orc_stats(C.byref(qc), sigma, 2, 0.0, nA, 1, nB, 1, gA, gB, TA, TB, out)
        qA, qB = cA * ratio ** 2 * g ** 2 * HBARC ** 2 / a ** 2, cB * ratio ** 2 * g ** 2 * HBARC ** 2 / a ** 2
        assert out[5] == count
        assert out[0] == pytest.approx(np.sqrt(max(qA, qB)), rel=1e-12)
        assert out[2] == pytest.approx(np.sqrt(min(qA, qB)), rel=1e-12)
        assert out[1] == pytest.approx(np.sqrt((qA + qB) / 2), rel=1e-12)
        assert out[4] == pytest.approx((S - 2) * min(qA, qB) * a ** 2 / HBARC ** 2, rel=1e-12)
        assert out[3] == pytest.approx(S * 0.3 * 0.7 * a ** 2 / HBARC ** 4, rel=1e-12)
        assert out[6] == pytest.approx(g * g / 4 / np.pi, rel=1e-15)
        assert out[7] == 1
        # the success test's thresholds: N_part >= 2 and Q_s^2 S_T above the minimum
        lib().orc_stats(C.byref(qc), sigma, 1, 0.0, nA, 1, nB, 1, gA, gB, TA, TB, out)
        assert out[7] == 0
        lib().orc_stats(C.byref(qc), sigma, 2, 1e9, nA, 1, nB, 1, gA, gB, TA, TB, out)
        assert out[7] == 0
    # a nucleon that did not collide opens no overlap region: no success
    nB[0].collided = 0
    lib().orc_stats(C.byref(qc), sigma, 2, 0.0, nA, 1, nB, 1, np.full(S, 0.02), np.full(S, 0.05),
                    np.full(S, 0.3), np.full(S, 0.7), out)
    assert out[5] == 0 and out[7] == 0


def test_newton_polish_converges_to_the_root():
    """the reference's root finder stops at Sum|F| <= 1e-6 with a
    forward-difference Jacobian (Init.cpp:3384-3476); the polish option adds
    exact Newton steps about the iterate.  On random link pairs these take
    the residual from the stopping level to round-off in a few steps and
    then leave the root where it is; with no polish the oracle returns the
    reference's iterate unchanged"""
    L = lib()
    L.orc_set_newton_polish.argtypes = [C.c_int]
    rng = np.random.default_rng(12)

    def F(U1, U2, U):
        W = U1 + U2
        M = W - W.conj().T + W @ U.conj().T - U @ W.conj().T
        return np.array([-np.imag(np.trace(t @ M)) for t in GEN])

    try:
        for trial in range(20):
            A, B = (np.einsum("a,aij->ij", rng.normal(size=8) * 0.6, GEN) for _ in range(2))
            U1, U2 = sl.expm(1j * A), sl.expm(1j * B)
            sols = {}
            for pol in (0, 2, 4):
                L.orc_set_newton_polish(pol)
                sols[pol], st = O.newton_link(U1, U2, seed=3, link=trial)
                assert st.converged
            r0, r2, r4 = (np.abs(F(U1, U2, sols[k])).sum() for k in (0, 2, 4))
            assert r0 < 1e-5
            assert r2 < 1e-13 and r4 < 1e-13, (trial, r0, r2, r4)
            assert np.abs(sols[4] - sols[2]).max() < 1e-13
            assert np.abs(sols[4] - sols[0]).max() < 1e-5
    finally:
        L.orc_set_newton_polish(0)
