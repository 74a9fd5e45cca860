"""Shared parity cases: parameter sets for the product and the matching
oracle configuration.  The nucleus constants are Glauber.cpp:19-245; the
collision setups are the Pb+Pb, Au+Au, O+O and p+Pb throughput cases,
scaled down in N and Ny so that the oracle finishes in seconds."""
import ctypes as C

import numpy as np

import _oracle as O

# Glauber.cpp:19-245: A, Z, R_WS, a_WS, beta2, beta3, beta4
NUCLEI = {
    "Au": (197, 79, 6.37, 0.535, -0.13, 0.0, -0.03),
    "Pb": (208, 82, 6.62, 0.546, 0.0, 0.0, 0.0),
    "O": (16, 8, 2.608, 0.513, -0.01, 0.0, -0.122),
    "p": (1, 1, 1.0, 1.0, 0.0, 0.0, 0.0),
    "d": (2, 1, 1.0, 0.228, 0.0, 0.0, 0.0),  # Hulthen: a in the a_WS slot, R_WS 1
}

# (projectile, target, R_WS, a_WS, sigma_NN, roots, L, bmin, bmax, linear_b)
SYSTEMS = {
    "AuAu": ("Au", "Au", 6.38, 0.535, 42.0, 200.0, 30.0, 5.0, 5.0, 1),
    "PbPb": ("Pb", "Pb", 6.62, 0.546, 70.0, 5020.0, 30.0, 0.0, 0.0, 1),
    "OO": ("O", "O", 2.61, 0.513, 70.0, 5020.0, 15.0, 0.0, 3.0, 1),
    "pPb": ("p", "Pb", 6.62, 0.546, 70.0, 5020.0, 10.0, 0.0, 5.0, 0),
    "PbPbdef": ("Pb", "Pb", 6.62, 0.546, 70.0, 5020.0, 30.0, 0.0, 4.0, 1),
    "dAu": ("d", "Au", 6.38, 0.535, 42.0, 200.0, 20.0, 0.0, 3.0, 1),
}


def params(system="AuAu", N=32, Ny=4, maxtime=0.4, seed=11, ws_override=True, **kw):
    import ipglasma_amd as I
    pr, tg, R, a, sig, roots, L, bmin, bmax, lin = SYSTEMS[system]
    p = I.default_params()
    p.set(size=N, Ny=Ny, L=L, g=2.0, m=0.2, BG=4.0, UVdamp=0.0, QsmuRatio=0.8, rmax=1e7,
          roots=roots, Rapidity=0.0, useFluctuatingx=1, xFromThisFactorTimesQs=1.0, SigmaNN=sig,
          maxtime=maxtime, Projectile=pr, Target=tg, setWSDeformParams=int(ws_override),
          R_WS=R, a_WS=a, beta2=0.0, beta3=0.0, beta4=0.0, gamma=0.0, force_dmin_flag=0,
          d_min=0.9, dR_np=0.0, da_np=0.0, bmin=bmin, bmax=bmax,
          samplebFromLinearDistribution=lin, rotateReactionPlane=1, gaussianWounding=1,
          computeGluonMultiplicity=0, writeOutputs=0, useTimeForSeed=0, seed=seed,
          minimumQs2ST=0, protonAnisotropy=0.0)
    p.set(**kw)
    return p


def glauber_cfg(p):
    """orc_glauber_cfg for a product parameter set (FindNucleusData2 with the
    setWSDeformParams override, Glauber.cpp:246-262)."""
    c = O.GlauberCfg()
    for tag, name in (("1", p.Projectile.decode()), ("2", p.Target.decode())):
        A, Z, R, a, b2, b3, b4 = NUCLEI[name]
        g = 0.0
        dR = da = 0.0
        if p.setWSDeformParams:
            R, a, b2, b3, b4, g, dR, da = p.R_WS, p.a_WS, p.beta2, p.beta3, p.beta4, p.gamma, p.dR_np, p.da_np
        setattr(c, "A" + tag, A)
        setattr(c, "Z" + tag, Z)
        for k, v in (("R", R), ("a", a), ("beta2_", b2), ("beta3_", b3), ("beta4_", b4),
                     ("gamma_", g), ("dR", dR), ("da", da)):
            setattr(c, k + tag, v)
    c.force_dmin = p.force_dmin_flag
    c.d_min = p.d_min
    c.bmin, c.bmax = p.bmin, p.bmax
    c.linear_b = p.samplebFromLinearDistribution
    c.rotate_rp = p.rotateReactionPlane
    c.xi = p.protonAnisotropy
    c.sigma_nn_mb = p.SigmaNN
    c.gaussian_wounding = p.gaussianWounding
    return c


def qs_cfg(p):
    c = O.QsCfg()
    c.N = p.size
    c.L, c.g, c.BG, c.rmax = p.L, p.g, p.BG, p.rmax
    c.QsmuRatio = c.QsmuRatioB = p.QsmuRatio
    c.xFactor, c.roots, c.rapidity = p.xFromThisFactorTimesQs, p.roots, p.Rapidity
    c.useFluctuatingx = p.useFluctuatingx
    return c


def load_table(path):
    d = np.loadtxt(path)
    T = np.ascontiguousarray(d[::44, 1])
    Q = np.ascontiguousarray(d[:, 2])
    return T, Q


def oracle_event(p, seed):
    """Oracle nucleons + colour charge for one attempt from a fresh MT stream."""
    cfg = glauber_cfg(p)
    st = O.MT()
    O.lib().orc_mt_init(C.byref(st), seed)
    nA = (O.Nucleon * cfg.A1)()
    nB = (O.Nucleon * cfg.A2)()
    info = O.EventInfo()
    rc = O.lib().orc_sample_event(C.byref(st), C.byref(cfg), nA, nB, C.byref(info))
    assert rc == 0
    return st, nA, nB, info


def nucleon_array(arr):
    return np.array([[q.x, q.y, q.z, q.phi, q.collided, q.proton] for q in arr])


def oracle_color_charge(p, nA, nB, table):
    N = p.size
    S = N * N
    TpA, TpB = np.zeros(S), np.zeros(S)
    O.lib().orc_tp_sum(N, p.L, p.BG, p.protonAnisotropy, nA, len(nA), TpA)
    O.lib().orc_tp_sum(N, p.L, p.BG, p.protonAnisotropy, nB, len(nB), TpB)
    gA, gB = np.zeros(S), np.zeros(S)
    qc = qs_cfg(p)
    T, Q = table
    O.lib().orc_g2mu2(C.byref(qc), T, Q, TpA, TpB, gA, gB)
    return TpA, TpB, gA, gB


def c3_params():
    """The 64-event heterogeneous batch: 16 p+Pb (N=256, L=10 fm, b uniform
    in [0,5] fm), 16 O+O (N=256, L=15 fm, R=2.61 fm, a=0.513 fm), 24 Pb+Pb
    (N=512, L=30 fm, b uniform in [0,12] fm), 8 Pb+Pb (N=768, L=30 fm, b=0);
    Ny=50, tau_max=0.4 fm/c."""
    def mk(N, L, pr, tg, R, a, bmax):
        return params("PbPb", N=N, Ny=50, L=L, Projectile=pr, Target=tg, R_WS=R, a_WS=a, bmin=0.0, bmax=bmax,
                      samplebFromLinearDistribution=0)
    return ([mk(256, 10.0, "p", "Pb", 6.62, 0.546, 5.0)] * 16 + [mk(256, 15.0, "O", "O", 2.61, 0.513, 0.0)] * 16 +
            [mk(512, 30.0, "Pb", "Pb", 6.62, 0.546, 12.0)] * 24 + [mk(768, 30.0, "Pb", "Pb", 6.62, 0.546, 0.0)] * 8)
