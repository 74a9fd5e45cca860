"""ctypes binding of the C oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the CPU restatement of the reference
hot path and is used as the checker by the tests, by __graft_entry__.smoke()
and by bench.py's cpu_baseline leg.  The product never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint64 * 312), ("mti", C.c_int), ("iset", C.c_int),
                ("gset", C.c_double)]


class Nucleon(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
                ("phi", C.c_double), ("collided", C.c_int), ("proton", C.c_int)]


class GlauberCfg(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("A1", "Z1", "A2", "Z2")] + \
        [(n, C.c_double) for n in ("R1", "a1", "beta2_1", "beta3_1", "beta4_1",
                                   "gamma_1", "dR1", "da1", "R2", "a2", "beta2_2",
                                   "beta3_2", "beta4_2", "gamma_2", "dR2", "da2")] + \
        [("force_dmin", C.c_int), ("d_min", C.c_double), ("bmin", C.c_double),
         ("bmax", C.c_double), ("linear_b", C.c_int), ("rotate_rp", C.c_int),
         ("xi", C.c_double), ("sigma_nn_mb", C.c_double),
         ("gaussian_wounding", C.c_int)]


class EventInfo(C.Structure):
    _fields_ = [("b", C.c_double), ("phiRP", C.c_double), ("Ncoll", C.c_int),
                ("Npart", C.c_int)]


class QsCfg(C.Structure):
    _fields_ = [("N", C.c_int)] + [(n, C.c_double) for n in (
        "L", "g", "BG", "rmax", "QsmuRatio", "QsmuRatioB", "xFactor", "roots",
        "rapidity")] + [("useFluctuatingx", C.c_int)]


class NewtonStat(C.Structure):
    _fields_ = [("iterations", C.c_int), ("restarts", C.c_int),
                ("converged", C.c_int), ("used_analytic_jacobian", C.c_int)]


_lib = None


def build():
    """Compile oracle/liboracle.so with its Makefile."""
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_mt_init.argtypes = [C.POINTER(MT), C.c_uint64]
        L.orc_mt_fill.argtypes = [C.POINTER(MT), C.c_int, C.c_int64, _dp]
        L.orc_philox4x32.argtypes = [C.POINTER(C.c_uint32 * 4), C.POINTER(C.c_uint32 * 2),
                                     C.POINTER(C.c_uint32 * 4)]
        L.orc_rho_normals_slice.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_newton_restart_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, _dp]
        L.orc_expm_coeff.argtypes = [_dp, _dp]
        L.orc_u_from_exponent.argtypes = [_dp, _dp]
        L.orc_fft2d.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_sample_event.argtypes = [C.POINTER(MT), C.POINTER(GlauberCfg),
                                       C.POINTER(Nucleon), C.POINTER(Nucleon),
                                       C.POINTER(EventInfo)]
        L.orc_tp_sum.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.POINTER(Nucleon), C.c_int, _dp]
        L.orc_nuclear_qs2.argtypes = [_dp, _dp, C.c_double, C.c_double]
        L.orc_nuclear_qs2.restype = C.c_double
        L.orc_g2mu2.argtypes = [C.POINTER(QsCfg), _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_stats.argtypes = [C.POINTER(QsCfg), C.c_double, C.c_int, C.c_double,
                                C.POINTER(Nucleon), C.c_int, C.POINTER(Nucleon), C.c_int,
                                _dp, _dp, _dp, _dp, _dp]
        L.orc_setv.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                               _dp, _dp, C.c_void_p, C.c_void_p, C.c_uint64, _dp]
        L.orc_newton_link.argtypes = [_dp, _dp, C.c_uint64, C.c_uint32, _dp,
                                      C.POINTER(NewtonStat)]
        L.orc_init_fields.argtypes = [C.c_int, C.c_double, _dp, _dp, C.c_uint64] + [_dp] * 7
        L.orc_evolve_u.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_evolve_phi.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_evolve_pi.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_evolve_e.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                   _dp, _dp, _dp]
        L.orc_run.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int] + [_dp] * 8
        L.orc_tmunu.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int] + [_dp] * 7
        L.orc_gauss_law.argtypes = [C.c_int] + [_dp] * 6
        L.orc_gauss_law.restype = C.c_double
        L.orc_dedeta.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.orc_dedeta.restype = C.c_double
        L.orc_time_steps.argtypes = [C.c_double, C.c_int, C.c_double,
                                     C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.orc_eig4.argtypes = [_dp] * 5
        L.orc_eig4.restype = C.c_int
        L.orc_landau.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp,
                                 C.POINTER(C.c_int)]
        L.orc_eccentricity.argtypes = [C.c_int] + [C.c_double] * 4 + [_dp] * 5
        L.orc_expm_pade.argtypes = [_dp, _dp]
        L.orc_expm_pade.restype = C.c_int
        L.orc_reunitarize.argtypes = [_dp, _dp]
        L.orc_fftn_shifted.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_gauge_fix.argtypes = [C.c_int] + [C.c_void_p] * 6 + [C.c_int, C.POINTER(C.c_int),
                                                                 C.POINTER(C.c_double), C.c_void_p]
        L.orc_multiplicity.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                       _dp, _dp, _dp, _dp]
        L.orc_run_multiplicity.argtypes = ([C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
                                           + [_dp] * 11)
        _lib = L
    return _lib


# ---- thin pythonic helpers ---------------------------------------------------

def mt_stream(seed, kind, n):
    st = MT()
    lib().orc_mt_init(C.byref(st), seed)
    out = np.zeros(n)
    lib().orc_mt_fill(C.byref(st), kind, n, out)
    return out


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32(C.byref(c), C.byref(k), C.byref(o))
    return list(o)


def rho_normals(seed, nucleus, k, S):
    z = np.zeros(S * 8)
    lib().orc_rho_normals_slice(seed, nucleus, k, S, z)
    return z.reshape(S, 8)


def expm_coeff(Q):
    out = np.zeros(18)
    lib().orc_expm_coeff(np.ascontiguousarray(Q, dtype=np.float64), out)
    return out[0::2] + 1j * out[1::2]


def exp_iQt(Q):
    out = np.zeros(18)
    lib().orc_u_from_exponent(np.ascontiguousarray(Q, dtype=np.float64), out)
    return (out[0::2] + 1j * out[1::2]).reshape(3, 3)


def fft2d(x, sign):
    N = x.shape[0]
    buf = np.ascontiguousarray(np.stack([x.real, x.imag], -1).reshape(-1))
    lib().orc_fft2d(N, buf, sign)
    b = buf.reshape(N, N, 2)
    return b[..., 0] + 1j * b[..., 1]


def to_c(m):
    """complex [..., 3, 3] -> interleaved float64 [..., 18]"""
    m = np.asarray(m, dtype=np.complex128)
    return np.ascontiguousarray(np.stack([m.real, m.imag], -1).reshape(m.shape[:-2] + (18,)))


def from_c(a):
    a = np.asarray(a).reshape(a.shape[:-1] + (9, 2))
    return (a[..., 0] + 1j * a[..., 1]).reshape(a.shape[:-2] + (3, 3))


def newton_link(U1, U2, seed=1, link=0):
    out = np.zeros(18)
    st = NewtonStat()
    lib().orc_newton_link(to_c(U1), to_c(U2), seed, link, out, C.byref(st))
    return from_c(out), st


def time_steps(L, N, maxtime):
    d = C.c_double()
    it = C.c_int()
    lib().orc_time_steps(L, N, maxtime, C.byref(d), C.byref(it))
    return d.value, it.value


def eig4(A):
    """eigenvalues (complex [4]) and eigenvectors (complex columns [4, 4])"""
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(16)
    wr, wi, vr, vi = (np.zeros(4), np.zeros(4), np.zeros(16), np.zeros(16))
    rc = lib().orc_eig4(A, wr, wi, vr, vi)
    return rc, wr + 1j * wi, (vr + 1j * vi).reshape(4, 4)


def landau(N, L, dtau, it, T):
    """T [10, S] -> (out [15, S], avg [4], failures)"""
    T = np.ascontiguousarray(T, dtype=np.float64).reshape(-1)
    out = np.zeros(15 * N * N)
    avg = np.zeros(4)
    nf = C.c_int()
    lib().orc_landau(N, L, dtau, it, T, out, avg, C.byref(nf))
    return out.reshape(15, N * N), avg, nf.value


def eccentricity(N, L, eps, utau, g2A=None, g2B=None, QsmuRatio=0.8, QsmuRatioB=0.8, cutoff=0.):
    S = N * N
    g2A = np.zeros(S) if g2A is None else np.ascontiguousarray(g2A, dtype=np.float64)
    g2B = np.zeros(S) if g2B is None else np.ascontiguousarray(g2B, dtype=np.float64)
    res = np.zeros(19)
    lib().orc_eccentricity(N, L, QsmuRatio, QsmuRatioB, cutoff,
                           np.ascontiguousarray(eps, dtype=np.float64),
                           np.ascontiguousarray(utau, dtype=np.float64), g2A, g2B, res)
    return res


def new_state(N):
    S = N * N
    return {k: np.zeros((S, 18)) for k in ("Ux", "Uy", "E1", "E2", "phi", "pi")}


FIELDS = ("Ux", "Uy", "E1", "E2", "phi", "pi")


# ---- Coulomb gauge and gluon spectra ------------------------------------------

def expm_pade(A):
    out = np.zeros(18)
    rc = lib().orc_expm_pade(to_c(A), out)
    assert rc == 0
    return from_c(out)


def reunitarize(M):
    out = np.zeros(18)
    lib().orc_reunitarize(to_c(M), out)
    return from_c(out)


def fftn_shifted(f, sign):
    """[S, 18] field -> its quadrant-swapped transform"""
    N = int(round(np.sqrt(f.shape[0])))
    w = np.ascontiguousarray(f, dtype=np.float64).copy()
    lib().orc_fftn_shifted(N, w, sign)
    return w


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def gauge_fix(N, st, max_iter=4000, history=False):
    """in place on the state dict; returns (iterations, residual[, history])"""
    it = C.c_int()
    res = C.c_double()
    h = np.zeros(max_iter) if history else None
    for k in FIELDS:
        assert st[k].flags.c_contiguous and st[k].dtype == np.float64
    lib().orc_gauge_fix(N, *[_ptr(st[k]) for k in FIELDS], max_iter, C.byref(it), C.byref(res), _ptr(h))
    if history:
        return it.value, res.value, h[:it.value + 1 if it.value < max_iter else max_iter]
    return it.value, res.value


def multiplicity(N, L, g, dtau, it, E1, E2, pi):
    out = np.zeros(308)
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    lib().orc_multiplicity(N, L, g, dtau, it, c(E1), c(E2), c(pi), out)
    return out


def run_multiplicity(N, g, L, dtau, itmax, st):
    T = np.zeros(10 * N * N)
    obs, m1, mmax, gf = np.zeros(3), np.zeros(308), np.zeros(308), np.zeros(4)
    lib().orc_run_multiplicity(N, g, L, dtau, itmax, *[st[k] for k in FIELDS], T, obs, m1, mmax, gf)
    return T.reshape(10, N * N), obs, m1, mmax, gf
