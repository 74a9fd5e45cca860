"""ipglasma_amd -- MI355X-native IP-Glasma solver core (Python binding).

The product is the C-ABI shared library ``libipglasma_amd.so`` (HIP kernels
for gfx950 plus host C++), declared in ``include/ipg.h``.  This module binds
it with ctypes for tests, the benchmark and scripting.  There is no CPU
fallback: every call that needs the device fails loudly when the library is
missing or no HIP device is present.

The classes mirror the reference's call surface (schenke/ipglasma):
``Setup`` (Setup.cpp:24-104) parses the input file into ``Parameters``;
``Event`` owns one Init/Evolution/Lattice triple (main.cpp:167-187) with
``init()`` (Init::init) and ``run()`` (Evolution::run); ``Batch`` runs many
events through one launch sequence.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libipglasma_amd.so")
# the Q_s^2(T_p, y) table the reference ships (data, Init.cpp:441-480)
QS_TABLE = os.path.join(_HERE, "data", "qs2Adj_vs_Tp_vs_Y_200.in")

FIELDS = {"Ux": 0, "Uy": 1, "E1": 2, "E2": 3, "phi": 4, "pi": 5, "VA": 6, "VB": 7,
          "g2mu2A": 8, "g2mu2B": 9, "TpA": 10, "TpB": 11}
TMUNU = ("tautau", "xx", "yy", "etaeta", "taux", "tauy", "taueta", "xy", "xeta", "yeta")


class IPGError(RuntimeError):
    pass


_INT_FIELDS_1 = ("size", "Nc", "Ny")
_DBL_FIELDS_1 = ("L", "g", "m", "UVdamp", "BG", "QsmuRatio", "rmax", "roots", "Rapidity",
                 "xFromThisFactorTimesQs", "SigmaNN", "maxtime")
_INT_FIELDS_2 = ("inverseQsForMaxTime", "useFluctuatingx", "usePseudoRapidity", "useNucleus",
                 "useSmoothNucleus", "useConstituentQuarkProton", "smearQs", "runningCoupling",
                 "gaussianWounding", "nucleonPositionsFromFile", "rotateReactionPlane",
                 "samplebFromLinearDistribution", "setWSDeformParams", "force_dmin_flag",
                 "useFixedNpart", "averageOverThisManyNuclei", "useSeedList", "useTimeForSeed",
                 "computeGluonMultiplicity", "writeOutputs", "readInitialWilsonLines",
                 "writeInitialWilsonLines", "SubNucleonParamType")
_DBL_FIELDS_2 = ("R_WS", "a_WS", "beta2", "beta3", "beta4", "gamma", "d_min", "dR_np", "da_np",
                 "bmin", "bmax", "protonAnisotropy", "minimumQs2ST", "smearingWidth",
                 "Jacobianm", "g2mu")


class Parameters(C.Structure):
    """ipg_params (include/ipg.h); the keys of readInput (main.cpp:448-578)."""
    _fields_ = ([(n, C.c_int) for n in _INT_FIELDS_1] + [(n, C.c_double) for n in _DBL_FIELDS_1] +
                [(n, C.c_int) for n in _INT_FIELDS_2] + [(n, C.c_double) for n in _DBL_FIELDS_2] +
                [("seed", C.c_ulonglong), ("Projectile", C.c_char * 16), ("Target", C.c_char * 16),
                 ("NucleusQsTableFileName", C.c_char * 512), ("dtau", C.c_double), ("itmax", C.c_int),
                 ("rho_rng", C.c_int), ("newton_polish", C.c_int), ("sizeOutput", C.c_int), ("etaSizeOutput", C.c_int),
                 ("LOutput", C.c_double), ("detaOutput", C.c_double),
                 ("muZero", C.c_double), ("c", C.c_double), ("runWithThisFactorTimesQs", C.c_double)])

    def set(self, **kw):
        for k, v in kw.items():
            if k in ("Projectile", "Target", "NucleusQsTableFileName"):
                v = v.encode() if isinstance(v, str) else v
            setattr(self, k, v)
        lib().ipg_derive(C.byref(self))
        return self

    def copy(self):
        c = Parameters()
        C.pointer(c)[0] = self
        return c


class Observables(C.Structure):
    _fields_ = ([("success", C.c_int), ("Npart", C.c_int), ("Ncoll", C.c_int)] +
                [(n, C.c_double) for n in ("b", "phiRP", "averageQs", "averageQsAvg", "averageQsmin",
                                           "Tpp", "Qs2minST", "alphas", "dEdeta", "gauss_max",
                                           "tau_fm")] +
                [("newton_iterations", C.c_longlong)] +
                [(n, C.c_int) for n in ("newton_max_iterations", "newton_restarts",
                                        "newton_unconverged", "newton_analytic_jacobian",
                                        "qs_serial_fallback")] +
                [("rho_seed", C.c_uint64), ("newton_seed", C.c_uint64)] +
                [("eccentricity", C.c_double * 6), ("Psi", C.c_double * 6)] +
                [(n, C.c_double) for n in ("area", "Rbar", "avg_eps", "avg_qs2A_qs2B", "rms_r")] +
                [("centroid", C.c_double * 2), ("u_rms", C.c_double * 3)] +
                [("landau_failures", C.c_int), ("landau_no_timelike", C.c_int)])

    def as_dict(self):
        d = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            d[k] = list(v) if isinstance(v, C.Array) else v
        return d


class Multiplicity(C.Structure):
    _fields_ = ([(n, C.c_double) for n in ("dNdy", "dEdy", "dNdy_kt", "dEdy_kt", "dNdy_cut3", "dEdy_cut3",
                                           "dNdy_cut6", "dEdy_cut6", "mean_pt", "kt_bin")] +
                [("n_kt", C.c_double * 100), ("E_kt", C.c_double * 100), ("counts", C.c_int * 100),
                 ("gauge_iterations", C.c_int), ("gauge_residual", C.c_double), ("tau_fm", C.c_double)])

    def as_dict(self):
        d = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            d[k] = np.array(v) if isinstance(v, C.Array) else v
        return d


_lib = None


def lib():
    """Load the in-tree library; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IPGError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        sig = {
            "ipg_default_params": (None, [C.POINTER(Parameters)]),
            "ipg_read_input": (C.c_int, [C.c_char_p, C.POINTER(Parameters)]),
            "ipg_derive": (None, [C.POINTER(Parameters)]),
            "ipg_last_error": (C.c_char_p, []),
            "ipg_version": (C.c_int, []),
            "ipg_release_cached_memory": (C.c_int, []),
            "ipg_struct_size": (C.c_int, [C.c_int]),
            "ipg_create": (vp, [C.POINTER(Parameters)]),
            "ipg_destroy": (None, [vp]),
            "ipg_load_qs_table": (C.c_int, [vp, C.c_char_p]),
            "ipg_seed": (C.c_int, [vp, C.c_uint64]),
            "ipg_init": (C.c_int, [vp, ip]),
            "ipg_init_nucleons": (C.c_int, [vp]),
            "ipg_init_color_charge": (C.c_int, [vp, ip]),
            "ipg_init_wilson_lines": (C.c_int, [vp]),
            "ipg_init_fields": (C.c_int, [vp]),
            "ipg_write_wilson_line_file": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_double, dp, C.c_int]),
            "ipg_read_wilson_line_file": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                                    dp]),
            "ipg_write_wilson_lines": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_int]),
            "ipg_write_eccentricities": (C.c_int, [vp, C.c_char_p, C.c_int]),
            "ipg_write_hydro": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dp]),
            "ipg_write_tmunu": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]),
            "ipg_write_jazma": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                          C.c_double]),
            "ipg_write_jazma_file": (C.c_int, [C.c_char_p, dp, dp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                               C.c_int, C.c_double, C.c_double, C.c_double]),
            "ipg_write_tmunu_file": (C.c_int, [C.c_char_p, dp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                               C.c_int, C.c_double, C.c_double]),
            "ipg_write_hydro_file": (C.c_int, [C.c_char_p, dp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                               C.c_int, C.c_double, C.c_double, dp]),
            "ipg_eccentricity_extents": (C.c_int, [dp, dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                                   dp, dp]),
            "ipg_read_wilson_lines": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_int]),
            "ipg_evolve": (C.c_int, [vp]),
            "ipg_evolve_step": (C.c_int, [vp, C.c_int]),
            "ipg_get_observables": (C.c_int, [vp, C.POINTER(Observables)]),
            "ipg_get_field": (C.c_int, [vp, C.c_int, dp]),
            "ipg_set_field": (C.c_int, [vp, C.c_int, dp]),
            "ipg_get_tmunu": (C.c_int, [vp, dp]),
            "ipg_get_landau": (C.c_int, [vp, dp]),
            "ipg_autotune_report": (C.c_int, [C.c_char_p, C.c_int]),
            "ipg_adjoint_dedeta": (C.c_int, [vp, dp, dp]),
            "ipg_adjoint_rho": (C.c_int, [vp, dp, dp]),
            "ipg_get_rho_grad": (C.c_int, [vp, dp]),
            "ipg_set_checkpoint_interval": (C.c_int, [vp, C.c_int]),
            "ipg_set_keep_slices": (C.c_int, [vp, C.c_int]),
            "ipg_get_checkpoint_interval": (C.c_int, [vp, ip]),
            "ipg_batch_adjoint": (C.c_int, [vp, C.c_int, dp, C.c_int, dp]),
            "ipg_batch_get_rho_grad": (C.c_int, [vp, C.c_int, dp]),
            "ipg_batch_set_checkpoint_interval": (C.c_int, [vp, C.c_int]),
            "ipg_batch_adjoint_info": (C.c_int, [vp, C.POINTER(C.c_ulonglong), ip]),
            "ipg_adjoint_phases": (C.c_int, [vp, dp]),
            "ipg_batch_adjoint_phases": (C.c_int, [vp, dp]),
            "ipg_adjoint_loss": (C.c_int, [vp, C.c_int, dp, C.c_int, dp, dp]),
            "ipg_landau": (C.c_int, [vp, dp, C.c_int]),
            "ipg_landau_noalloc": (C.c_int, [C.POINTER(Parameters), vp, vp, vp, C.c_int, vp,
                                             C.POINTER(Observables)]),
            "ipg_gauge_fix": (C.c_int, [vp, C.c_int, ip, dp]),
            "ipg_gluon_multiplicity": (C.c_int, [vp, C.c_int, C.POINTER(Multiplicity)]),
            "ipg_write_multiplicity": (C.c_int, [vp, C.c_char_p, C.POINTER(Multiplicity)]),
            "ipg_get_nucleons": (C.c_int, [vp, dp, ip, ip]),
            "ipg_set_nucleons": (C.c_int, [vp, dp, C.c_int, C.c_int]),
            "ipg_get_newton_iterations": (C.c_int, [vp, ip]),
            "ipg_set_rho_normals": (C.c_int, [vp, dp]),
            "ipg_get_rho_normals": (C.c_int, [vp, C.c_int, C.c_int, dp]),
            "ipg_enable_timing": (C.c_int, [vp, C.c_int]),
            "ipg_get_timing": (C.c_int, [vp, dp]),
            "ipg_batch_create": (vp, [C.POINTER(Parameters), C.c_int, C.POINTER(C.c_uint64)]),
            "ipg_batch_destroy": (None, [vp]),
            "ipg_batch_load_qs_table": (C.c_int, [vp, C.c_char_p]),
            "ipg_batch_run": (C.c_int, [vp, C.c_int, C.c_int]),
            "ipg_batch_get_observables": (C.c_int, [vp, C.c_int, C.POINTER(Observables)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if (L.ipg_struct_size(0) != C.sizeof(Parameters) or L.ipg_struct_size(1) != C.sizeof(Observables) or
                L.ipg_struct_size(2) != C.sizeof(Multiplicity)):
            raise IPGError("ipg_params / ipg_observables / ipg_multiplicity layout mismatch between ipg.h and the "
                           "binding")
        _lib = L
    return _lib


def release_cached_memory():
    """hand the device blocks kept for later contexts back to the device"""
    _check(lib().ipg_release_cached_memory())


def _check(rc):
    if rc != 0:
        raise IPGError(f"status {rc}: {lib().ipg_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params(**kw):
    p = Parameters()
    lib().ipg_default_params(C.byref(p))
    return p.set(**kw)


class Setup:
    """Input-file reader with the reference's key/value semantics
    (Setup::StringFind/DFind/IFind/ULLIFind, Setup.cpp:24-104)."""

    @staticmethod
    def read(path, base=None):
        p = base.copy() if base is not None else default_params()
        _check(lib().ipg_read_input(path.encode(), C.byref(p)))
        return p


class Event:
    """One event: the reference's Init(nn), Evolution(nn) and Lattice for one
    parameter set (main.cpp:167-187) on the GPU."""

    def __init__(self, params, qs_table=None, seed=None):
        self.p = params
        self.N = params.size
        self.S = self.N * self.N
        self._c = lib().ipg_create(C.byref(params))
        if not self._c:
            raise IPGError(lib().ipg_last_error().decode())
        table = qs_table or QS_TABLE
        _check(lib().ipg_load_qs_table(self._c, table.encode()))
        _check(lib().ipg_seed(self._c, params.seed if seed is None else seed))

    def close(self):
        if self._c:
            lib().ipg_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # Init::init and its stages
    def init(self):
        ok = C.c_int(0)
        _check(lib().ipg_init(self._c, C.byref(ok)))
        return bool(ok.value)

    def init_nucleons(self):
        _check(lib().ipg_init_nucleons(self._c))

    def init_color_charge(self):
        ok = C.c_int(0)
        _check(lib().ipg_init_color_charge(self._c, C.byref(ok)))
        return bool(ok.value)

    def init_wilson_lines(self):
        _check(lib().ipg_init_wilson_lines(self._c))

    # Init.cpp:1857-1980: the writeInitialWilsonLines output of setV
    def write_wilson_lines(self, path_a, path_b, fmt=2):
        _check(lib().ipg_write_wilson_lines(self._c, os.fsencode(path_a), os.fsencode(path_b), fmt))

    # Init::init with READFROMFILE > 0 (Init.cpp:2319-2369, readV 1988-2301),
    # in place of init_nucleons, init_color_charge and init_wilson_lines: b
    # from the first draw, no nucleons; then init_fields
    def read_wilson_lines(self, path_a, path_b, fmt=2):
        _check(lib().ipg_read_wilson_lines(self._c, os.fsencode(path_a), os.fsencode(path_b), fmt))

    # one line of eccentricities<id>.dat (Evolution.cpp:1797-1809) for the
    # Landau frame of step it (default: itmax, where Evolution::run takes it)
    def write_eccentricities(self, path, it=None):
        it = self.p.itmax if it is None else it
        _check(lib().ipg_write_eccentricities(self._c, os.fsencode(path), it))

    # epsilon-u-Hydro file (MyEigen.cpp:326-800) of the Landau frame of step
    # it (default itmax); returns the reference's printed Etot
    def write_hydro(self, path, size_out, eta_size_out, L_out, deta_out, it=None):
        it = self.p.itmax if it is None else it
        e = C.c_double()
        _check(lib().ipg_write_hydro(self._c, os.fsencode(path), it, size_out, eta_size_out, L_out, deta_out,
                                     C.byref(e)))
        return e.value

    # Tmunu-t file (MyEigen.cpp:801-1046) of T^{mu nu} at itmax
    def write_tmunu(self, path, size_out, eta_size_out, L_out, deta_out, it=None):
        it = self.p.itmax if it is None else it
        _check(lib().ipg_write_tmunu(self._c, os.fsencode(path), it, size_out, eta_size_out, L_out, deta_out))

    # Jazma-Hydro file (MyEigen.cpp:1047-1293), scaled to Etot
    def write_jazma(self, path, size_out, eta_size_out, L_out, deta_out, Etot, it=None):
        it = self.p.itmax if it is None else it
        _check(lib().ipg_write_jazma(self._c, os.fsencode(path), it, size_out, eta_size_out, L_out, deta_out, Etot))

    def init_fields(self):
        _check(lib().ipg_init_fields(self._c))

    # Evolution::run
    def run(self):
        _check(lib().ipg_evolve(self._c))

    def step(self, it):
        _check(lib().ipg_evolve_step(self._c, it))

    def observables(self):
        o = Observables()
        _check(lib().ipg_get_observables(self._c, C.byref(o)))
        return o.as_dict()

    # GaugeFix::FFTChi on the current fields
    def gauge_fix(self, max_iterations=4000):
        it, res = C.c_int(), C.c_double()
        _check(lib().ipg_gauge_fix(self._c, max_iterations, C.byref(it), C.byref(res)))
        return it.value, res.value

    # Evolution::multiplicity at step it (gauge fix included)
    def multiplicity(self, it):
        m = Multiplicity()
        _check(lib().ipg_gluon_multiplicity(self._c, it, C.byref(m)))
        self._last_mult = m
        return m.as_dict()

    def write_multiplicity(self, path, m=None):
        m = self._last_mult if m is None else m
        _check(lib().ipg_write_multiplicity(self._c, os.fsencode(path), C.byref(m)))

    def run_with_multiplicity(self):
        """Evolution::run with computeGluonMultiplicity: steps 0..itmax with
        the gauge fix and spectra after steps 1 and itmax; returns both"""
        itmax = self.p.itmax
        out = {}
        for it in range(itmax + 1):
            self.step(it)
            if it >= 1 and (it == 1 or it == itmax):
                out[it] = self.multiplicity(it)
        return out

    def field(self, name):
        which = FIELDS[name]
        out = np.zeros((self.S, 18) if which < 8 else self.S)
        _check(lib().ipg_get_field(self._c, which, _dptr(out)))
        return out

    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _check(lib().ipg_set_field(self._c, FIELDS[name], _dptr(a)))

    def tmunu(self):
        out = np.zeros((10, self.S))
        _check(lib().ipg_get_tmunu(self._c, _dptr(out)))
        return out

    def adjoint_dedeta(self):
        """dE/deta and its gradient with respect to the tau = 0 fields
        {name: [S, 18]} (consumes the state; call after init)"""
        g = np.zeros((6, self.S, 18))
        v = C.c_double()
        _check(lib().ipg_adjoint_dedeta(self._c, C.byref(v), _dptr(g)))
        return v.value, {k: g[i] for i, k in enumerate(("Ux", "Uy", "E1", "E2", "phi", "pi"))}

    LOSSES = {"dEdeta": 0, "eps2": 1, "tmunu": 2}

    def set_checkpoint_interval(self, K):
        """steps between the reverse mode's checkpoints (0: chosen by the
        free device memory, the full tape when it fits; 1: every state
        stored, nothing recomputed; K: every K-th state)"""
        _check(lib().ipg_set_checkpoint_interval(self._c, int(K)))

    def keep_slices(self, on=True):
        """keep the filtered colour fields of the next Wilson lines for the
        next reverse mode to rho (which then skips recomputing them)"""
        _check(lib().ipg_set_keep_slices(self._c, int(bool(on))))

    def checkpoint_interval(self):
        """the interval the last reverse pass used"""
        k = C.c_int(0)
        _check(lib().ipg_get_checkpoint_interval(self._c, C.byref(k)))
        return k.value

    def adjoint(self, loss="dEdeta", weights=None, wrt="fields", fetch=True):
        """Reverse mode with a choice of loss at tau_max: "dEdeta", "eps2"
        (eps_2 of the Landau-frame weights) or "tmunu" (sum over sites of
        weights[k] T_k, ten weights in T's order).  wrt "fields": gradient
        {name: [S, 18]} with respect to the tau = 0 fields; wrt "rho":
        [2, Ny, S, 8] (fetch=False leaves it on the device); wrt "tmunu":
        [10, S], the gradient with respect to T^{mu nu} at itmax.  Consumes
        the state; call after init."""
        v = C.c_double()
        w = None
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.float64)
            if w.shape != (10,):
                raise ValueError("weights: ten entries in T's order")
        if wrt == "fields":
            g = np.zeros((6, self.S, 18))
            _check(lib().ipg_adjoint_loss(self._c, self.LOSSES[loss], None if w is None else _dptr(w), 0,
                                          C.byref(v), _dptr(g)))
            return v.value, {k: g[i] for i, k in enumerate(("Ux", "Uy", "E1", "E2", "phi", "pi"))}
        if wrt == "rho":
            _check(lib().ipg_adjoint_loss(self._c, self.LOSSES[loss], None if w is None else _dptr(w), 1,
                                          C.byref(v), None))
            return v.value, (self.rho_grad() if fetch else None)
        if wrt == "tmunu":
            g = np.zeros((10, self.S))
            _check(lib().ipg_adjoint_loss(self._c, self.LOSSES[loss], None if w is None else _dptr(w), 2,
                                          C.byref(v), _dptr(g)))
            return v.value, g
        raise ValueError("wrt: 'fields', 'rho' or 'tmunu'")

    def adjoint_rho(self, fetch=True):
        """dE/deta and d(dE/deta)/d rho as [2, Ny, S, 8] (consumes the state;
        call after init).  fetch=False leaves the gradient on the device and
        returns None in its place (rho_grad() copies it later)."""
        v = C.c_double()
        _check(lib().ipg_adjoint_rho(self._c, C.byref(v), None))
        return v.value, (self.rho_grad() if fetch else None)

    def rho_grad(self):
        """the last adjoint_rho gradient, [2, Ny, S, 8]"""
        g = np.empty((2, self.p.Ny, self.S, 8))
        _check(lib().ipg_get_rho_grad(self._c, _dptr(g)))
        return g

    def landau(self):
        """[15, S]: eps, u^tau, u^x, u^y, u^eta, pi^{mu nu} (T's order)"""
        out = np.zeros((15, self.S))
        _check(lib().ipg_get_landau(self._c, _dptr(out)))
        return out

    def landau_from(self, T, it):
        """Landau frame and eccentricities of a given T[10, S] at step it"""
        a = None if T is None else np.ascontiguousarray(T, dtype=np.float64)
        _check(lib().ipg_landau(self._c, None if a is None else _dptr(a), it))

    def nucleons(self):
        a1, a2 = C.c_int(), C.c_int()
        _check(lib().ipg_get_nucleons(self._c, None, C.byref(a1), C.byref(a2)))
        out = np.zeros((a1.value + a2.value, 6))
        _check(lib().ipg_get_nucleons(self._c, _dptr(out), C.byref(a1), C.byref(a2)))
        return out[:a1.value], out[a1.value:]

    def set_nucleons(self, A, B):
        arr = np.ascontiguousarray(np.vstack([A, B]), dtype=np.float64)
        _check(lib().ipg_set_nucleons(self._c, _dptr(arr), len(A), len(B)))

    def newton_iterations(self):
        out = np.zeros(2 * self.S, dtype=np.int32)
        _check(lib().ipg_get_newton_iterations(self._c, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def set_rho_normals(self, z):
        if z is None:
            _check(lib().ipg_set_rho_normals(self._c, None))
            return
        a = np.ascontiguousarray(z, dtype=np.float64)
        assert a.size == 2 * self.p.Ny * self.S * 8
        self._z = a
        _check(lib().ipg_set_rho_normals(self._c, _dptr(a)))

    def rho_normals(self, nucleus, k):
        out = np.zeros((self.S, 8))
        _check(lib().ipg_get_rho_normals(self._c, nucleus, k, _dptr(out)))
        return out

    def adjoint_phases(self):
        """wall time of the last reverse-mode call's phases, ms"""
        out = np.zeros(5)
        _check(lib().ipg_adjoint_phases(self._c, _dptr(out)))
        return dict(zip(ADJOINT_PHASES, out.tolist()))

    def enable_timing(self, on=True):
        _check(lib().ipg_enable_timing(self._c, int(on)))

    def timing(self):
        out = np.zeros(3)
        _check(lib().ipg_get_timing(self._c, _dptr(out)))
        return {"step_ms_total": out[0], "tmunu_ms": out[1], "nsteps": int(out[2])}


ADJOINT_PHASES = ("workspace", "forward", "reverse_evolution", "reverse_initial_conditions", "batch_init")


class Batch:
    """Many events in one launch sequence (heterogeneous sizes allowed)."""

    def __init__(self, params_list, seeds, qs_table=None):
        n = len(params_list)
        arr = (Parameters * n)(*params_list)
        sd = (C.c_uint64 * n)(*seeds)
        self.n = n
        self._b = lib().ipg_batch_create(arr, n, sd)
        if not self._b:
            raise IPGError(lib().ipg_last_error().decode())
        _check(lib().ipg_batch_load_qs_table(self._b, (qs_table or QS_TABLE).encode()))

    def run(self, use_graph=True, max_retries=20):
        _check(lib().ipg_batch_run(self._b, int(use_graph), max_retries))

    def observables(self, e):
        o = Observables()
        _check(lib().ipg_batch_get_observables(self._b, e, C.byref(o)))
        return o.as_dict()

    def adjoint(self, loss="dEdeta", weights=None, max_retries=20):
        """Init::init for all events, then each event's loss at tau_max and
        its gradient with respect to rho, kept on the device per event
        (rho_grad(e) copies one out).  Returns the losses [nev]."""
        v = np.zeros(self.n)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        _check(lib().ipg_batch_adjoint(self._b, Event.LOSSES[loss], None if w is None else _dptr(w), max_retries,
                                       _dptr(v)))
        return v

    def rho_grad(self, e, Ny, S):
        g = np.empty((2, Ny, S, 8))
        _check(lib().ipg_batch_get_rho_grad(self._b, e, _dptr(g)))
        return g

    def set_checkpoint_interval(self, K):
        _check(lib().ipg_batch_set_checkpoint_interval(self._b, int(K)))

    def adjoint_info(self):
        """(largest device memory in use during the last adjoint, bytes;
        the checkpoint interval it used)"""
        peak = C.c_ulonglong(0)
        k = C.c_int(0)
        _check(lib().ipg_batch_adjoint_info(self._b, C.byref(peak), C.byref(k)))
        return peak.value, k.value

    def adjoint_phases(self):
        """wall time of the last adjoint's phases summed over the events, ms"""
        out = np.zeros(5)
        _check(lib().ipg_batch_adjoint_phases(self._b, _dptr(out)))
        return dict(zip(ADJOINT_PHASES, out.tolist()))

    def close(self):
        if self._b:
            lib().ipg_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def wilson_line_file_names(seed, event_id=0, mpi_size=1, fmt=2):
    """The reference's names of the two Wilson-line files (Init.cpp:1863-1877)."""
    ext = ".txt" if fmt == 1 else ""
    return (f"V-{event_id + 2 * seed * mpi_size}{ext}", f"V-{event_id + (1 + 2 * seed) * mpi_size}{ext}")


def write_wilson_line_file(path, V, L, rapidity=0.0, fmt=2):
    """One nucleus's Wilson lines V [N*N, 18] to a file (format 1 text, 2 binary)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    N = int(round(np.sqrt(V.shape[0])))
    if V.shape != (N * N, 18):
        raise IPGError(f"V must be [N*N, 18], got {V.shape}")
    _check(lib().ipg_write_wilson_line_file(os.fsencode(path), N, L, rapidity, _dptr(V), fmt))


def read_wilson_line_file(path, N, L, b=0.0, nucleus=0, fmt=2, V=None):
    """Init::readV for one nucleus: V [N*N, 18] (the identity where the file,
    shifted by -+b/2 along x, does not reach, unless V is given)."""
    if V is None:
        V = np.zeros((N * N, 18))
        V[:, [0, 8, 16]] = 1.0
    V = np.ascontiguousarray(V, dtype=np.float64)
    _check(lib().ipg_read_wilson_line_file(os.fsencode(path), N, L, b, nucleus, fmt, _dptr(V)))
    return V


def eccentricity_extents(eps, utau, L, avx, avy, cutoff=0.0):
    """maxX, maxY of Evolution::eccentricity (Evolution.cpp:1492-1640) for
    eps, u^tau [N*N] and the centroid (avx, avy)."""
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    utau = np.ascontiguousarray(utau, dtype=np.float64)
    N = int(round(np.sqrt(eps.size)))
    mx, my = C.c_double(), C.c_double()
    _check(lib().ipg_eccentricity_extents(_dptr(eps), _dptr(utau), N, L, avx, avy, cutoff, C.byref(mx), C.byref(my)))
    return mx.value, my.value


def write_hydro_file(path, land, L, it, dtau, size_out, eta_size_out, L_out, deta_out):
    """epsilon-u-Hydro file from a host Landau frame land [15, N*N]; returns Etot."""
    land = np.ascontiguousarray(land, dtype=np.float64)
    N = int(round(np.sqrt(land.shape[1])))
    e = C.c_double()
    _check(lib().ipg_write_hydro_file(os.fsencode(path), _dptr(land), N, L, it, dtau, size_out, eta_size_out, L_out,
                                      deta_out, C.byref(e)))
    return e.value


def write_tmunu_file(path, T, L, it, dtau, size_out, eta_size_out, L_out, deta_out):
    """Tmunu-t file from a host T [10, N*N]."""
    T = np.ascontiguousarray(T, dtype=np.float64)
    N = int(round(np.sqrt(T.shape[1])))
    _check(lib().ipg_write_tmunu_file(os.fsencode(path), _dptr(T), N, L, it, dtau, size_out, eta_size_out, L_out,
                                      deta_out))


def write_jazma_file(path, g2mu2A, g2mu2B, L, it, dtau, size_out, eta_size_out, L_out, deta_out, Etot):
    """Jazma-Hydro file from host g2mu2A, g2mu2B [N*N]."""
    A = np.ascontiguousarray(g2mu2A, dtype=np.float64)
    B = np.ascontiguousarray(g2mu2B, dtype=np.float64)
    N = int(round(np.sqrt(A.size)))
    _check(lib().ipg_write_jazma_file(os.fsencode(path), _dptr(A), _dptr(B), N, L, it, dtau, size_out, eta_size_out,
                                      L_out, deta_out, Etot))


def hydro_file_name(tau, event_id=0, final=False):
    """The reference's name of the hydro file (MyEigen.cpp:379-386)."""
    return f"epsilon-u-Hydro-TauHydro-{event_id}.dat" if final else f"epsilon-u-Hydro-t{tau:g}-{event_id}.dat"


def autotune_report():
    """The autotuner's table: {key: {"times_ms": [...], "chosen": k or None}}"""
    n = lib().ipg_autotune_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().ipg_autotune_report(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        key, rest = line.split(":", 1)
        parts = rest.split()
        times = [float(p.split("=")[1].rstrip("ms")) for p in parts if p[0].isdigit()]
        chosen = [int(p.split("=")[1]) for p in parts if p.startswith("chosen=")]
        out[key] = {"times_ms": times, "chosen": chosen[0] if chosen else None}
    return out

