"""ipg_landau_noalloc (the Landau frame and eccentricities of a device-resident
T^{mu nu} with no device allocation) against ipg_landau, the context entry
point it varies, on the same T^{mu nu}: two calls in a row, the second on a
smaller lattice, through the C ABI.  Device buffers come from the HIP
runtime the library links (tests/_hipmem.py)."""
import ctypes as C

import numpy as np
import pytest

import _cases as K
from _hipmem import DeviceArray

pytestmark = pytest.mark.gpu

OBS_FIELDS = ("area", "Rbar", "avg_eps", "avg_qs2A_qs2B", "rms_r", "landau_failures", "landau_no_timelike")


def _event_T(system, N, Ny, seed):
    import ipglasma_amd as I
    p = K.params(system, N=N, Ny=Ny, seed=seed)
    ev = I.Event(p)
    for _ in range(20):
        if ev.init():
            break
    else:
        raise RuntimeError("no successful event")
    ev.run()
    return p, ev


def _compare(p, ev):
    import ipglasma_amd as I
    L = I.lib()
    it = p.itmax
    T = ev.tmunu()
    g2A, g2B = ev.field("g2mu2A"), ev.field("g2mu2B")
    # the context entry point on the same T^{mu nu}
    ev.landau_from(T, it)
    land_ref = ev.landau()
    o_ref = ev.observables()
    Td = DeviceArray((10, ev.S), T)
    gA = DeviceArray((ev.S,), g2A)
    gB = DeviceArray((ev.S,), g2B)
    land = DeviceArray((15, ev.S), np.full((15, ev.S), np.nan))
    o = I.Observables()
    rc = L.ipg_landau_noalloc(C.byref(p), Td.ptr, gA.ptr, gB.ptr, it, land.ptr, C.byref(o))
    assert rc == 0, L.ipg_last_error().decode()
    got = land.download()
    # same kernels on the same input: the same bits
    assert np.array_equal(got, land_ref), np.nanmax(np.abs(got - land_ref))
    d = o.as_dict()
    for k in ("eccentricity", "Psi", "centroid", "u_rms"):
        assert np.array_equal(np.array(d[k]), np.array(o_ref[k])), (k, d[k], o_ref[k])
    for k in OBS_FIELDS:
        assert d[k] == o_ref[k], (k, d[k], o_ref[k])
    assert abs(d["tau_fm"] - o_ref["tau_fm"]) <= 1e-15 * o_ref["tau_fm"]
    return d


def test_noalloc_landau_matches_context_entry_point_twice():
    first = _compare(*_event_T("AuAu", 48, 3, 5))
    second = _compare(*_event_T("pPb", 32, 2, 3))  # smaller lattice, same scratch array
    assert first["area"] != second["area"]


def test_noalloc_landau_without_colour_charge_and_size_limit():
    import ipglasma_amd as I
    L = I.lib()
    p, ev = _event_T("pPb", 32, 2, 4)
    T = ev.tmunu()
    Td = DeviceArray((10, ev.S), T)
    land = DeviceArray((15, ev.S), np.zeros((15, ev.S)))
    o = I.Observables()
    rc = L.ipg_landau_noalloc(C.byref(p), Td.ptr, None, None, p.itmax, land.ptr, C.byref(o))
    assert rc == 0, L.ipg_last_error().decode()
    assert o.avg_qs2A_qs2B == 0.0
    assert np.array_equal(land.download(), ev.landau())
    # beyond the scratch array's size: refused before any device work
    big = p.copy()
    big.size = 2048
    rc = L.ipg_landau_noalloc(C.byref(big), Td.ptr, None, None, p.itmax, land.ptr, C.byref(o))
    assert rc == -1 and b"2048" in L.ipg_last_error()
