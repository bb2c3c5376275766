"""The C ABI bound with raw ctypes and the HIP runtime (no torch), the way the
reference package would bind it, checked against the oracle."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
vp, i64, c_int, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double


def _libs():
    hip = ctypes.CDLL("libamdhip64.so")
    ss = ctypes.CDLL(os.path.join(ROOT, "spinsim_amd", "libspinsim_hip.so"))
    ss.ss_last_error.restype = ctypes.c_char_p
    ss.ss_compile.argtypes = [ctypes.c_char_p] + [c_int] * 7 + [ctypes.POINTER(vp)]
    ss.ss_module_free.argtypes = [vp]
    ss.ss_time_evolution.argtypes = [vp, vp, dbl, dbl, dbl, i64, vp, vp, vp, c_int, vp]
    ss.ss_state_scan_workspace.restype = ctypes.c_size_t
    ss.ss_state_scan_workspace.argtypes = [i64, c_int]
    ss.ss_state_scan.argtypes = [vp, vp, vp, vp, i64, c_int, vp, ctypes.c_size_t, vp]
    ss.ss_spin.argtypes = [vp, vp, i64, c_int, vp]
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, c_int]
    return hip, ss


class _Dev:
    def __init__(self, hip, nbytes):
        self.hip, self.ptr = hip, vp()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(nbytes, 1)) == 0

    def put(self, a):
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, 2) == 0
        return a

    def free(self):
        self.hip.hipFree(self.ptr)


def test_raw_ctypes_binding_matches_oracle():
    print("case: raw ctypes binding", flush=True)
    hip, ss = _libs()
    field = F.RABI_SWEEP_ONE
    mod = vp()
    rc = ss.ss_compile(field.src.encode(), 3, 0, 1, 1, 24, 0, 0, ctypes.byref(mod))
    assert rc == 0, ss.ss_last_error()
    dt, n, nf, _ = O.timing(0.0, 3e-4, 1e-7, 1e-6)
    sweep = np.asarray(field.sweep, np.float64)
    time = np.empty(n)
    u = np.empty((n, 3, 3), np.complex128)
    state = np.empty((n, 3), np.complex128)
    spin = np.empty((n, 3))
    bufs = [_Dev(hip, x.nbytes) for x in (sweep, time, u, state, spin)]
    d_sw, d_t, d_u, d_s, d_o = bufs
    d_sw.put(sweep)
    assert ss.ss_time_evolution(mod, d_sw.ptr, 0.0, dt, 1e-6, n, d_t.ptr, d_u.ptr, None, 0,
                                None) == 0, ss.ss_last_error()
    ws = ss.ss_state_scan_workspace(n, 3)
    d_ws = _Dev(hip, ws)
    psi0 = np.ascontiguousarray(O.EIGENSTATES["one"]["plus_x"])
    assert ss.ss_state_scan(psi0.ctypes.data, d_u.ptr, d_s.ptr, None, n, 3, d_ws.ptr, ws,
                            None) == 0
    assert ss.ss_spin(d_s.ptr, d_o.ptr, n, 3, None) == 0
    d_t.get(time), d_u.get(u), d_s.get(state), d_o.get(spin)
    for b in bufs + [d_ws]:
        b.free()
    ss.ss_module_free(mod)
    co = COracle(field.src, 3)
    t_ref, u_ref = co.time_evolution(sweep, 0.0, dt, 1e-6, nf, n)
    s_ref = co.get_state(psi0, u_ref)
    assert np.array_equal(time, t_ref)
    assert np.abs(u - u_ref).max() <= 1e-9
    assert np.abs(state - s_ref).max() <= 1e-9
    assert np.abs(spin - co.get_spin(s_ref)).max() <= 1e-9
