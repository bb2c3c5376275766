"""ctypes binding of libspinsim_hip.so (declared in include/spinsim_hip.h).

The library is built in-tree by ``spinsim_amd._build``.  If it is missing,
every entry point raises: there is no CPU fallback behind the HIP path.
"""
from __future__ import annotations

import ctypes
import os
import threading

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libspinsim_hip.so")

_lock = threading.Lock()
_lib = None

c_i64 = ctypes.c_int64
c_vp = ctypes.c_void_p
c_int = ctypes.c_int
c_dbl = ctypes.c_double

# name -> (restype, argtypes); mirrors include/spinsim_hip.h
SIGNATURES = {
    "ss_last_error": (ctypes.c_char_p, []),
    "ss_abi_version": (c_int, []),
    "ss_source_hash": (ctypes.c_uint64, []),
    "ss_time_evolution_events": (c_int, [c_vp, c_vp]),
    "ss_compile_code": (c_int, [ctypes.c_char_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                ctypes.POINTER(c_vp), ctypes.POINTER(ctypes.c_size_t),
                                ctypes.POINTER(c_vp)]),
    "ss_free": (None, [c_vp]),
    "ss_module_load": (c_int, [c_vp, ctypes.c_size_t, c_int, c_int, c_int, c_int,
                               ctypes.POINTER(c_vp)]),
    "ss_compile": (c_int, [ctypes.c_char_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                           ctypes.POINTER(c_vp)]),
    "ss_module_free": (c_int, [c_vp]),
    "ss_module_info": (c_int, [c_vp, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                               ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "ss_time_evolution": (c_int, [c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_i64, c_vp, c_vp, c_vp,
                                  c_int, c_vp]),
    "ss_batch_time_evolution": (c_int, [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp,
                                        c_i64, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int,
                                        c_int, c_vp]),
    "ss_state_scan_workspace": (ctypes.c_size_t, [c_i64, c_int]),
    "ss_state_scan": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_vp, ctypes.c_size_t,
                              c_vp]),
    "ss_batch_scan_workspace": (ctypes.c_size_t, [c_i64, c_int, c_i64]),
    "ss_time_evolution_plan": (c_int, [c_i64, c_int, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "ss_batch_state_scan": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_int, c_i64,
                                    c_int, c_vp, ctypes.c_size_t, c_vp]),
    "ss_batch_state_scan_static": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_int,
                                           c_i64, c_vp]),
    "ss_batch_costate_scan": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_int, c_i64,
                                      c_vp, ctypes.c_size_t, c_vp]),
    "ss_batch_final_state": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_vp]),
    "ss_batch_adjoint": (c_int, [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64,
                                 c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_i64, c_vp,
                                 c_vp, ctypes.c_size_t, c_vp]),
    "ss_batch_adjoint_workspace": (ctypes.c_size_t, [c_i64, c_int]),
    "ss_interp_eval": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_i64, c_vp]),
    "ss_spin": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp]),
    "ss_fma_probe": (c_int, [c_vp, c_i64, c_int, c_vp]),
}


# Version of the C ABI these signatures describe (ss_abi_version()).
ABI_VERSION = 9

# Operator row layouts of the batch entry points (SS_U_* of the C header).
U_FULL = 0
U_SU2_COLUMN = 1
U_SU2_FULL = 2


class NativeError(RuntimeError):
    """A failure reported by libspinsim_hip.so."""


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise NativeError(
                    f"{LIB_PATH} is missing: build it with `python -m spinsim_amd._build` "
                    "(spinsim_amd has no CPU fallback for the HIP path)")
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            version = handle.ss_abi_version()
            if version != ABI_VERSION:
                raise NativeError(
                    f"{LIB_PATH} has C ABI version {version}, this package needs {ABI_VERSION}: "
                    "rebuild it with `python -m spinsim_amd._build`")
            _lib = handle
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().ss_last_error().decode(errors="replace")
        raise NativeError(f"{what}: {msg}")


def compile_code(field_src: str, dim: int, method: int, expo: int, rotating: bool,
                 number_of_squares: int, field_has_context: bool, precision: int = 0):
    """hipRTC-compile a field callback with the integrator; returns (code, log)."""
    L = lib()
    code = c_vp()
    size = ctypes.c_size_t()
    log = c_vp()
    rc = L.ss_compile_code(field_src.encode(), dim, method, expo, int(rotating),
                           int(number_of_squares), int(field_has_context), int(precision),
                           ctypes.byref(code),
                           ctypes.byref(size), ctypes.byref(log))
    text = ctypes.string_at(log.value).decode(errors="replace") if log.value else ""
    if log.value:
        L.ss_free(log)
    if rc != 0:
        msg = L.ss_last_error().decode(errors="replace")
        raise NativeError(msg)
    blob = ctypes.string_at(code.value, size.value)
    L.ss_free(code)
    return blob, text


class Module:
    """A loaded integrator module (one field callback, one method set)."""

    def __init__(self, code: bytes, dim: int, method: int, expo: int, rotating: bool):
        L = lib()
        self._code = ctypes.create_string_buffer(code, len(code))
        handle = c_vp()
        check(L.ss_module_load(self._code, len(code), dim, method, expo, int(rotating),
                               ctypes.byref(handle)), "ss_module_load")
        self.handle = handle
        self.dim = dim

    def info(self):
        v, s, scratch, t = c_int(), c_int(), c_int(), c_int()
        check(lib().ss_module_info(self.handle, ctypes.byref(v), ctypes.byref(s),
                                   ctypes.byref(scratch), ctypes.byref(t)), "ss_module_info")
        return {"vgprs": v.value, "scratch_bytes": scratch.value, "max_threads": t.value}

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            try:
                _lib.ss_module_free(h)
            except Exception:  # interpreter shutdown
                pass
            self.handle = None
