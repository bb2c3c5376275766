"""CPU devices (Device.CPU, CPU_SINGLE, PYTHON): a host build of the same
integrator text the GPU kernels are compiled from.

The reference JIT-compiles its integrator with numba for the CPU as well as
for CUDA (spinsim/__init__.py:144-211) and runs the coarse steps in parallel
with ``nb.prange`` (:883-885), resized per call from ``number_of_threads``
(:1035-1047).  Here the device headers (``csrc/device``) are compiled with
g++ and OpenMP, specialised by the same macros as the hipRTC build, once per
field callback (cached on disk like the GPU code objects):

* ``libspinsim_host.so`` (built by ``spinsim_amd._build``): the state
  product (sequential, like :1103-1111), the Bloch vectors and the
  interpolation sampler;
* a host module per callback: ``ss_host_time_evolution``, one coarse step per
  loop iteration, the loop split over the OpenMP threads.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
import threading

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
HOST_LIB_PATH = os.path.join(PKG, "libspinsim_host.so")

# Sources a host module is built from (their text is part of the cache key).
MODULE_HEADERS = ("host/ss_host_prelude.h", "device/ss_args.h", "device/ss_complex.h",
                  "device/ss_prelude.h", "device/ss_integrator.h", "host/ss_offsets.h",
                  "host/ss_host_module.h")
# No fast-math and no contraction beyond the explicit fma() calls of the
# device text: g++ would otherwise fuse a*b + c by default in C++, e.g. the
# time stamps t0 + dt_out*i, which the reference rounds twice (:818).
MODULE_FLAGS = ["-O3", "-std=c++17", "-march=native", "-ffp-contract=off", "-fopenmp", "-shared",
                "-fPIC", "-Wno-unknown-pragmas", "-Wno-unused-function"]
HOST_LIB_FLAGS = ["-O3", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-shared", "-fPIC",
                  "-Wno-unknown-pragmas", "-Wno-unused-function"]
HOST_ABI = 3
MODULE_ABI = 1

_lock = threading.Lock()
_host_lib = None
_modules: dict = {}

c_vp = ctypes.c_void_p
c_int = ctypes.c_int
c_ll = ctypes.c_longlong
c_dbl = ctypes.c_double


class HostError(RuntimeError):
    """A failure of the host (CPU device) build or kernels."""


def build_host_library(verbose: bool = False) -> str:
    """Compile libspinsim_host.so in-tree; returns its path."""
    tmp = HOST_LIB_PATH + f".tmp{os.getpid()}"
    cmd = ["g++", *HOST_LIB_FLAGS, "-I", CSRC, os.path.join(CSRC, "host", "ss_host.cpp"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise HostError("g++ failed building the host library:\n" + res.stdout + res.stderr)
    os.replace(tmp, HOST_LIB_PATH)
    return HOST_LIB_PATH


def host_lib():
    """The loaded callback-independent host library; raises if it is not built."""
    global _host_lib
    if _host_lib is not None:
        return _host_lib
    with _lock:
        if _host_lib is None:
            if not os.path.exists(HOST_LIB_PATH):
                raise HostError(f"{HOST_LIB_PATH} is missing: build it with "
                                "`python -m spinsim_amd._build`")
            lib = ctypes.CDLL(HOST_LIB_PATH)
            lib.ss_host_abi_version.restype = c_int
            if lib.ss_host_abi_version() != HOST_ABI:
                raise HostError(f"{HOST_LIB_PATH} has ABI {lib.ss_host_abi_version()}, "
                                f"this package needs {HOST_ABI}: rebuild it")
            lib.ss_host_state.argtypes = [c_vp, c_vp, c_vp, c_vp, c_ll, c_int]
            lib.ss_host_state.restype = c_int
            lib.ss_host_spin.argtypes = [c_vp, c_vp, c_ll, c_int, c_int]
            lib.ss_host_spin.restype = c_int
            lib.ss_host_interp_eval.argtypes = [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_ll]
            lib.ss_host_interp_eval.restype = c_int
            lib.ss_host_f32_eval.argtypes = [c_int, c_vp, c_vp, c_ll]
            lib.ss_host_f32_eval.restype = c_int
            lib.ss_host_f64_eval.argtypes = [c_int, c_vp, c_vp, c_ll]
            lib.ss_host_f64_eval.restype = c_int
            _host_lib = lib
    return _host_lib


def _check(rc, what):
    if rc != 0:
        raise HostError(f"{what} rejected its arguments")


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _compiler_version():
    res = subprocess.run(["g++", "--version"], capture_output=True, text=True)
    return res.stdout.splitlines()[0] if res.returncode == 0 and res.stdout else "g++"


def module_source(field_src, dim, method, expo, rotating, number_of_squares, has_context,
                  precision):
    """C++ text of one host module: the specialisation macros (as the hipRTC
    build sets them), the host prelude, the device headers, the callback and
    the host entry points."""
    hyper = (number_of_squares + 1) // 2 if number_of_squares > 0 else 0  # :1547
    lie = 3 if dim == 2 else (8 if expo == 2 else 4)
    head = (f"#define SS_DIM {dim}\n#define SS_LIE {lie}\n#define SS_METHOD {method}\n"
            f"#define SS_EXPO {expo}\n#define SS_ROT {1 if rotating else 0}\n"
            f"#define SS_HYPER {hyper}\n#define SS_CTX {1 if has_context else 0}\n"
            f"#define SS_PREC {precision}\n")
    # the callback's sin and cos: the table-driven float64 functions, or with
    # precision fp32 float32 ones with exp as well (ss_prelude.h), as on the GPU
    if precision == 3:
        on = "#define sin(x) ss_sin32(x)\n#define cos(x) ss_cos32(x)\n#define exp(x) ss_exp32(x)\n"
        off = "#undef sin\n#undef cos\n#undef exp\n"
    else:
        on = "#define sin(x) ss_sin64(x)\n#define cos(x) ss_cos64(x)\n"
        off = "#undef sin\n#undef cos\n"
    return (head + '#include "host/ss_host_prelude.h"\n#include "device/ss_args.h"\n'
            '#include "device/ss_complex.h"\n#include "device/ss_prelude.h"\n'
            + on + '#line 1 "get_field"\n' + field_src + "\n" + off +
            '#include "device/ss_integrator.h"\n#include "host/ss_host_module.h"\n')


def _cache_dir():
    d = os.environ.get("SPINSIM_AMD_CACHE")
    if d is None:
        d = os.path.join(os.path.expanduser("~"), ".cache", "spinsim_amd")
    return os.path.join(d, "host")


class HostModule:
    """One compiled host module (one callback, one method set)."""

    def __init__(self, path):
        self.path = path
        lib = ctypes.CDLL(path)
        lib.ss_host_module_abi.restype = c_int
        if lib.ss_host_module_abi() != MODULE_ABI:
            raise HostError(f"{path}: host module ABI {lib.ss_host_module_abi()} != {MODULE_ABI}")
        lib.ss_host_time_evolution.argtypes = [c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_ll, c_vp, c_vp,
                                               c_int]
        lib.ss_host_time_evolution.restype = c_int
        if hasattr(lib, "ss_host_residual_squares"):
            lib.ss_host_residual_squares.argtypes = [c_dbl]
            lib.ss_host_residual_squares.restype = c_int
        self.lib = lib

    def time_evolution(self, sweep, tables, t0, dt, dt_out, time_out, u_out, threads):
        """Fill time_out (N,) float64 and u_out (N, d, d) complex128 in place."""
        sweep = np.ascontiguousarray(np.atleast_1d(np.asarray(sweep, np.float64)))
        n = int(time_out.shape[0])
        if not (time_out.flags.c_contiguous and u_out.flags.c_contiguous
                and time_out.dtype == np.float64 and u_out.dtype == np.complex128
                and u_out.shape[0] == n):
            raise ValueError("time_out must be (N,) float64 and u_out (N, d, d) complex128, "
                             "C-contiguous")
        _check(self.lib.ss_host_time_evolution(
            sweep.ctypes.data, tables, float(t0), float(dt), float(dt_out), n,
            time_out.ctypes.data, u_out.ctypes.data, int(threads or 0)),
            "ss_host_time_evolution")


def compile_module(field_src, dim, method, expo, rotating, number_of_squares, has_context,
                   precision) -> HostModule:
    """Compile (or load from the disk cache) the host module of one callback."""
    src = module_source(field_src, dim, method, expo, rotating, number_of_squares, has_context,
                        precision)
    h = hashlib.sha256()
    h.update(src.encode())
    for name in MODULE_HEADERS:
        with open(os.path.join(CSRC, name), "rb") as fh:
            h.update(fh.read())
    h.update("\x00".join(MODULE_FLAGS + [_compiler_version(), _cpu_model()]).encode())
    key = h.hexdigest()[:32]
    with _lock:
        if key in _modules:
            return _modules[key]
    d = _cache_dir()
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, f"{key}.so")
    if not os.path.exists(path):
        cpp = os.path.join(d, f"{key}.{os.getpid()}.cpp")
        tmp = f"{path}.tmp{os.getpid()}"
        with open(cpp, "w") as fh:
            fh.write(src)
        try:
            res = subprocess.run(["g++", *MODULE_FLAGS, "-I", CSRC, cpp, "-o", tmp],
                                 capture_output=True, text=True)
            if res.returncode != 0:
                raise HostError("g++ could not compile get_field with the integrator:\n"
                                + res.stdout + res.stderr)
            os.replace(tmp, path)
        finally:
            for p in (cpp, tmp):
                if os.path.exists(p):
                    os.remove(p)
    mod = HostModule(path)
    with _lock:
        _modules[key] = mod
    return mod


def f32_eval(fn, x):
    """The field callbacks' float32 sin / cos / exp (precision "fp32"), host
    build: fn in ("sin", "cos", "exp")."""
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, np.float64)))
    out = np.empty_like(x)
    _check(host_lib().ss_host_f32_eval(("sin", "cos", "exp").index(fn), x.ctypes.data,
                                       out.ctypes.data, x.size), "ss_host_f32_eval")
    return out


def f64_eval(fn, x):
    """The field callbacks' table-driven float64 sin / cos (every precision
    but "fp32"), host build: fn in ("sin", "cos")."""
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, np.float64)))
    out = np.empty_like(x)
    _check(host_lib().ss_host_f64_eval(("sin", "cos").index(fn), x.ctypes.data,
                                       out.ctypes.data, x.size), "ss_host_f64_eval")
    return out


def state(psi0, u, with_spin=False):
    """get_state on the host: (state (N, d), spin (N, 3) or None)."""
    u = np.ascontiguousarray(u, np.complex128)
    n, d = int(u.shape[0]), int(u.shape[1])
    psi = np.ascontiguousarray(np.asarray(psi0, np.complex128).reshape(d))
    out = np.empty((n, d), np.complex128)
    spin = np.empty((n, 3), np.float64) if with_spin else None
    _check(host_lib().ss_host_state(psi.ctypes.data, u.ctypes.data, out.ctypes.data,
                                    None if spin is None else spin.ctypes.data, n, d),
           "ss_host_state")
    return out, spin


def spin(state_, threads=None):
    """Bloch vectors (N, 3) of a state time series on the host."""
    st = np.ascontiguousarray(state_, np.complex128)
    n, d = int(st.shape[0]), int(st.shape[1])
    out = np.empty((n, 3), np.float64)
    _check(host_lib().ss_host_spin(st.ctypes.data, out.ctypes.data, n, d, int(threads or 0)),
           "ss_host_spin")
    return out
