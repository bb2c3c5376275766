"""Build and call the C restatement of spinsim's integration path.

TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may use it.  See ``oracle/oracle_core.c`` for what is restated and where.

One shared library is compiled per (field source, spin, method, exponential,
frame, squares) with ``gcc -std=c11 -O3 -fopenmp`` into a cache directory
outside the source tree (``$SPINSIM_ORACLE_BUILD``, default
``~/.cache/spinsim_amd/oracle``), so the libraries a process maps from the
tree are the product's own; ``-march=native`` only when built on the machine
that runs it (the cache key includes the host CPU model).
"""
from __future__ import annotations

import ctypes
import hashlib
import math
import os
import platform
import subprocess
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.environ.get("SPINSIM_ORACLE_BUILD") or os.path.join(
    os.path.expanduser("~"), ".cache", "spinsim_amd", "oracle")
_LOCK = threading.Lock()
_CACHE: dict = {}

METHOD_INDEX = {"magnus_cf4": 0, "euler": 1, "heun": 2}
EXPO_INDEX = {"analytic": 0, "lie_trotter": 1, "lie_trotter_8": 2}

PRELUDE = r"""
#include <math.h>
#include <stdint.h>
#define __device__
#define __host__
#define __forceinline__ inline
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define SS_PI 3.14159265358979323846
#define SS_TAU 6.28318530717958647692
typedef struct {
    const double *time;
    const double *amplitude;
    int64_t size;
    int64_t amplitude_stride;
    int order;
    int zero_boundary;
    int axes;
    int reserved;
    int64_t axis_stride;
} ss_table;
typedef struct {
    const ss_table *tables;
    int sim;
} ss_ctx;
double ss_interp_sample(const ss_ctx *ctx, int k, double t);
void ss_interp_sample_axes(const ss_ctx *ctx, int k, double t, double *out, int n);
double ss_interp_sample_axis(const ss_ctx *ctx, int k, int a, double t);
#define SS_CONTEXT const ss_ctx *ss_ctx_ref
#define SS_SAMPLE(k, t) ss_interp_sample(ss_ctx_ref, (k), (t))
#define SS_SAMPLE_AXES(k, t, out, n) ss_interp_sample_axes(ss_ctx_ref, (k), (t), (out), (n))
#define SS_SAMPLE_AXIS(k, a, t) ss_interp_sample_axis(ss_ctx_ref, (k), (a), (t))
"""


class SsTable(ctypes.Structure):
    _fields_ = [("time", ctypes.c_void_p), ("amplitude", ctypes.c_void_p),
                ("size", ctypes.c_int64), ("amplitude_stride", ctypes.c_int64),
                ("order", ctypes.c_int), ("zero_boundary", ctypes.c_int),
                ("axes", ctypes.c_int), ("reserved", ctypes.c_int),
                ("axis_stride", ctypes.c_int64)]


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.machine()


def _hyper(number_of_squares):
    h = math.ceil(number_of_squares / 2)
    return max(h, 0)


class COracle:
    """A compiled oracle for one field callback and one method set."""

    def __init__(self, field_src, dim, method="magnus_cf4", expo=None, rotating=True,
                 number_of_squares=24, sampler_names=(), native=True):
        if expo is None:
            expo = "lie_trotter" if dim == 3 else "analytic"
        if expo == "analytic" and dim == 3:
            expo = "lie_trotter"
        self.dim = dim
        self.sampler_names = tuple(sampler_names)
        defines = {
            "SS_DIM": dim, "SS_LIE": 3 if dim == 2 else (8 if expo == "lie_trotter_8" else 4),
            "SS_METHOD": METHOD_INDEX[method], "SS_EXPO": EXPO_INDEX[expo],
            "SS_ROT": int(bool(rotating)), "SS_HYPER": _hyper(number_of_squares),
            "SS_CTX": int(bool(self.sampler_names)),
        }
        head = "".join(f"#define {k} {v}\n" for k, v in defines.items())
        names = "".join(f"#define {n}(t) SS_SAMPLE({k}, (t))\n"
                        for k, n in enumerate(self.sampler_names))
        with open(os.path.join(_HERE, "oracle_core.c")) as fh:
            core = fh.read()
        src = head + PRELUDE + names + "\n" + field_src + "\n" + core
        flags = ["-std=c11", "-O3", "-fopenmp", "-fPIC", "-shared"]
        if native:
            flags.append("-march=native")
        key = hashlib.sha256((src + " ".join(flags) + _cpu_model()).encode()).hexdigest()[:20]
        self.lib = _load(src, flags, key)
        f64p, i64 = ctypes.c_void_p, ctypes.c_int64
        self.lib.oracle_time_evolution.argtypes = [
            f64p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
            i64, i64, f64p, f64p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        self.lib.oracle_get_state.argtypes = [f64p, f64p, f64p, i64]
        self.lib.oracle_get_spin.argtypes = [f64p, f64p, i64, ctypes.c_int]
        self._tables = None
        self._keep = []

    def set_samplers(self, samplers):
        """samplers: list of (time, amplitude, order, zero_boundary[, multi_axis]).
        amplitude (M,): one component; (B, M): one row per simulation; with
        multi_axis=True, (A, M): A components, or (B, A, M) per simulation."""
        self._tables, self._keep = tables_of(samplers)

    def interp_eval(self, k, t, sim=0, axes=1):
        """Sampler k (row `sim`) at times t: (n, axes) values (n values when axes = 1)."""
        self.lib.oracle_interp_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_int64]
        t = np.ascontiguousarray(np.atleast_1d(t), np.float64)
        out = np.zeros((t.size, axes), np.float64)
        self.lib.oracle_interp_eval(ctypes.addressof(self._tables), int(k), int(sim), int(axes),
                                    t.ctypes.data, out.ctypes.data, t.size)
        return out[:, 0] if axes == 1 else out

    def time_evolution(self, sweep, t0, dt, dt_out, n_fine, n_coarse, i_begin=0, i_end=None,
                       sim=0, threads=None, out=None):
        i_end = n_coarse if i_end is None else i_end
        threads = threads or len(os.sched_getaffinity(0))
        sweep = np.ascontiguousarray(sweep, np.float64)
        if out is None:
            time = np.zeros(n_coarse, np.float64)
            u = np.zeros((n_coarse, self.dim, self.dim), np.complex128)
        else:
            time, u = out
        tables = ctypes.addressof(self._tables) if self._tables is not None else None
        self.lib.oracle_time_evolution(sweep.ctypes.data, t0, dt, dt_out, int(n_fine),
                                       int(i_begin), int(i_end), time.ctypes.data,
                                       u.ctypes.data, tables, int(sim), int(threads))
        return time, u

    def get_state(self, psi0, u):
        u = np.ascontiguousarray(u, np.complex128)
        psi0 = np.ascontiguousarray(psi0, np.complex128)
        state = np.zeros((u.shape[0], self.dim), np.complex128)
        self.lib.oracle_get_state(psi0.ctypes.data, u.ctypes.data, state.ctypes.data,
                                  u.shape[0])
        return state

    def get_spin(self, state, threads=None):
        threads = threads or len(os.sched_getaffinity(0))
        state = np.ascontiguousarray(state, np.complex128)
        spin = np.zeros((state.shape[0], 3), np.float64)
        self.lib.oracle_get_spin(state.ctypes.data, spin.ctypes.data, state.shape[0],
                                 int(threads))
        return spin


def tables_of(samplers):
    """(ctypes ss_table array, the arrays its records point at) for samplers
    given as (time, amplitude, order, zero_boundary[, multi_axis])."""
    arr = (SsTable * max(len(samplers), 1))()
    keep = []
    for k, entry in enumerate(samplers):
        time, amp, order, zero_boundary = entry[:4]
        multi = bool(entry[4]) if len(entry) > 4 else False
        t = np.ascontiguousarray(time, np.float64)
        a = np.ascontiguousarray(amp, np.float64)
        keep += [t, a]
        m = t.size
        if multi:
            axes = a.shape[-2]
            rows_stride = axes * m if a.ndim == 3 else 0
        else:
            axes = 1
            rows_stride = m if a.ndim == 2 else 0
        arr[k].time = t.ctypes.data
        arr[k].amplitude = a.ctypes.data
        arr[k].size = m
        arr[k].amplitude_stride = rows_stride
        arr[k].order = int(order)
        arr[k].zero_boundary = int(bool(zero_boundary))
        arr[k].axes = axes
        arr[k].reserved = 0
        arr[k].axis_stride = m
    return arr, keep


def _load(src, flags, key):
    with _LOCK:
        if key in _CACHE:
            return _CACHE[key]
        os.makedirs(_BUILD, exist_ok=True)
        so = os.path.join(_BUILD, f"oracle_{key}.so")
        if not os.path.exists(so):
            c_path = os.path.join(_BUILD, f"oracle_{key}.c")
            with open(c_path, "w") as fh:
                fh.write(src)
            tmp = so + f".tmp{os.getpid()}"
            cmd = ["gcc", *flags, "-o", tmp, c_path, "-lm"]
            res = subprocess.run(cmd, capture_output=True, text=True)
            if res.returncode != 0:
                raise RuntimeError("oracle build failed:\n" + res.stderr)
            os.replace(tmp, so)
        lib = ctypes.CDLL(so)
        _CACHE[key] = lib
        return lib
