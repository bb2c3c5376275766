"""Drop-in ``Simulator`` / ``Results`` (spinsim/__init__.py:271-1111) on MI355X.

``Simulator(get_field, spin_quantum_number, ...)`` compiles the field callback
with the integrator for gfx950 (hipRTC) at construction, as the reference
JIT-compiles with numba (:447-451, :453-991).  ``evaluate(...)`` runs three
device steps on the current HIP stream -- the per-coarse-step time-evolution
kernel, the parallel state prefix product and (lazily) the Bloch vectors --
and returns numpy arrays like the reference, or device tensors on request.
"""
from __future__ import annotations

import ctypes
import hashlib
import math
import os
import threading
import warnings

import numpy as np

from . import _native
from .enums import Device, ExponentiationMethod, IntegrationMethod, SpinQuantumNumber
from .interpolation import InterpolationSampler, build_table_array
from .transpile import transpile

_RED = "\033[31m"
_YELLOW = "\033[33m"
_RESET = "\033[0m"

_code_cache: dict = {}
_code_lock = threading.Lock()


def _cache_dir():
    d = os.environ.get("SPINSIM_AMD_CACHE")
    if d is None:
        d = os.path.join(os.path.expanduser("~"), ".cache", "spinsim_amd")
    return d


PRECISIONS = {"fp64": 0, "fp32_bulk": 1, "fp64_series": 2, "fp32": 3}


def compile_field(src: str, dim: int, method: int, expo: int, rotating: bool,
                  number_of_squares: int, has_context: bool, precision: int = 0) -> bytes:
    """hipRTC code object for one field source and method set (memory and disk cached)."""
    L = _native.lib()
    key_text = "\x00".join(map(str, (src, dim, method, expo, int(rotating), number_of_squares,
                                      int(has_context), precision, L.ss_source_hash(),
                                      L.ss_abi_version())))
    key = hashlib.sha256(key_text.encode()).hexdigest()[:32]
    with _code_lock:
        if key in _code_cache:
            return _code_cache[key]
    path = os.path.join(_cache_dir(), f"{key}.co")
    code = None
    if os.path.exists(path):
        with open(path, "rb") as fh:
            code = fh.read()
    if not code:
        code, _log = _native.compile_code(src, dim, method, expo, rotating, number_of_squares,
                                          has_context, precision)
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = f"{path}.tmp{os.getpid()}"
            with open(tmp, "wb") as fh:
                fh.write(code)
            os.replace(tmp, path)
        except OSError:
            pass
    with _code_lock:
        _code_cache[key] = code
    return code


class FieldSource:
    """A field callback as HIP-C text plus the samplers it reads.

    ``get_field`` may be a Python function (translated by
    :func:`spinsim_amd.transpile.transpile`), a string of HIP-C defining
    ``__device__ void get_field(double t, const double* sweep, double* field)``
    (``..., SS_CONTEXT)`` when it reads samplers), or a FieldSource.
    ``samplers`` maps names usable in HIP-C text to samplers (or is a list,
    read with ``SS_SAMPLE(k, t)``); a multi-axis sampler of A components reads
    as ``name(t, out)`` into ``double out[A]`` (or ``SS_SAMPLE_AXES(k, t, out,
    A)`` / ``SS_SAMPLE_AXIS(k, a, t)``).
    """

    def __init__(self, get_field, samplers=None):
        if isinstance(get_field, FieldSource):
            self.source, self.samplers, self.has_context = (
                get_field.source, list(get_field.samplers), get_field.has_context)
            return
        if isinstance(get_field, str):
            names = []
            if isinstance(samplers, dict):
                names = list(samplers.keys())
                self.samplers = list(samplers.values())
            else:
                self.samplers = list(samplers or [])
            for s in self.samplers:
                if not isinstance(s, InterpolationSampler):
                    raise TypeError("samplers must be InterpolationSampler objects")
            # a one-axis sampler reads as name(t); a multi-axis one as name(t, out),
            # filling out[0 .. A) from one index search
            macros = "".join(
                f"#define {n}(t) SS_SAMPLE({k}, (t))\n" if self.samplers[k].axes == 1 else
                f"#define {n}(t, out) SS_SAMPLE_AXES({k}, (t), (out), {self.samplers[k].axes})\n"
                for k, n in enumerate(names))
            self.source = macros + get_field
            self.has_context = bool(self.samplers) or "SS_CONTEXT" in get_field
            return
        if callable(get_field):
            if samplers:
                raise TypeError("samplers= is only used with HIP-C source; Python field "
                                "functions reference their samplers directly")
            self.source, self.samplers, self.has_context = transpile(get_field)
            return
        raise TypeError("get_field must be a Python function, HIP-C source text or a FieldSource")


# Largest evaluate() result (time, U and state) staged through pinned host
# memory; larger results are copied to pageable memory directly.
PINNED_RESULT_BYTES = 1 << 31


class Results:
    """The results of an evaluation (spinsim/__init__.py:271-332).

    ``time`` (N,) float64, ``time_evolution`` (N, d, d) complex128,
    ``state`` (N, d) complex128, and ``spin`` (N, 3) float64 computed on first
    access with ``spin_calculator``.  With ``output="torch"`` these are device
    tensors.
    """

    def __init__(self, time, time_evolution, state, spin_calculator, _device_state=None):
        self.time = time
        self.time_evolution = time_evolution
        self.state = state
        self.spin_calculator = spin_calculator
        self._device_state = _device_state

    def __getattr__(self, attr_name: str):
        if attr_name == "spin":
            src = self.__dict__.get("_device_state")
            if src is not None and isinstance(self.__dict__["state"], np.ndarray):
                spin = self.spin_calculator(src).cpu().numpy()
            else:
                spin = self.spin_calculator(self.__dict__["state"])
            setattr(self, attr_name, spin)
            return spin
        raise AttributeError("{} has no attribute called {}.".format(self, attr_name))


def _default_device():
    """Device.ROCM when a HIP device is visible, else Device.CPU (as the
    reference picks CUDA when available, else CPU, :435-439)."""
    try:
        import torch
        if torch.cuda.is_available():
            return Device.ROCM
    except ImportError:
        pass
    return Device.CPU


HOST_DEVICES = (Device.CPU, Device.CPU_SINGLE, Device.PYTHON)


def resolve_exponentiation(spin_quantum_number, exponentiation_method, warn=True):
    """Defaults and the ANALYTIC -> LIE_TROTTER fallback of :529-559."""
    if not exponentiation_method:
        if spin_quantum_number == SpinQuantumNumber.ONE:
            exponentiation_method = ExponentiationMethod.LIE_TROTTER
        else:
            exponentiation_method = ExponentiationMethod.ANALYTIC
    if (exponentiation_method == ExponentiationMethod.ANALYTIC
            and spin_quantum_number != SpinQuantumNumber.HALF):
        if warn:
            print(f"{_RED}spinsim warning!!!\n_attempting to use an analytic exponentiation "
                  f"method outside of spin-half. Switching to a Lie Trotter method.{_RESET}")
        exponentiation_method = ExponentiationMethod.LIE_TROTTER
    if (exponentiation_method == ExponentiationMethod.LIE_TROTTER_8
            and spin_quantum_number != SpinQuantumNumber.ONE):
        # "Only available for use with spin-one systems", :1429-1432
        raise ValueError("ExponentiationMethod.LIE_TROTTER_8 is only available for spin-one systems")
    return exponentiation_method


def derive_steps(time_start, time_end, time_step_integration, time_step_output, warn=True):
    """evaluate()'s step arithmetic, :1025-1034: (dt, N, time_end_points)."""
    dt_old = time_step_integration
    dt = time_step_output / round(max(time_step_output / time_step_integration, 1))
    if math.fabs(time_step_output / dt_old - round(time_step_output / dt_old)) > 1e-6 and warn:
        print(f"{_YELLOW}spinsim warning!!!\ntime_step_output ({time_step_output:8.4e}) not an "
              f"integer multiple of time_step_integration ({dt_old:8.4e}). Resetting "
              f"time_step_integration to {dt:8.4e}.{_RESET}\n")
    ends = np.asarray([time_start, time_end], np.float64)
    n = int((ends[1] - ends[0]) / time_step_output)
    return float(dt), n, ends


class Simulator:
    """Drop-in for spinsim.Simulator (spinsim/__init__.py:334-1111).

    Extra keyword arguments: ``samplers`` (for HIP-C field text),
    ``lanes_per_step`` (lanes sharing one coarse step: an int in 1..64;
    None or 0 for the default launch, which deals the rows to the resident
    waves in phases of decreasing item length so they finish together; -k
    for the same on k waves per SIMD; ``"auto"`` to time the candidates on
    first use of each shape) and
    ``precision``: ``"fp64"`` (default; the reference's arithmetic),
    ``"fp32_bulk"`` -- the higher-order residual exp(A) - I - A of each step
    operator and the products of residuals within a lane's chunk of steps in
    float32 (a scaled Taylor series instead of the exponentiation method),
    while field samples, the frame, first-order terms, the products across
    lanes, the outputs and the state scan stay float64 -- or ``"fp32"``:
    float32-bulk, and the field function's sin, cos and exp evaluated in
    float32 on arguments reduced in float64 (its other arithmetic, the time
    and the frame stay float64).  Accurate when each step's rotation angle
    is small (rotating frame); check with :func:`spinsim_amd.precision_error`.
    ``"fp64_series"`` runs the float32-bulk series in float64 (for
    measurement).
    ``threads_per_block`` and ``max_registers`` are accepted for
    compatibility and do not change the gfx950 launch.

    Devices: ``Device.ROCM`` (alias ``ROC``) runs the gfx950 HIP kernels; the
    default when a HIP device is visible.  ``Device.CPU`` runs a host build of
    the same integrator text (g++, OpenMP over the coarse steps like the
    reference's ``nb.prange``, :883-885) on ``number_of_threads`` threads (all
    host cores when None, :1035-1047); the default without a HIP device.
    ``Device.CPU_SINGLE`` runs it on one thread, and ``Device.PYTHON`` -- in
    the reference the uncompiled Python integrator -- runs it on one thread
    too.  ``Device.CUDA`` is not supported.  The CPU devices compute
    ``evaluate``, ``get_time_evolution`` and ``spin_calculator``; batches and
    gradients run on ROCm devices.
    """

    def __init__(self, get_field, spin_quantum_number: SpinQuantumNumber, device: Device = None,
                 exponentiation_method: ExponentiationMethod = None, use_rotating_frame: bool = True,
                 integration_method: IntegrationMethod = IntegrationMethod.MAGNUS_CF4,
                 number_of_squares: int = 24, threads_per_block: int = 64, max_registers: int = None,
                 number_of_threads: int = None, *, samplers=None, lanes_per_step: int = None,
                 precision: str = "fp64"):
        if not device:
            device = _default_device()
        if device == Device.CUDA:
            raise NotImplementedError(
                "Device.CUDA is the reference's numba.cuda back end; spinsim_amd runs on ROCm "
                "devices (Device.ROCM) and on the host (Device.CPU, CPU_SINGLE, PYTHON)")
        self.threads_per_block = threads_per_block
        self.spin_quantum_number = spin_quantum_number
        self.device = device
        self.number_of_threads = number_of_threads
        self.use_rotating_frame = bool(use_rotating_frame)
        self.integration_method = integration_method
        self.number_of_squares = int(number_of_squares)
        self.lanes_per_step = lanes_per_step
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = precision
        self.exponentiation_method = resolve_exponentiation(spin_quantum_number,
                                                            exponentiation_method)
        if precision != "fp64" and self.exponentiation_method == ExponentiationMethod.LIE_TROTTER_8:
            raise ValueError(f"precision={precision!r} covers the x, y, z and q generators only; "
                             "use precision='fp64' with ExponentiationMethod.LIE_TROTTER_8")
        # what computes each step operator (the residual series replaces the
        # exponentiation method in the float32-bulk modes)
        if precision == "fp64":
            self.step_exponential = self.exponentiation_method.value
        else:
            self.step_exponential = ("residual series, " +
                                     ("float64" if precision == "fp64_series" else "float32"))
            if exponentiation_method is not None or number_of_squares != 24:
                warnings.warn(
                    f"precision={precision!r} computes each step operator as exp(A) = I + A + "
                    "a Taylor residual series (scaled and squared when large); "
                    f"exponentiation_method={self.exponentiation_method.name} and "
                    f"number_of_squares={number_of_squares} are not used", stacklevel=2)
        self.dimension = spin_quantum_number.dimension
        self._modules = {}
        self._tables = {}
        self._host = device in HOST_DEVICES
        try:
            self.field = FieldSource(get_field, samplers)
            if self._host:
                from . import host
                self._code = None
                self._host_module = host.compile_module(
                    self.field.source, self.dimension, integration_method.index,
                    self.exponentiation_method.index, self.use_rotating_frame,
                    self.number_of_squares, self.field.has_context, PRECISIONS[precision])
            else:
                self._code = compile_field(self.field.source, self.dimension,
                                           integration_method.index,
                                           self.exponentiation_method.index,
                                           self.use_rotating_frame, self.number_of_squares,
                                           self.field.has_context, PRECISIONS[precision])
        except Exception:
            compiler = "g++" if self._host else "hipRTC"
            print(f"{_RED}spinsim error!!!\n{compiler} could not compile get_field() into a device "
                  f"function.{_RESET}\n")
            raise
        self.get_time_evolution = self._get_time_evolution
        self.spin_calculator = self._spin_calculator

    # ---- host (CPU device) plumbing ----------------------------------------
    def _host_threads(self):
        """Threads of a CPU-device call: number_of_threads (all host cores when
        None) for Device.CPU, one for CPU_SINGLE and PYTHON."""
        if self.device != Device.CPU:
            return 1
        if self.number_of_threads:
            return int(self.number_of_threads)
        return len(os.sched_getaffinity(0))

    def _host_tables(self):
        if not self.field.samplers:
            return None
        versions = tuple(s.version for s in self.field.samplers)
        cached = self._tables.get("host")
        if cached is None or cached[0] != versions:
            from .interpolation import build_host_table_array
            cached = (versions, build_host_table_array(self.field.samplers))
            self._tables["host"] = cached
        return ctypes.addressof(cached[1][0])  # (array, arrays it points at)

    def _host_time_evolution(self, sweep_parameters, time_coarse, time_end_points,
                             time_step_integration, time_step_output, time_evolution_output):
        t0 = float(np.asarray(time_end_points, np.float64)[0])
        self._host_module.time_evolution(sweep_parameters, self._host_tables(), t0,
                                         time_step_integration, time_step_output, time_coarse,
                                         time_evolution_output, self._host_threads())

    def _evaluate_host(self, time_start, time_end, time_step_integration, time_step_output,
                       state_init, sweep_parameters, output):
        from . import host
        dt, n, ends = derive_steps(time_start, time_end, time_step_integration, time_step_output)
        if n < 0:
            raise ValueError("time_end is before time_start")
        d = self.dimension
        time = np.empty(n, np.float64)
        u = np.empty((n, d, d), np.complex128)
        self._host_time_evolution(sweep_parameters, time, ends, dt, time_step_output, u)
        state, _ = host.state(state_init, u)
        if output == "torch":
            import torch
            return Results(torch.from_numpy(time), torch.from_numpy(u), torch.from_numpy(state),
                           self.spin_calculator)
        return Results(time, u, state, self.spin_calculator)

    # ---- device plumbing --------------------------------------------------
    @staticmethod
    def _torch():
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("spinsim_amd: no HIP device is visible (Device.ROCM needs one)")
        return torch

    def _module(self, device):
        if self._host:
            raise RuntimeError(f"this Simulator runs on {self.device}; device launches need "
                               "Device.ROCM")
        key = device.index
        if key not in self._modules:
            torch = self._torch()
            with torch.cuda.device(device):
                self._modules[key] = _native.Module(
                    self._code, self.dimension, self.integration_method.index,
                    self.exponentiation_method.index, self.use_rotating_frame)
        return self._modules[key]

    def _table_ptr(self, device):
        if not self.field.samplers:
            return None
        versions = tuple(s.version for s in self.field.samplers)
        cached = self._tables.get(device.index)
        if cached is None or cached[0] != versions:
            cached = (versions, build_table_array(self.field.samplers, device))
            self._tables[device.index] = cached
        return cached[1].data_ptr()

    def module_info(self):
        torch = self._torch()
        return self._module(torch.device("cuda", torch.cuda.current_device())).info()

    # ---- launches on resident device buffers (no allocation, no copies) -----
    def launch_time_evolution(self, sweep, t0, dt, dt_out, time_out, u_out, lanes_per_step=None,
                              stream=None):
        """Enqueue ss_time_evolution: sweep (P,) float64, time_out (N,) float64 and
        u_out (N, d, d) complex128 are contiguous device tensors."""
        torch = self._torch()
        dev = u_out.device
        n = int(time_out.shape[0])
        lanes = lanes_per_step if lanes_per_step is not None else (self.lanes_per_step or 0)
        if lanes == "auto":
            from .autotune import tune_lanes
            n_fine = math.floor(dt_out / dt + 0.5)
            lanes, _ = tune_lanes(self, n, n_fine, dt, dt_out, sweep, t0) if n else (0, None)
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        mod = self._module(dev)
        _native.check(_native.lib().ss_time_evolution(
            mod.handle, sweep.data_ptr(), float(t0), float(dt), float(dt_out), n,
            time_out.data_ptr(), u_out.data_ptr(), self._table_ptr(dev), int(lanes), stream),
            "ss_time_evolution")

    def launch_state_scan(self, psi0, u, state, workspace, spin=None, stream=None):
        """Enqueue ss_state_scan (and the Bloch vectors into ``spin`` when given);
        psi0 is a host complex128 array of d values."""
        torch = self._torch()
        n, d = int(u.shape[0]), self.dimension
        psi = np.ascontiguousarray(np.asarray(psi0, np.complex128).reshape(d))
        if stream is None:
            stream = torch.cuda.current_stream(u.device).cuda_stream
        _native.check(_native.lib().ss_state_scan(
            psi.ctypes.data, u.data_ptr(), state.data_ptr(),
            None if spin is None else spin.data_ptr(), n, d, workspace.data_ptr(),
            workspace.numel(), stream), "ss_state_scan")

    def launch_spin(self, state, spin, stream=None):
        torch = self._torch()
        n, d = int(state.shape[0]), int(state.shape[1])
        if stream is None:
            stream = torch.cuda.current_stream(state.device).cuda_stream
        _native.check(_native.lib().ss_spin(state.data_ptr(), spin.data_ptr(), n, d, stream),
                      "ss_spin")

    def scan_workspace(self, n, device):
        torch = self._torch()
        nbytes = _native.lib().ss_state_scan_workspace(int(n), self.dimension)
        return torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)

    # ---- reference kernels ----------------------------------------------------
    def _get_time_evolution(self, sweep_parameters, time_coarse, time_end_points,
                            time_step_integration, time_step_output, time_evolution_output,
                            lanes_per_step=None):
        """get_time_evolution(sweep, time_coarse, time_end_points, dt, dt_out, U), :854-896.

        ``time_coarse`` (N,) float64 and ``time_evolution_output`` (N, d, d)
        complex128 are device tensors filled in place (numpy arrays are copied
        through the device); on the CPU devices they are numpy arrays.
        """
        if self._host:
            self._host_time_evolution(sweep_parameters, time_coarse, time_end_points,
                                      time_step_integration, time_step_output,
                                      time_evolution_output)
            return None
        torch = self._torch()
        out_np = isinstance(time_evolution_output, np.ndarray)
        if out_np:
            dev = torch.device("cuda", torch.cuda.current_device())
            t_dev = torch.empty(time_coarse.shape[0], dtype=torch.float64, device=dev)
            u_dev = torch.empty(time_evolution_output.shape, dtype=torch.complex128, device=dev)
        else:
            t_dev, u_dev = time_coarse, time_evolution_output
            dev = u_dev.device
        n = int(t_dev.shape[0])
        d = self.dimension
        if tuple(u_dev.shape) != (n, d, d) or u_dev.dtype != torch.complex128 \
                or not u_dev.is_contiguous() or t_dev.dtype != torch.float64 \
                or not t_dev.is_contiguous() or t_dev.device != dev:
            raise ValueError(f"time_coarse must be ({n},) float64 and time_evolution_output "
                             f"({n}, {d}, {d}) complex128, contiguous, on one device")
        sweep = torch.as_tensor(np.atleast_1d(np.asarray(sweep_parameters, np.float64)),
                                device=dev)
        t0 = float(np.asarray(time_end_points, np.float64)[0])
        self.launch_time_evolution(sweep, t0, time_step_integration, time_step_output, t_dev,
                                   u_dev, lanes_per_step)
        if out_np:
            time_coarse[...] = t_dev.cpu().numpy()
            time_evolution_output[...] = u_dev.cpu().numpy()
        return None

    def state_scan(self, state_init, time_evolution):
        """get_state on the device, :1087-1111: state[i] = U[i-1] state[i-1]."""
        torch = self._torch()
        u = time_evolution
        n, d = int(u.shape[0]), self.dimension
        state = torch.empty((n, d), dtype=torch.complex128, device=u.device)
        self.launch_state_scan(state_init, u, state, self.scan_workspace(n, u.device))
        return state

    def _spin_calculator(self, state):
        """Bloch vectors of a state time series (spin_calculator, :961-988)."""
        if self._host:
            from . import host
            if not isinstance(state, np.ndarray):
                return torch_from_host(host.spin(state.cpu().numpy(), self._host_threads()))
            return host.spin(state, self._host_threads())
        torch = self._torch()
        host = isinstance(state, np.ndarray)
        if host:
            st = torch.as_tensor(np.ascontiguousarray(state, np.complex128),
                                 device=torch.device("cuda", torch.cuda.current_device()))
        else:
            st = state.contiguous()
        spin = torch.empty((int(st.shape[0]), 3), dtype=torch.float64, device=st.device)
        self.launch_spin(st, spin)
        return spin.cpu().numpy() if host else spin

    # ---- evaluate ---------------------------------------------------------
    def evaluate(self, time_start: float, time_end: float, time_step_integration: float,
                 time_step_output: float, state_init, sweep_parameters=[0], *,
                 output: str = "numpy", lanes_per_step: int = None) -> Results:
        """Integrate and return the state time series (spinsim/__init__.py:993-1085)."""
        if output not in ("numpy", "torch"):
            raise ValueError("output must be 'numpy' or 'torch'")
        if self._host:
            return self._evaluate_host(time_start, time_end, time_step_integration,
                                       time_step_output, state_init, sweep_parameters, output)
        torch = self._torch()
        dt, n, ends = derive_steps(time_start, time_end, time_step_integration, time_step_output)
        if n < 0:
            raise ValueError("time_end is before time_start")
        d = self.dimension
        dev = torch.device("cuda", torch.cuda.current_device())
        time = torch.empty(n, dtype=torch.float64, device=dev)
        u = torch.empty((n, d, d), dtype=torch.complex128, device=dev)
        self._get_time_evolution(sweep_parameters, time, ends, dt, time_step_output, u,
                                 lanes_per_step=lanes_per_step)
        if output == "torch":
            return Results(time, u, self.state_scan(state_init, u), self.spin_calculator)
        # numpy results: time and U go to pinned host memory on a side stream
        # while the state scan runs, then the states follow
        pinned = n * (d * d + d + 1) * 16 <= PINNED_RESULT_BYTES
        if not pinned:
            state = self.state_scan(state_init, u)
            return Results(time.cpu().numpy(), u.cpu().numpy(), state.cpu().numpy(),
                           self.spin_calculator, _device_state=state)
        main = torch.cuda.current_stream(dev)
        side = self._copy_stream(dev)
        h_time = torch.empty(time.shape, dtype=time.dtype, pin_memory=True)
        h_u = torch.empty(u.shape, dtype=u.dtype, pin_memory=True)
        h_state = torch.empty((n, d), dtype=u.dtype, pin_memory=True)
        side.wait_stream(main)
        with torch.cuda.stream(side):
            h_time.copy_(time, non_blocking=True)
            h_u.copy_(u, non_blocking=True)
        state = self.state_scan(state_init, u)
        h_state.copy_(state, non_blocking=True)
        main.synchronize()
        side.synchronize()
        return Results(h_time.numpy(), h_u.numpy(), h_state.numpy(), self.spin_calculator,
                       _device_state=state)

    def _copy_stream(self, device):
        """A side stream (default priority) per device for result copies."""
        streams = self.__dict__.setdefault("_copy_streams", {})
        if device.index not in streams:
            streams[device.index] = self._torch().cuda.Stream(device)
        return streams[device.index]

    def evaluate_batch(self, time_start, time_end, time_step_integration, time_step_output,
                       state_init, sweep_parameters, **kw):
        """Evaluate B simulations (sweep rows) in one persistent launch.

        Any of the times/steps may be per-simulation (B,) arrays and
        ``state_init`` may be (B, d); returns :class:`spinsim_amd.batch.BatchResults`.
        """
        if self._host:
            raise NotImplementedError("batched evaluation runs on ROCm devices (Device.ROCM)")
        from .batch import evaluate_batch
        return evaluate_batch(self, time_start, time_end, time_step_integration,
                              time_step_output, state_init, sweep_parameters, **kw)


def torch_from_host(a):
    import torch
    return torch.from_numpy(a)


def precision_error(sim: Simulator, time_start, time_end, time_step_integration, time_step_output,
                    state_init, sweep_parameters=[0], **kw) -> dict:
    """Error of a float32-bulk Simulator against float64 on the same inputs.

    Builds a float64 twin of ``sim`` (same field, spin, methods and frame), runs
    both on the device and returns the maximum absolute differences of the state
    and of the coarse step operators, and the float32-bulk run's unitarity
    deviation max |U^+ U - I|.
    """
    import torch

    twin = Simulator(sim.field, sim.spin_quantum_number,
                     exponentiation_method=sim.exponentiation_method,
                     use_rotating_frame=sim.use_rotating_frame,
                     integration_method=sim.integration_method,
                     number_of_squares=sim.number_of_squares, precision="fp64")
    args = (time_start, time_end, time_step_integration, time_step_output, state_init,
            sweep_parameters)
    a = sim.evaluate(*args, output="torch", **kw)
    b = twin.evaluate(*args, output="torch", **kw)
    u = a.time_evolution
    eye = torch.eye(sim.dimension, dtype=u.dtype, device=u.device)
    unitarity = (torch.matmul(u.conj().transpose(1, 2).contiguous(), u) - eye).abs()
    return {"state": float((a.state - b.state).abs().max()) if a.state.numel() else 0.0,
            "time_evolution": float((u - b.time_evolution).abs().max()) if u.numel() else 0.0,
            "unitarity": float(unitarity.max()) if u.numel() else 0.0}
