"""Benchmark: Magnus-CF4 steps/s of spinsim_amd's integration path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--no-variants] [--copy-probe MB]

Headline workload (``value``): the spin-one Rabi drive -- x = 2pi*2 kHz
cos(2pi*700 kHz t), z = 2pi*700 kHz, quadratic shift 2pi*1 kHz; t = 0 -> 100 ms,
integration step 100 ns, output step 1 us: 1e5 coarse x 10 fine = 1e6 CF4 steps;
Lie-Trotter exponentials (24 squares), rotating frame, plain float64.  A step
is one evaluate() pass over inputs resident in HBM: the time-evolution kernel,
the state prefix product with the Bloch vectors fused into its last pass
(outputs stay on the device).

Also reported (``variants``): the same workload in float32-bulk, in the
float64 run of the float32-bulk series (the same algorithm in float64: the
ratio of the two is what float32 itself buys) and in fp32 (float32-bulk plus
the field function's sin/cos/exp in float32 on float64-reduced arguments);
the 4096-simulation spin-half Rabi ensemble with uneven durations (1-50 ms,
50 ns steps, 1 us outputs) on identical inputs through the batched
float32-bulk, fp32 and float64 graphs (one persistent launch + batch state
scan, HIP-graph replay) and through a loop of one plain float64 evaluate per
simulation (the reference's way through a sweep), each checked on 64 sampled
simulations against the float64 oracle, with the batched/plain ratios, and
the batched fp32 and float64 graphs again with operators stored as first
columns (states and Bloch vectors only); the
differentiation workload (1024 spin-one pulses, forward + adjoint gradient,
float64 and float32-bulk forward, and the deterministic gradient reduction),
with its gradient error against central differences; the long noisy spin-half
run (1e7 steps; float64, float32-bulk, fp32); and the spin-half Larmor
precession (1e6 steps, analytic exponentials).  The CPU baselines (the
oracle's C restatement) run on every host core this process may use
(reported with the CPU model), timed over the oracle runs they check
against (``cpu_steps_per_s``, with the sample named).  ``value`` is withheld
(null, exit status 1) if the headline's state error exceeds 1e-6.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VARIANTS = ("fp32_bulk", "batched", "adjoint", "long_noisy", "larmor")
METRIC = "Magnus-CF4 steps/sec on 1 GPU at state err ≤1e-6 vs FP64 oracle; batched/plain"

RABI_ONE = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = SS_TAU * 2e3 * cos(SS_TAU * 7e5 * t); f[1] = 0; f[2] = SS_TAU * 7e5; f[3] = SS_TAU * 1e3; }
"""
T0, T1, DT, DT_OUT = 0.0, 100e-3, 100e-9, 1e-6

LARMOR = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = 0; f[1] = 0; f[2] = SS_TAU * 1e3; }
"""

# Swept spin-half Rabi drive: sweep = (carrier, Rabi amplitude, detuning), rad/s.
RABI_SWEEP = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = 2 * s[1] * cos(s[0] * t); f[1] = 0; f[2] = s[0] + s[2]; }
"""

# Algorithmic float64 flops per spin-one Lie-Trotter CF4 fine step, counted from
# the reference's expressions (1 flop per real +,-,*,/,sqrt; complex * = 6,
# complex + = 2; transcendentals excluded): 2 exponentials x (24 residual
# squarings x 207 + 143 set-up) + 2 matrix products x 198 + frame 24 + CF4
# combinations 32 + sampling 14 + time update 1.
FLOPS_PER_STEP = 2 * (24 * 207 + 143) + 2 * 198 + 24 + 32 + 14 + 1
# The same count for the kernel's method: its residual square forms two rows
# of the product and derives the third from special unitarity, 185 flops (84
# FMA, 12 MUL, 5 ADD) where the reference's full square counts 207 (and
# compiles to 201).  rocprofv3's counters read 9,929 per step for the
# three-wave launch, which also does the two-lane row products.
FLOPS_PER_STEP_EXECUTED = 2 * (24 * 185 + 143) + 2 * 198 + 24 + 32 + 14 + 1
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per time-evolution launch of the headline workload (phased launch
# on one and on three waves per SIMD read the same), from rocprofv3 FETCH_SIZE /
# WRITE_SIZE counter passes (KiB per dispatch) scaled by what the same counters
# read for a 256 MiB device copy (fetch x2.0, write x1.0); kept with the
# profile summaries in profiles/.
PMC_FETCH_KIB, PMC_WRITE_KIB, FETCH_SCALE, WRITE_SCALE = 176.0, 14843.75, 2.0, 1.0
TRAFFIC_BYTES = (PMC_FETCH_KIB * FETCH_SCALE + PMC_WRITE_KIB * WRITE_SCALE) * 1024


def host_cores():
    """Host threads of the CPU baselines: every core this process may run on
    (len(os.sched_getaffinity(0))), lowered only to a cgroup CPU quota when one
    is set (the core time actually available to it), with the CPU model."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, period = fh.read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    threads = affinity if quota is None else max(1, min(affinity, int(quota)))
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": threads, "affinity": affinity, "cgroup_quota_cpus": quota,
            "model": model, "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(co, psi0, dt, n, nf, cores, reps=3):
    """The oracle's C restatement on the host cores, whole workload, median of reps."""
    rates = []
    state = None
    for _ in range(reps):
        t = time.perf_counter()
        _, u = co.time_evolution([0.0], T0, dt, DT_OUT, nf, n, threads=cores)
        state = co.get_state(psi0, u)
        co.get_spin(state, threads=cores)
        rates.append(n * nf / (time.perf_counter() - t))
    return statistics.median(rates), state


def cpu_single_core(co, dt, nf, rows=10000, reps=3):
    """The C restatement on one host thread over the first `rows` coarse steps
    (the CPU_SINGLE analogue), median steps/s of reps."""
    rates = []
    for _ in range(reps):
        t = time.perf_counter()
        co.time_evolution([0.0], T0, dt, DT_OUT, nf, rows, threads=1)
        rates.append(rows * nf / (time.perf_counter() - t))
    return statistics.median(rates)


def evaluate_end_to_end(ss, torch, lanes, reps=3):
    """Median ms of Simulator.evaluate() returning numpy arrays (time evolution,
    scan, device-to-host copies of time, U and state) on the same workload."""
    sim = ss.Simulator(RABI_ONE, ss.SpinQuantumNumber.ONE)
    psi0 = ss.SpinQuantumNumber.ONE.plus_z
    sim.evaluate(T0, T1, DT, DT_OUT, psi0, lanes_per_step=lanes)
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        sim.evaluate(T0, T1, DT, DT_OUT, psi0, lanes_per_step=lanes)
        times.append(1e3 * (time.perf_counter() - t))
    return statistics.median(times)


class TimingEvent:
    """A HIP event for timing only, created on the HIP runtime torch has
    loaded (which owns the streams).  The time-evolution kernels are timed by
    handing a pair to ss_time_evolution_events: the launch itself stamps the
    kernel's own start and end, so the pair reads the kernel alone rather
    than the span between two markers around it (0.5-2% shorter, same pass
    time).  Where an event is recorded on a stream it is recorded without the
    system-scope memory fence (hipEventDisableSystemFence), so it writes its
    timestamp without the cache writeback and invalidation a default event
    adds between the kernels it separates."""

    DISABLE_SYSTEM_FENCE = 0x20000000
    _hip = None

    @classmethod
    def _runtime(cls):
        import ctypes

        if cls._hip is None:
            path = None
            with open("/proc/self/maps") as fh:
                for line in fh:
                    if "libamdhip64.so" in line:
                        path = line.split()[-1]
                        break
            if path is None:
                raise RuntimeError("the HIP runtime is not loaded")
            hip = ctypes.CDLL(path)
            hip.hipEventCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
            hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                                ctypes.c_void_p]
            hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
            cls._hip = hip
        return cls._hip

    def __init__(self):
        import ctypes

        self.hip = self._runtime()
        self.handle = ctypes.c_void_p()
        self._check(self.hip.hipEventCreateWithFlags(ctypes.byref(self.handle),
                                                     self.DISABLE_SYSTEM_FENCE), "create")

    @staticmethod
    def _check(rc, what):
        if rc != 0:
            raise RuntimeError(f"hipEvent{what} failed with HIP error {rc}")

    def record(self, stream):
        self._check(self.hip.hipEventRecord(self.handle, stream.cuda_stream), "Record")

    def elapsed_time(self, end):
        """Milliseconds from this event to `end` (both recorded and complete)."""
        import ctypes

        ms = ctypes.c_float()
        self._check(self.hip.hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle),
                    "ElapsedTime")
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            self.hip.hipEventDestroy(self.handle)


def resident_passes(ss, torch, sim, psi0, t0, t1, dt_in, dt_out, steps, warmup, lanes):
    """Time K evaluate() passes of one simulation over device-resident buffers
    (time evolution, then the state scan with the Bloch vectors fused);
    returns a dict, the states and the time stamps."""
    from spinsim_amd import _native
    from spinsim_amd.simulator import derive_steps

    dev = torch.device("cuda", 0)
    d = sim.dimension
    dt, n, _ = derive_steps(t0, t1, dt_in, dt_out)
    nf = math.floor(dt_out / dt + 0.5)
    sweep = torch.zeros(1, dtype=torch.float64, device=dev)
    time_out = torch.empty(n, dtype=torch.float64, device=dev)
    u = torch.empty((n, d, d), dtype=torch.complex128, device=dev)
    state = torch.empty((n, d), dtype=torch.complex128, device=dev)
    spin = torch.empty((n, 3), dtype=torch.float64, device=dev)
    ws = sim.scan_workspace(n, dev)
    stream = torch.cuda.current_stream(dev)
    table = None
    if lanes == "auto":
        from spinsim_amd.autotune import tune_lanes
        lanes, table = tune_lanes(sim, n, nf, dt, dt_out, sweep, t0)

    L = _native.lib()

    def step(events=None):
        if events is not None:
            # the launch stamps the kernel's own start and end into the events
            # (no marker packets between the pass's kernels)
            L.ss_time_evolution_events(events[0].handle, events[1].handle)
        sim.launch_time_evolution(sweep, t0, dt, dt_out, time_out, u, lanes)
        sim.launch_state_scan(psi0, u, state, ws, spin=spin)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [(TimingEvent(), TimingEvent()) for _ in range(steps)]
    t_start = time.perf_counter()
    for k in range(steps):
        step(ev[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    kernel_ms = statistics.mean(a.elapsed_time(b) for a, b in ev)
    out = {"steps_per_s": n * nf / (elapsed / steps), "ms_per_step": 1e3 * elapsed / steps,
           "kernel_ms": kernel_ms, "lanes_per_step": lanes, "coarse": n, "fine": n * nf,
           "autotune_ms": None if table is None else {str(k): v for k, v in table.items()},
           "dt": dt, "n_fine": nf}
    return out, state.cpu().numpy(), time_out.cpu().numpy()


def single_workload(ss, torch, precision, steps, warmup, lanes):
    """Time K evaluate() passes of the spin-one workload; returns a dict and the state."""
    sim = ss.Simulator(RABI_ONE, ss.SpinQuantumNumber.ONE, precision=precision)
    psi0 = ss.SpinQuantumNumber.ONE.plus_z
    out, state, _ = resident_passes(ss, torch, sim, psi0, T0, T1, DT, DT_OUT, steps, warmup,
                                    lanes)
    return out, state, psi0


def larmor_workload(ss, torch, steps, warmup, cpu_cores):
    """Spin-half Larmor precession, field (0, 0, 2pi 1 kHz), t = 0 - 100 ms, dt 100 ns,
    output 500 ns (2e5 x 5 = 1e6 CF4 steps), +x, analytic exponentials, rotating
    frame, float64: steps/s per device-resident evaluate pass, state error against
    the closed form psi(t) = (exp(-i pi 1e3 t), exp(i pi 1e3 t)) / sqrt 2 and
    against the oracle, and the oracle's C restatement timed on the host cores."""
    import numpy as np

    from oracle.c_oracle import COracle

    sim = ss.Simulator(LARMOR, ss.SpinQuantumNumber.HALF,
                       exponentiation_method=ss.ExponentiationMethod.ANALYTIC)
    psi0 = ss.SpinQuantumNumber.HALF.plus_x
    out, state, t = resident_passes(ss, torch, sim, psi0, 0.0, 100e-3, 100e-9, 500e-9, steps,
                                    warmup, "auto")
    exact = np.stack([np.exp(-1j * math.pi * 1e3 * t), np.exp(1j * math.pi * 1e3 * t)], 1)
    exact /= math.sqrt(2)
    co = COracle(LARMOR, 2, expo="analytic")
    n, nf, dt = out["coarse"], out["n_fine"], out["dt"]
    rates, ref = [], None
    for _ in range(3 if cpu_cores else 1):
        c0 = time.perf_counter()
        _, u = co.time_evolution([0.0], 0.0, dt, 500e-9, nf, n, threads=cpu_cores or None)
        ref = co.get_state(psi0, u)
        co.get_spin(ref, threads=cpu_cores or None)
        rates.append(n * nf / (time.perf_counter() - c0))
    res = {k: out[k] for k in ("steps_per_s", "ms_per_step", "kernel_ms", "lanes_per_step",
                               "coarse", "fine")}
    res["state_err_vs_closed_form"] = float(np.abs(state - exact).max())
    res["state_err_vs_oracle"] = float(np.abs(state - ref).max())
    if cpu_cores:
        res["cpu_steps_per_s"] = statistics.median(rates)
        res["cpu_sample"] = (f"whole workload (time evolution, state product, spin) on "
                             f"{cpu_cores} host threads, median of 3")
    return res


def ensemble_inputs(n_sims=4096):
    """The 4096-simulation spin-half Rabi ensemble: sweep rows (carrier, Rabi
    amplitude, detuning; rad/s) and durations, seeded."""
    import numpy as np

    rng = np.random.default_rng(20261202)
    tau = 2 * math.pi
    sweep = np.stack([np.full(n_sims, tau * 7e5), tau * rng.uniform(500, 5000, n_sims),
                      tau * rng.normal(0, 200, n_sims)], 1)
    dur = rng.uniform(1e-3, 50e-3, n_sims)
    return sweep, dur


def ensemble_check_sims(dur, check):
    """Sampled simulations of the ensemble: the first, the last, the longest and
    the shortest, and the rest spread evenly over the CSR range."""
    import numpy as np

    n = len(dur)
    picks = {0, n - 1, int(np.argmax(dur)), int(np.argmin(dur))}
    for s in np.linspace(0, n - 1, check).astype(int):
        if len(picks) >= check:
            break
        picks.add(int(s))
    return sorted(picks)


def ensemble_oracle_states(sweep, dur, sims, psi0, cpu_cores):
    """Float64 oracle states (C restatement) of the sampled simulations, and the
    host seconds and CF4 steps they took."""
    from oracle import spinsim_oracle as O
    from oracle.c_oracle import COracle

    co = COracle(RABI_SWEEP, 2)
    refs, steps, cpu_s = {}, 0, 0.0
    for s in sims:
        dt, n, nf, _ = O.timing(0.0, dur[s], 50e-9, 1e-6)
        c0 = time.perf_counter()
        _, u = co.time_evolution(sweep[s], 0.0, dt, 1e-6, nf, n, threads=cpu_cores)
        refs[s] = co.get_state(psi0, u)
        cpu_s += time.perf_counter() - c0
        steps += n * nf
    return refs, steps, cpu_s


def batched_workload(ss, torch, precision, sweep, dur, refs, steps, warmup, n_sims=4096,
                     compact=False):
    """The ensemble in one persistent launch + batch state scan (Bloch vectors
    fused), replayed as a HIP graph, autotuned; steps/s and the state error of
    the sampled simulations against the float64 oracle.  compact: operators
    stored as first columns (32 B a row instead of 64)."""
    import numpy as np

    from spinsim_amd.autotune import tune_batch

    sim = ss.Simulator(RABI_SWEEP, ss.SpinQuantumNumber.HALF, precision=precision)
    psi0 = ss.SpinQuantumNumber.HALF.plus_z
    plan = ss.BatchPlan(sim, 0.0, dur, 50e-9, 1e-6, psi0, sweep, with_spin=True, compact=compact)
    tune_batch(plan, reps=2, warmup=1, lanes=(1, 2, 4), items=(1, 4, 8, 16))
    plan.capture()
    for _ in range(warmup):
        plan.replay()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(steps):
        plan.replay()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    errs = []
    for s, ref in refs.items():
        a, b = int(plan.offsets_host[s]), int(plan.offsets_host[s + 1])
        errs.append(float(np.abs(plan.state[a:b].cpu().numpy() - ref).max()))
    out = {"steps_per_s": plan.fine_steps / (elapsed / steps), "ms_per_step": 1e3 * elapsed / steps,
           "fine_steps": plan.fine_steps, "coarse_rows": plan.total, "simulations": n_sims,
           "lanes_per_step": plan.lanes, "items_per_grab": plan.items_per_grab,
           "scan_chunk_rows": plan.scan_chunk_rows,
           "state_err_vs_oracle_sampled": max(errs), "sampled_simulations": len(refs),
           "precision": precision, "graph": True,
           "operators": "first columns" if compact else "full"}
    del plan
    torch.cuda.empty_cache()
    return out


def plain_loop_workload(ss, torch, sweep, dur, refs, reps=2):
    """The reference's way through a sweep (one evaluate() per sweep point,
    spinsim/__init__.py:993): the plain float64 single-simulation path -- time
    evolution, state scan and Bloch vectors -- once per ensemble member, into
    device-resident buffers, one stream; steps/s over the whole ensemble and
    the state error of the sampled simulations against the oracle."""
    import numpy as np

    from spinsim_amd.simulator import derive_steps

    n_sims = len(dur)
    dev = torch.device("cuda", 0)
    sim = ss.Simulator(RABI_SWEEP, ss.SpinQuantumNumber.HALF)
    psi0 = ss.SpinQuantumNumber.HALF.plus_z
    shapes = [derive_steps(0.0, float(dur[s]), 50e-9, 1e-6, warn=False) for s in range(n_sims)]
    rows = np.array([n for _, n, _ in shapes], np.int64)
    offsets = np.concatenate([[0], np.cumsum(rows)])
    total = int(offsets[-1])
    fine = int(sum(n * math.floor(1e-6 / dt + 0.5) for dt, n, _ in shapes))
    sweep_dev = torch.as_tensor(np.ascontiguousarray(sweep), device=dev)
    time_out = torch.empty(total, dtype=torch.float64, device=dev)
    u = torch.empty((total, 2, 2), dtype=torch.complex128, device=dev)
    state = torch.empty((total, 2), dtype=torch.complex128, device=dev)
    spin = torch.empty((total, 3), dtype=torch.float64, device=dev)
    ws = sim.scan_workspace(int(rows.max()), dev)

    def one_pass():
        for s in range(n_sims):
            a, b = int(offsets[s]), int(offsets[s + 1])
            if a == b:
                continue
            dt = shapes[s][0]
            sim.launch_time_evolution(sweep_dev[s], 0.0, dt, 1e-6, time_out[a:b], u[a:b], 0)
            sim.launch_state_scan(psi0, u[a:b], state[a:b], ws, spin=spin[a:b])

    one_pass()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(reps):
        one_pass()
    torch.cuda.synchronize()
    elapsed = (time.perf_counter() - t_start) / reps
    errs = []
    for s, ref in refs.items():
        a, b = int(offsets[s]), int(offsets[s + 1])
        errs.append(float(np.abs(state[a:b].cpu().numpy() - ref).max()))
    out = {"steps_per_s": fine / elapsed, "ms_per_pass": 1e3 * elapsed, "fine_steps": fine,
           "simulations": n_sims, "launches_per_pass": 3 * n_sims, "lanes_per_step": 0,
           "state_err_vs_oracle_sampled": max(errs), "sampled_simulations": len(refs),
           "precision": "fp64"}
    del u, state, spin, time_out
    torch.cuda.empty_cache()
    return out


def ensemble_same_inputs(ss, torch, steps, cpu_cores, check=64):
    """The ensemble's identical inputs through the batched float32-bulk and
    batched float64 graphs and through the plain float64 loop of one evaluate
    per simulation; the ratios the batched/plain target is stated on."""
    sweep, dur = ensemble_inputs()
    sims = ensemble_check_sims(dur, check)
    psi0 = ss.SpinQuantumNumber.HALF.plus_z
    refs, cpu_steps, cpu_s = ensemble_oracle_states(sweep, dur, sims, psi0, cpu_cores)
    out = {"batched_fp32_bulk": batched_workload(ss, torch, "fp32_bulk", sweep, dur, refs, steps, 1),
           "batched_fp32": batched_workload(ss, torch, "fp32", sweep, dur, refs, steps, 1),
           "batched_fp64": batched_workload(ss, torch, "fp64", sweep, dur, refs, steps, 1),
           "plain_fp64_loop": plain_loop_workload(ss, torch, sweep, dur, refs),
           # states and Bloch vectors only: operators kept as first columns
           "batched_fp32_compact": batched_workload(ss, torch, "fp32", sweep, dur, refs, steps, 1,
                                                    compact=True),
           "batched_fp64_compact": batched_workload(ss, torch, "fp64", sweep, dur, refs, steps, 1,
                                                    compact=True)}
    plain = out["plain_fp64_loop"]["steps_per_s"]
    out["ratio_batched_fp32_bulk_over_plain_fp64"] = out["batched_fp32_bulk"]["steps_per_s"] / plain
    out["ratio_batched_fp64_over_plain_fp64"] = out["batched_fp64"]["steps_per_s"] / plain
    out["ratio_fp32_bulk_over_fp64_batched"] = (out["batched_fp32_bulk"]["steps_per_s"]
                                                / out["batched_fp64"]["steps_per_s"])
    out["ratio_batched_fp32_over_plain_fp64"] = out["batched_fp32"]["steps_per_s"] / plain
    out["ratio_fp32_over_fp32_bulk_batched"] = (out["batched_fp32"]["steps_per_s"]
                                                / out["batched_fp32_bulk"]["steps_per_s"])
    out["ratio_compact_over_full_fp32"] = (out["batched_fp32_compact"]["steps_per_s"]
                                           / out["batched_fp32"]["steps_per_s"])
    out["ratio_compact_over_full_fp64"] = (out["batched_fp64_compact"]["steps_per_s"]
                                           / out["batched_fp64"]["steps_per_s"])
    out["target_2x_met"] = bool(out["ratio_batched_fp32_bulk_over_plain_fp64"] >= 2.0 and
                                out["batched_fp32_bulk"]["state_err_vs_oracle_sampled"] <= 1e-6)
    if cpu_cores:
        out["cpu_steps_per_s"] = cpu_steps / cpu_s
        out["cpu_sample"] = (f"the {len(sims)} sampled simulations ({cpu_steps} CF4 steps, "
                             f"float64 time evolution and state product) on {cpu_cores} host "
                             "threads")
    return out


def long_noisy_inputs(seed=20261202):
    """Spin-half long noisy run: z = 2pi (700 kHz + 8 Gaussian pulses) + 1/f noise,
    x = 2pi 1 kHz cos(2pi 700 kHz t); t = 0 - 100 ms.  Returns (field, sweep, sampler
    args): pulses in sweep (amplitudes Hz, centres s, widths s), the noise as a
    cubic-Hermite table on a 1 us grid (std 2pi 50 rad/s, PSD ~ 1/f, no DC)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    amp = rng.uniform(1e3, 1e4, 8)
    centre = rng.uniform(0, 0.1, 8)
    width = rng.uniform(5e-6, 50e-6, 8)
    grid = np.arange(100001) * 1e-6
    spec = np.fft.rfft(rng.normal(size=grid.size))
    freq = np.fft.rfftfreq(grid.size, 1e-6)
    spec[0] = 0
    spec[1:] /= np.sqrt(freq[1:])
    noise = np.fft.irfft(spec, grid.size)
    noise *= 2 * math.pi * 50 / noise.std()
    return np.concatenate([amp, centre, width]), grid, noise


def make_long_noisy_field(noise_sampler):
    def get_field(time_sample, sweep_parameters, field_sample):
        z = 7e5
        for k in range(8):
            d = time_sample - sweep_parameters[8 + k]
            z += sweep_parameters[k] * math.exp(-d * d / (2 * sweep_parameters[16 + k] ** 2))
        field_sample[2] = 2 * math.pi * z + noise_sampler(time_sample)
        field_sample[0] = 2 * math.pi * 1e3 * math.cos(2 * math.pi * 7e5 * time_sample)
        field_sample[1] = 0
    return get_field


def long_noisy_workload(ss, torch, reps=3, cpu_cores=None):
    """1e7 CF4 steps (dt 10 ns, output 1 us), float64 and float32-bulk: steps/s,
    unitarity drift max |U^+ U - I| of the coarse operators, the float32-bulk
    state error against float64, both state errors against the oracle, and the
    oracle's C restatement timed on cpu_cores host threads."""
    import numpy as np

    from oracle.c_oracle import COracle

    sweep, grid, noise = long_noisy_inputs()
    sampler = ss.generate_interpolation_sampler(grid, noise)
    field = make_long_noisy_field(sampler)
    psi0 = ss.SpinQuantumNumber.HALF.plus_z
    out, states = {}, {}
    for prec in ("fp64", "fp32_bulk", "fp32"):
        sim = ss.Simulator(field, ss.SpinQuantumNumber.HALF, precision=prec)
        res = sim.evaluate(0.0, 0.1, 10e-9, 1e-6, psi0, sweep, output="torch",
                           lanes_per_step="auto")
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            res = sim.evaluate(0.0, 0.1, 10e-9, 1e-6, psi0, sweep, output="torch",
                               lanes_per_step="auto")
        torch.cuda.synchronize()
        el = (time.perf_counter() - t) / reps
        u = res.time_evolution
        eye = torch.eye(2, dtype=u.dtype, device=u.device)
        drift = float((torch.matmul(u.conj().transpose(1, 2).contiguous(), u) - eye).abs().max())
        states[prec] = res.state.cpu().numpy()
        out[prec] = {"steps_per_s": 1e7 / el, "ms_per_evaluate": 1e3 * el,
                     "unitarity_drift": drift}
    for prec in ("fp32_bulk", "fp32"):
        out[prec]["state_err_vs_fp64"] = float(np.abs(states[prec] - states["fp64"]).max())
    src, _, _ = ss.transpile(field)
    co = COracle(src, 2, sampler_names=("noise_table",))
    co.set_samplers([(grid, noise, 3, True)])
    c0 = time.perf_counter()
    _, u = co.time_evolution(sweep, 0.0, 1e-8, 1e-6, 100, 100000, threads=cpu_cores)
    ref = co.get_state(psi0, u)
    cpu_s = time.perf_counter() - c0
    for prec in ("fp64", "fp32_bulk", "fp32"):
        out[prec]["state_err_vs_oracle"] = float(np.abs(states[prec] - ref).max())
    if cpu_cores:
        out["cpu_steps_per_s"] = 1e7 / cpu_s
        out["cpu_sample"] = (f"whole workload (1e7 CF4 steps, float64 time evolution and state "
                             f"product) on {cpu_cores} host threads, one run")
    return out


def _z_bias(time_sample, sweep_parameters, field_sample):
    field_sample[2] = 2 * math.pi * 1e5


# The differentiation workload's field for the oracle: two per-simulation
# cubic control tables on x and y and the z bias.
CONTROLLED_ONE = """
__device__ void get_field(double t, const double* s, double* f, SS_CONTEXT)
{ f[0] = ctl_x(t); f[1] = ctl_y(t); f[2] = SS_TAU * 1e5; f[3] = 0; }
"""


def adjoint_workload(ss, torch, precision, n_sims=1024, reps=3, fd_sims=8, fd_coords=8,
                     cpu_cores=None, check=8):
    """1024 spin-one state-transfer pulses (2 ms, 20 ns steps, 2000 cubic control
    samples on x and y), loss 1 - |psi(T)[-1]|^2: forward + adjoint gradient,
    checked against central differences of the float64 forward; psi(T) of
    `check` sampled pulses against the oracle, whose float64 forward runs (C
    restatement on cpu_cores host threads) give the CPU rate."""
    import numpy as np

    from oracle.c_oracle import COracle

    rng = np.random.default_rng(20261202)
    t_end, m, sigma = 2e-3, 2000, 2 * math.pi * 5e3
    tk = np.linspace(0.0, t_end, m)
    amp = rng.normal(0, sigma, (n_sims, 2, m))
    psi0 = ss.SpinQuantumNumber.ONE.plus_z
    cs = ss.ControlledSimulator(ss.SpinQuantumNumber.ONE, [ss.Control("x", tk),
                                                           ss.Control("y", tk)], _z_bias,
                                precision=precision)
    amps = torch.tensor(amp, device="cuda")

    def run(deterministic=False):
        a = amps.clone().requires_grad_(True)
        psi = cs.final_state(a, 0.0, t_end, 20e-9, 1e-6, psi0, deterministic=deterministic)
        (1 - psi[:, -1].abs() ** 2).sum().backward()
        return a.grad, psi

    grad, psi_t = run()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        grad, psi_t = run()
    torch.cuda.synchronize()
    elapsed = (time.perf_counter() - t) / reps
    # the deterministic gradient reduction: its cost, its agreement with the
    # atomic sums, and its repeatability
    det = {}
    if precision == "fp64":
        g_det, _ = run(True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            g_det2, _ = run(True)
        torch.cuda.synchronize()
        det = {"deterministic_forward_backward_ms": 1e3 * (time.perf_counter() - t) / reps,
               "deterministic_vs_atomic_rel": float((g_det - grad).abs().max() / grad.abs().max()),
               "deterministic_repeat_identical": bool(torch.equal(g_det, g_det2))}
    psi_t = psi_t.detach().cpu().numpy()
    co = COracle(CONTROLLED_ONE, 3, sampler_names=("ctl_x", "ctl_y"))
    co.set_samplers([(tk, amp[:, 0], 3, True), (tk, amp[:, 1], 3, True)])
    psi_err, cpu_s = 0.0, 0.0
    for s_ in np.linspace(0, n_sims - 1, check).astype(int):
        c0 = time.perf_counter()
        _, u = co.time_evolution([0.0], 0.0, 20e-9, 1e-6, 50, 2000, sim=int(s_),
                                 threads=cpu_cores)
        st = co.get_state(psi0, u)
        cpu_s += time.perf_counter() - c0
        psi_err = max(psi_err, float(np.abs(u[-1] @ st[-1] - psi_t[s_]).max()))
    steps = n_sims * 100000
    # central differences (float64 forward), all perturbations in one batch
    ref = cs if precision == "fp64" else ss.ControlledSimulator(
        ss.SpinQuantumNumber.ONE, [ss.Control("x", tk), ss.Control("y", tk)], _z_bias)
    h = 1e-4 * sigma
    coords, rows = [], []
    for s_ in rng.choice(n_sims, fd_sims, False):
        for _ in range(fd_coords):
            k, j = int(rng.integers(2)), int(rng.integers(m))
            coords.append((int(s_), k, j))
            for sign in (1, -1):
                a = amp[s_].copy()
                a[k, j] += sign * h
                rows.append(a)
    with torch.no_grad():
        psi_p = ref.final_state(torch.tensor(np.stack(rows), device="cuda"), 0.0, t_end, 20e-9,
                                1e-6, psi0).cpu().numpy()
    loss_p = 1 - np.abs(psi_p[:, -1]) ** 2
    fd = (loss_p[0::2] - loss_p[1::2]) / (2 * h)
    g = grad.cpu().numpy()
    ad = np.array([g[s_, k, j] for (s_, k, j) in coords])
    rel = float(np.abs(fd - ad).max() / np.abs(fd).max())
    cpu = {}
    if cpu_cores:
        cpu = {"cpu_forward_steps_per_s": check * 100000 / cpu_s,
               "cpu_sample": f"float64 forward (time evolution and state product) of the {check} "
                             f"sampled pulses ({check * 100000} CF4 steps) on {cpu_cores} host "
                             "threads; the oracle has no adjoint"}
    return {"forward_backward_ms": elapsed * 1e3, "steps_per_s": 2 * steps / elapsed,
            "fine_steps_per_pass": steps, "simulations": n_sims, "control_samples": 2 * m,
            "grad_rel_err_vs_central_fd": rel, "fd_coordinates": len(coords),
            "final_state_err_vs_oracle_sampled": psi_err, "sampled_simulations": check,
            "precision": precision, **det, **cpu}


def fma_probe(torch, threads=256 * 4 * 8 * 64, iters=8192, reps=3):
    """Sustained float64 FMA rate of the device (TFLOP/s): 8 waves per SIMD,
    8 independent FMA chains per lane (ss_fma_probe), best of `reps`."""
    from spinsim_amd import _native

    out = torch.empty(threads, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    best = 0.0
    for _ in range(reps + 1):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        _native.check(_native.lib().ss_fma_probe(out.data_ptr(), threads, iters, stream.cuda_stream),
                      "ss_fma_probe")
        b.record(stream)
        b.synchronize()
        best = max(best, 64.0 * iters * threads / (a.elapsed_time(b) * 1e-3) / 1e12)
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--lanes", default="auto")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--variants", default=",".join(VARIANTS),
                    help="comma-separated subset of the variants to run beside the headline")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fp32_bulk"),
                    help="arithmetic of the headline run (profiling the other variant)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run one device copy of this many MiB (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); spinsim_amd runs the "
              "integration path on a single device", file=sys.stderr)
        sys.exit(2)
    lanes = args.lanes if args.lanes == "auto" else int(args.lanes)
    chosen = set() if args.no_variants else {v for v in args.variants.split(",") if v}
    unknown = chosen - set(VARIANTS)
    if unknown:
        ap.error(f"unknown variants {sorted(unknown)}; choose from {list(VARIANTS)}")

    import numpy as np
    import torch

    import spinsim_amd as ss
    from oracle.c_oracle import COracle

    torch.cuda.set_device(0)
    main_run, state, psi0 = single_workload(ss, torch, args.precision, args.steps, args.warmup,
                                            lanes)
    n, nf, dt = main_run["coarse"], main_run["n_fine"], main_run["dt"]
    achieved = main_run["fine"] * FLOPS_PER_STEP / (main_run["kernel_ms"] * 1e-3) / 1e12
    executed = main_run["fine"] * FLOPS_PER_STEP_EXECUTED / (main_run["kernel_ms"] * 1e-3) / 1e12
    probe = fma_probe(torch)
    e2e_ms = evaluate_end_to_end(ss, torch, main_run["lanes_per_step"]) \
        if args.precision == "fp64" else None

    host = host_cores()
    cores = host["threads"]
    cpu_cores = None if args.no_cpu_baseline else cores
    variants = {}
    other_states = {}
    if "fp32_bulk" in chosen:
        # the same workload in float32-bulk, and the float32-bulk algorithm (residual
        # series in place of the 24 Lie-Trotter squares) carried out in float64: the
        # ratio of the two is what float32 itself buys, the ratio of the latter to
        # the headline what the algorithm buys
        for prec in ("fp32_bulk", "fp64_series", "fp32"):
            v, other_states[prec], _ = single_workload(ss, torch, prec, args.steps, args.warmup,
                                                       lanes)
            variants[prec] = {k: v[k] for k in ("steps_per_s", "ms_per_step", "kernel_ms",
                                                "lanes_per_step")}
    if "batched" in chosen:
        variants["ensemble_4096_spin_half_same_inputs"] = ensemble_same_inputs(
            ss, torch, max(3, args.steps // 4), cpu_cores)
    if "adjoint" in chosen:
        for prec in ("fp64", "fp32_bulk"):
            variants[f"adjoint_1024_spin_one_{prec}"] = adjoint_workload(
                ss, torch, prec, cpu_cores=cpu_cores if prec == "fp64" else None)
    if "long_noisy" in chosen:
        variants["long_noisy_1e7_spin_half"] = long_noisy_workload(ss, torch,
                                                                   cpu_cores=cpu_cores)
    if "larmor" in chosen:
        variants["larmor_1e6_spin_half"] = larmor_workload(ss, torch, args.steps, args.warmup,
                                                           cpu_cores)

    if args.copy_probe:
        src = torch.empty(args.copy_probe << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    # Accuracy against the float64 oracle (C restatement) on the same inputs.
    co = COracle(RABI_ONE, 3)
    base = None
    if args.no_cpu_baseline:
        _, u_ref = co.time_evolution([0.0], T0, dt, DT_OUT, nf, n, threads=cores)
        ref_state = co.get_state(psi0, u_ref)
    else:
        rate, ref_state = cpu_baseline(co, psi0, dt, n, nf, cores)
        base = {"value": rate, "unit": "steps/s", "cores": cores, "kind": "port",
                "sample": f"whole workload ({n * nf} CF4 steps: time evolution, state "
                          f"product, spin) on {cores} host threads, median of 3",
                "host": host}
    base1 = None
    if not args.no_cpu_baseline:
        rows = min(n, 10000)
        base1 = {"value": cpu_single_core(co, dt, nf, rows), "unit": "steps/s", "cores": 1,
                 "kind": "port", "sample": f"first {rows} coarse steps ({rows * nf} CF4 steps, "
                                           "time evolution only) on one host thread, median of 3"}
    state_err = float(np.abs(state - ref_state).max())
    for prec, st in other_states.items():
        variants[prec]["state_err_vs_oracle"] = float(np.abs(st - ref_state).max())
    if "fp32_bulk" in variants:
        variants["fp32_bulk"]["speedup_over_fp64"] = (variants["fp32_bulk"]["steps_per_s"]
                                                      / main_run["steps_per_s"])
        variants["fp32_bulk"]["speedup_over_fp64_series"] = (
            variants["fp32_bulk"]["steps_per_s"] / variants["fp64_series"]["steps_per_s"])
        variants["fp64_series"]["speedup_over_fp64"] = (variants["fp64_series"]["steps_per_s"]
                                                        / main_run["steps_per_s"])

    line = {
        "metric": METRIC, "value": main_run["steps_per_s"], "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 residuals / f64",
        "data": "synthetic",
        "config": {"workload": "spin-one Rabi drive, Lie-Trotter (24 squares), rotating frame, "
                               + ("plain float64" if args.precision == "fp64" else "float32-bulk")
                               + ", t 0-100 ms, dt 100 ns, output 1 us",
                   "coarse_steps": n, "fine_steps": n * nf,
                   "lanes_per_step": main_run["lanes_per_step"], "parallelism": "single GPU"},
        "autotune_ms": main_run["autotune_ms"],
        "state_err_vs_oracle": state_err,
        "time_evolution_kernel_ms": main_run["kernel_ms"],
        # float64 flops the kernel executes per step (its residual squares derive
        # one row), over the time-evolution kernel's event-timed launches
        "roofline": {"bound": "valu", "achieved": executed, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": executed / FP64_PEAK_TFLOPS,
                     "traffic": TRAFFIC_BYTES, "flops_per_step": FLOPS_PER_STEP_EXECUTED},
        # the same time priced at the reference's own flop count per step
        "roofline_effective": {"flops_per_step": FLOPS_PER_STEP, "tflops": achieved,
                               "frac_of_peak": achieved / FP64_PEAK_TFLOPS},
        "fp64_fma_probe": {"tflops": probe, "frac_of_published_peak": probe / FP64_PEAK_TFLOPS,
                           "kernel_frac_of_probe": executed / probe},
        "evaluate_numpy_ms": e2e_ms,
        "cpu_baseline": base,
        "cpu_baseline_1core": base1,
        "variants": variants,
    }
    ok = state_err <= 1e-6
    if not ok:  # the metric is quoted at state error <= 1e-6 against the oracle
        line["value"] = None
        line["value_withheld"] = f"state error {state_err:.3e} against the oracle exceeds 1e-6"
    print(json.dumps(line), flush=True)
    if not ok:
        sys.exit(1)


if __name__ == "__main__":
    main()
