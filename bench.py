"""Benchmark: Magnus-CF4 steps/s of spinsim_amd's integration path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

A step is one evaluate() pass of the spin-one Rabi workload (x = 2pi*2 kHz
cos(2pi*700 kHz t), z = 2pi*700 kHz, quadratic shift 2pi*1 kHz; t = 0 -> 100 ms,
integration step 100 ns, output step 1 us: 1e5 coarse x 10 fine = 1e6 CF4
steps; Lie-Trotter exponentials, rotating frame, float64) over inputs already
resident in HBM: the time-evolution kernel, the state prefix product and the
Bloch vectors.  Prints one JSON line.
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

METRIC = "Magnus-CF4 steps/sec on 1 GPU at state err ≤1e-6 vs FP64 oracle; batched/plain"

RABI_ONE = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = SS_TAU * 2e3 * cos(SS_TAU * 7e5 * t); f[1] = 0; f[2] = SS_TAU * 7e5; f[3] = SS_TAU * 1e3; }
"""
T0, T1, DT, DT_OUT = 0.0, 100e-3, 100e-9, 1e-6

# Algorithmic float64 flops per spin-one Lie-Trotter CF4 fine step, counted from
# the reference's expressions (1 flop per real +,-,*,/,sqrt; complex * = 6,
# complex + = 2; transcendentals excluded): 2 exponentials x (24 residual
# squarings x 207 + 143 set-up) + 2 matrix products x 198 + frame 24 + CF4
# combinations 32 + sampling 14 + time update 1.
FLOPS_PER_STEP = 2 * (24 * 207 + 143) + 2 * 198 + 24 + 32 + 14 + 1
FP64_PEAK_TFLOPS = 78.6


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--lanes", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the integration path "
              "is data-parallel over coarse steps and simulations and is benchmarked on one "
              "device", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import spinsim_amd as ss
    from spinsim_amd.simulator import derive_steps

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    sim = ss.Simulator(RABI_ONE, ss.SpinQuantumNumber.ONE)
    dt, n, _ = derive_steps(T0, T1, DT, DT_OUT)
    nf = math.floor(DT_OUT / dt + 0.5)
    fine_steps = n * nf
    psi0 = ss.SpinQuantumNumber.ONE.plus_z

    sweep = torch.zeros(1, dtype=torch.float64, device=dev)
    time_out = torch.empty(n, dtype=torch.float64, device=dev)
    u = torch.empty((n, 3, 3), dtype=torch.complex128, device=dev)
    state = torch.empty((n, 3), dtype=torch.complex128, device=dev)
    spin = torch.empty((n, 3), dtype=torch.float64, device=dev)
    ws = sim.scan_workspace(n, dev)
    stream = torch.cuda.current_stream(dev)

    def step(events=None):
        if events is not None:
            events[0].record(stream)
        sim.launch_time_evolution(sweep, T0, dt, DT_OUT, time_out, u, args.lanes)
        if events is not None:
            events[1].record(stream)
        sim.launch_state_scan(psi0, u, state, ws)
        sim.launch_spin(state, spin)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t_start = time.perf_counter()
    for k in range(args.steps):
        step(ev[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    kernel_ms = statistics.mean(a.elapsed_time(b) for a, b in ev)

    ms_per_step = 1e3 * elapsed / args.steps
    value = fine_steps / (elapsed / args.steps)
    achieved = fine_steps * FLOPS_PER_STEP / (kernel_ms * 1e-3) / 1e12

    # Accuracy against the float64 oracle (C restatement) on the same inputs.
    from oracle.c_oracle import COracle
    co = COracle(RABI_ONE, 3)
    cores = min(len(os.sched_getaffinity(0)), 16)
    base = None
    if args.no_cpu_baseline:
        _, u_ref = co.time_evolution([0.0], T0, dt, DT_OUT, nf, n, threads=cores)
        ref_state = co.get_state(psi0, u_ref)
    else:
        rate, ref_state = cpu_baseline(co, psi0, dt, n, nf, cores)
        base = {"value": rate, "unit": "steps/s", "cores": cores, "kind": "port",
                "sample": f"whole workload ({fine_steps} CF4 steps: time evolution, state "
                          f"product, spin) on {cores} host threads, median of 3"}
    state_err = float(np.abs(state.cpu().numpy() - ref_state).max())

    line = {
        "metric": METRIC, "value": value, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "spin-one Rabi drive, Lie-Trotter (24 squares), rotating frame, "
                               "plain float64, t 0-100 ms, dt 100 ns, output 1 us",
                   "coarse_steps": n, "fine_steps": fine_steps, "parallelism": "single GPU"},
        "state_err_vs_oracle": state_err,
        "time_evolution_kernel_ms": kernel_ms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None},
        "cpu_baseline": base,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
