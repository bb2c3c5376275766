"""ss_time_evolution_events: the next time-evolution launch on the thread
stamps its kernel's start and end into the caller's HIP events (the
hipExtModuleLaunchKernel path), computes exactly what the plain launch
computes, and leaves later launches untouched."""
from __future__ import annotations

import pytest

import fields as F

pytestmark = pytest.mark.gpu


def test_events_stamp_one_launch_and_results_match():
    import torch

    import bench
    import spinsim_amd as ss
    from spinsim_amd import _native

    print("case: time-evolution launch with kernel timestamps", flush=True)
    sim = ss.Simulator(F.RABI_ONE.src, ss.SpinQuantumNumber.ONE)
    dev = torch.device("cuda", 0)
    n, dt, dt_out = 20000, 1e-7, 1e-6
    sweep = torch.zeros(1, dtype=torch.float64, device=dev)
    outs = []
    for lanes in (0, 4):
        for with_events in (False, True):
            t = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
            u = torch.full((n, 3, 3), float("nan"), dtype=torch.complex128, device=dev)
            ev = (bench.TimingEvent(), bench.TimingEvent())
            if with_events:
                _native.check(_native.lib().ss_time_evolution_events(ev[0].handle, ev[1].handle),
                              "ss_time_evolution_events")
            sim.launch_time_evolution(sweep, 0.0, dt, dt_out, t, u, lanes)
            torch.cuda.synchronize()
            if with_events:
                ms = ev[0].elapsed_time(ev[1])
                print(f"  lanes={lanes}: kernel {ms:.4f} ms")
                assert 0.0 < ms < 1000.0
            outs.append((t.clone(), u.clone()))
        # the events applied to one launch only: this one runs the plain path
        t2 = torch.empty_like(t)
        u2 = torch.empty_like(u)
        sim.launch_time_evolution(sweep, 0.0, dt, dt_out, t2, u2, lanes)
        torch.cuda.synchronize()
        assert torch.equal(t2, outs[-1][0]) and torch.equal(u2, outs[-1][1])
    for k in (0, 2):
        assert torch.equal(outs[k][0], outs[k + 1][0]) and torch.equal(outs[k][1], outs[k + 1][1])
        assert torch.isfinite(outs[k][1].real).all()


def test_events_on_the_batch_launch():
    import numpy as np
    import torch

    import bench
    import spinsim_amd as ss
    from spinsim_amd import _native

    print("case: batch time-evolution launch with kernel timestamps", flush=True)
    sim = ss.Simulator(F.RABI_SWEEP_HALF.src, ss.SpinQuantumNumber.HALF)
    b = 64
    sweep = np.tile(np.array(F.RABI_SWEEP_HALF.sweep), (b, 1))
    dur = np.linspace(1e-5, 2e-4, b)
    plan = ss.BatchPlan(sim, 0.0, dur, 5e-8, 1e-6, ss.SpinQuantumNumber.HALF.plus_z, sweep)
    plan.run_time_evolution()
    torch.cuda.synchronize()
    ref = plan.u.clone()
    ev = (bench.TimingEvent(), bench.TimingEvent())
    _native.check(_native.lib().ss_time_evolution_events(ev[0].handle, ev[1].handle), "events")
    plan.run_time_evolution()
    torch.cuda.synchronize()
    assert ev[0].elapsed_time(ev[1]) > 0.0
    assert torch.equal(plan.u, ref)
