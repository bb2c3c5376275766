"""The autotuner's choices are valid launch shapes and the tuned launches keep
parity with the oracle (lanes per step, rows per grab, scan chunk size)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu
TOL = 1e-9


def test_tune_lanes_single_simulation():
    import torch

    import spinsim_amd as ss
    from spinsim_amd.autotune import PHASED_CANDIDATES, lane_candidates, tune_lanes

    print("case: tune_lanes rabi_one 2 ms", flush=True)
    sim = ss.Simulator(F.RABI_ONE.src, ss.SpinQuantumNumber.ONE)
    dt, n, nf, _ = O.timing(0.0, 2e-3, 100e-9, 1e-6)
    sweep = torch.zeros(1, dtype=torch.float64, device="cuda")
    lanes, table = tune_lanes(sim, n, nf, dt, 1e-6, sweep, 0.0, reps=1, use_cache=False)
    assert set(table) == {*PHASED_CANDIDATES, *lane_candidates(nf)}
    assert lanes in table and all(t > 0 for t in table.values())
    res = sim.evaluate(0.0, 2e-3, 100e-9, 1e-6, O.EIGENSTATES["one"]["plus_z"],
                       lanes_per_step=lanes)
    co = COracle(F.RABI_ONE.src, 3)
    _, u = co.time_evolution([0.0], 0.0, dt, 1e-6, nf, n)
    assert np.abs(res.state - co.get_state(O.EIGENSTATES["one"]["plus_z"], u)).max() <= TOL


def test_tune_batch_with_scan_chunks():
    import torch

    import spinsim_amd as ss
    from spinsim_amd.autotune import scan_chunk_candidates, tune_batch

    print("case: tune_batch + scan chunks", flush=True)
    rng = np.random.default_rng(31)
    b = 48
    sweep = np.tile(np.array(F.RABI_SWEEP_HALF.sweep), (b, 1))
    sweep[:, 1] *= rng.uniform(0.5, 2.0, b)
    sweep[:, 2] += rng.normal(0, 2 * math.pi * 200, b)
    t1 = rng.uniform(0, 4e-3, b)
    t1[5] = 0.0
    psi0 = O.EIGENSTATES["half"]["plus_z"]
    sim = ss.Simulator(F.RABI_SWEEP_HALF.src, ss.SpinQuantumNumber.HALF)
    plan = ss.BatchPlan(sim, 0.0, t1, 50e-9, 1e-6, psi0, sweep, with_spin=True)
    (lanes, items), table = tune_batch(plan, reps=1, warmup=0, use_cache=False, lanes=(1, 2),
                                       items=(1, 4))
    assert set(table) == {(1, 1), (1, 4), (2, 1), (2, 4)}
    assert (plan.lanes, plan.items_per_grab) == (lanes, items)
    assert plan.scan_chunk_rows in scan_chunk_candidates(plan.total)
    assert plan.scan_ws.numel() >= plan.scan_workspace_bytes(plan.scan_chunk_rows)
    co = COracle(F.RABI_SWEEP_HALF.src, 2)
    refs = {}
    for s in (0, 5, 17, b - 1):
        dt, n, nf, _ = O.timing(0.0, t1[s], 50e-9, 1e-6)
        _, u = co.time_evolution(sweep[s], 0.0, dt, 1e-6, nf, n)
        refs[s] = co.get_state(psi0, u)

    def check():
        st = plan.state.cpu().numpy()
        for s, ref in refs.items():
            a, e = int(plan.offsets_host[s]), int(plan.offsets_host[s + 1])
            assert e - a == ref.shape[0]
            assert np.abs(st[a:e] - ref).max(initial=0) <= TOL, s

    plan.run()
    torch.cuda.synchronize()
    check()
    plan.capture()
    for _ in range(3):
        plan.state.zero_()
        plan.replay()
        torch.cuda.synchronize()
        check()
