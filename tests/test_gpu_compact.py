"""Spin-half batches with operators stored as first columns (SS_U_SU2_COLUMN):
states, Bloch vectors and times against the full-operator batch of the same
inputs and against the oracle; the expanded operators against both; graph
replays against the eager run."""
from __future__ import annotations

import math

import numpy as np
import pytest

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu


def _batch(rng, b, max_dur):
    sweep = np.tile(np.array(F.RABI_SWEEP_HALF.sweep), (b, 1))
    sweep[:, 1] *= rng.uniform(0.5, 2.0, b)
    sweep[:, 2] += rng.normal(0, 2 * math.pi * 200, b)
    t0 = rng.uniform(0, 1e-5, b)
    dur = rng.uniform(0, max_dur, b)
    dur[1] = 0.0       # an empty simulation
    dur[2] = 0.4e-6    # shorter than one output step
    psi = rng.normal(size=(b, 2)) + 1j * rng.normal(size=(b, 2))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    return sweep, t0, t0 + dur, psi


@pytest.mark.parametrize("precision,lanes,tol_full,tol_oracle", [
    ("fp64", 1, 1e-12, 1e-9),
    ("fp64", 8, 1e-12, 1e-9),
    ("fp32_bulk", 4, 1e-7, 1e-6),
    ("fp32", 4, 1e-7, 1e-6),
], ids=["fp64_l1", "fp64_l8", "fp32_bulk", "fp32"])
def test_compact_matches_full_and_oracle(precision, lanes, tol_full, tol_oracle):
    import torch

    import spinsim_amd as ss

    print(f"case: compact batch {precision} lanes={lanes}", flush=True)
    rng = np.random.default_rng(77)
    b = 29
    sweep, t0, t1, psi = _batch(rng, b, 3e-4)
    sim = ss.Simulator(F.RABI_SWEEP_HALF.src, ss.SpinQuantumNumber.HALF, precision=precision)
    args = (t0, t1, 5e-8, 1e-6, psi, sweep)
    full = ss.BatchPlan(sim, *args, lanes_per_step=lanes, with_spin=True).run()
    comp = ss.BatchPlan(sim, *args, lanes_per_step=lanes, with_spin=True, compact=True).run()
    torch.cuda.synchronize()
    assert comp.u.shape == (comp.total, 2)
    assert torch.equal(comp.time, full.time)
    # the stored column is the full operator's first column
    assert torch.equal(comp.u, full.u[:, :, 0])
    rf, rc = full.results(), comp.results()
    st_diff = float((rc.state - rf.state).abs().max())
    spin_diff = float((rc.spin - rf.spin).abs().max())
    u_diff = float((rc.time_evolution - rf.time_evolution).abs().max())
    print(f"  compact - full: state {st_diff:.3e}, spin {spin_diff:.3e}, operators {u_diff:.3e}")
    assert st_diff <= tol_full and spin_diff <= 2 * tol_full and u_diff <= tol_full
    co = COracle(F.RABI_SWEEP_HALF.src, 2)
    worst = 0.0
    for s in range(b):
        dt2, n, nf, _ = O.timing(t0[s], t1[s], 5e-8, 1e-6)
        _, u = co.time_evolution(sweep[s], float(t0[s]), dt2, 1e-6, nf, n)
        st = co.get_state(psi[s], u)
        r = rc[s]
        assert r.state.shape == st.shape and r.time_evolution.shape == u.shape
        worst = max(worst, float(np.abs(r.state - st).max(initial=0)),
                    float(np.abs(r.time_evolution - u).max(initial=0)))
    print(f"  compact vs oracle: {worst:.3e}")
    assert worst <= tol_oracle


def test_compact_graph_replays_match_eager():
    import torch

    import spinsim_amd as ss

    print("case: compact batch graph replay", flush=True)
    rng = np.random.default_rng(78)
    sweep, t0, t1, psi = _batch(rng, 16, 4e-4)
    sim = ss.Simulator(F.RABI_SWEEP_HALF.src, ss.SpinQuantumNumber.HALF)
    plan = ss.BatchPlan(sim, t0, t1, 5e-8, 1e-6, psi, sweep, with_spin=True, compact=True)
    plan.run()
    torch.cuda.synchronize()
    eager = [x.clone() for x in (plan.u, plan.state, plan.spin_out)]
    plan.capture()
    for k in range(3):
        for x in (plan.u, plan.state, plan.spin_out):
            x.zero_()
        plan.replay()
        torch.cuda.synchronize()
        for x, e in zip((plan.u, plan.state, plan.spin_out), eager):
            assert torch.equal(x, e), k
        y = torch.randn(1000, device="cuda")
        _ = (y * 2).sum().item()


def test_compact_rejected_for_spin_one():
    import spinsim_amd as ss

    sim = ss.Simulator(F.RABI_SWEEP_ONE.src, ss.SpinQuantumNumber.ONE)
    with pytest.raises(ValueError, match="spin half"):
        ss.BatchPlan(sim, 0.0, 1e-5, 5e-8, 1e-6, ss.SpinQuantumNumber.ONE.plus_z,
                     F.RABI_SWEEP_ONE.sweep, compact=True)
