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


def _su2(rng, n):
    """Random special unitaries [[a, -conj b], [b, conj a]]."""
    v = rng.normal(size=(n, 4))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    a, b = v[:, 0] + 1j * v[:, 1], v[:, 2] + 1j * v[:, 3]
    u = np.empty((n, 2, 2), complex)
    u[:, 0, 0], u[:, 1, 0], u[:, 0, 1], u[:, 1, 1] = a, b, -b.conj(), a.conj()
    return u


@pytest.mark.parametrize("chunk_rows", [0, 1024, 4096])
def test_su2_scan_layouts_agree(chunk_rows):
    """ss_batch_state_scan on special unitaries stored full (SS_U_FULL, the general
    products; SS_U_SU2_FULL, first columns) and as first columns
    (SS_U_SU2_COLUMN): states and Bloch vectors agree to 1e-12 and match the
    sequential product; uneven simulations, empty ones, starts inside tiles."""
    import torch

    from spinsim_amd import _native

    print(f"case: SU(2) scan layouts, chunk_rows={chunk_rows}", flush=True)
    rng = np.random.default_rng(79)
    lengths = np.concatenate([rng.integers(0, 9000, 40), [0, 1, 2, 1023, 1024, 1025, 20000]])
    rng.shuffle(lengths)
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    total, b = int(offsets[-1]), len(lengths)
    u = _su2(rng, total)
    psi = rng.normal(size=(b, 2)) + 1j * rng.normal(size=(b, 2))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    L = _native.lib()
    doff = torch.as_tensor(offsets, device="cuda")
    dpsi = torch.as_tensor(psi, device="cuda")
    ws = torch.empty(max(int(L.ss_batch_scan_workspace(total, 2, chunk_rows)), 1), dtype=torch.uint8,
                     device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    outs = {}
    for layout, arr in ((_native.U_FULL, u), (_native.U_SU2_FULL, u),
                        (_native.U_SU2_COLUMN, np.ascontiguousarray(u[:, :, 0]))):
        du = torch.as_tensor(arr, device="cuda")
        st = torch.full((total, 2), float("nan"), dtype=torch.complex128, device="cuda")
        sp = torch.empty((total, 3), dtype=torch.float64, device="cuda")
        _native.check(L.ss_batch_state_scan(dpsi.data_ptr(), du.data_ptr(), st.data_ptr(), sp.data_ptr(),
                                            doff.data_ptr(), b, total, 2, chunk_rows, layout,
                                            ws.data_ptr(), ws.numel(), stream), "scan")
        outs[layout] = (st.cpu().numpy(), sp.cpu().numpy())
    ref_st, ref_sp = outs[_native.U_FULL]
    assert np.isfinite(ref_st).all()
    for layout in (_native.U_SU2_FULL, _native.U_SU2_COLUMN):
        st, sp = outs[layout]
        d_st, d_sp = float(np.abs(st - ref_st).max()), float(np.abs(sp - ref_sp).max())
        print(f"  layout {layout} vs full products: state {d_st:.2e}, spin {d_sp:.2e}")
        assert d_st <= 1e-12 and d_sp <= 1e-12
    # sequential product on a few simulations
    for s in rng.choice(b, 6, replace=False):
        a, e = offsets[s], offsets[s + 1]
        if a == e:
            continue
        v = psi[s].copy()
        seq = [v]
        for i in range(a + 1, e):
            v = u[i - 1] @ v
            seq.append(v)
        assert np.abs(outs[_native.U_SU2_COLUMN][0][a:e] - np.array(seq)).max() <= 1e-9
