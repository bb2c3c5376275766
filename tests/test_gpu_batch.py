"""Batched persistent launch vs the oracle, one simulation at a time."""
from __future__ import annotations

import math

import numpy as np
import pytest

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu
TOL = 1e-9


def _oracle_sim(co, sweep, t0, t1, dt, dt_out, psi0, sim_row=0):
    dt2, n, nf, _ = O.timing(t0, t1, dt, dt_out)
    t, u = co.time_evolution(sweep, float(t0), dt2, float(dt_out), nf, n, sim=sim_row)
    st = co.get_state(psi0, u)
    return t, u, st


@pytest.mark.parametrize("field", [F.RABI_SWEEP_HALF, F.RABI_SWEEP_ONE], ids=lambda f: f.name)
@pytest.mark.parametrize("lanes", [1, 4, 10])
def test_uneven_batch_matches_oracle(field, lanes):
    import spinsim_amd as ss

    print(f"case: batch {field.name} lanes={lanes}", flush=True)
    rng = np.random.default_rng(21)
    b = 37
    spin = ss.SpinQuantumNumber.HALF if field.spin == "half" else ss.SpinQuantumNumber.ONE
    base = np.array(field.sweep)
    sweep = np.tile(base, (b, 1))
    sweep[:, 1] *= rng.uniform(0.5, 2.0, b)
    sweep[:, 2] += rng.normal(0, 2 * math.pi * 200, b)
    t0 = rng.uniform(0, 1e-5, b)
    dur = rng.uniform(0, 3e-4, b)
    dur[3] = 0.0          # empty simulation
    dur[4] = 0.7e-6       # shorter than one output step
    t1 = t0 + dur
    dt = np.where(rng.uniform(size=b) < 0.5, 5e-8, 1e-7)
    dt_out = 1e-6
    psi = rng.normal(size=(b, spin.dimension)) + 1j * rng.normal(size=(b, spin.dimension))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    sim = ss.Simulator(field.src, spin)
    res = sim.evaluate_batch(t0, t1, dt, dt_out, psi, sweep, lanes_per_step=lanes)
    co = COracle(field.src, spin.dimension)
    spin_all = res.spin.cpu().numpy()
    for s in range(b):
        r = res[s]
        t, u, st = _oracle_sim(co, sweep[s], t0[s], t1[s], dt[s], dt_out, psi[s])
        assert r.time.shape == t.shape
        assert np.array_equal(r.time, t)
        assert np.abs(r.time_evolution - u).max(initial=0) <= TOL
        assert np.abs(r.state - st).max(initial=0) <= TOL
        a, e = res.offsets[s], res.offsets[s + 1]
        assert np.abs(spin_all[a:e] - co.get_spin(st)).max(initial=0) <= TOL


def test_batch_sampler_rows():
    """A per-simulation (B, M) amplitude table: simulation s reads row s."""
    import spinsim_amd as ss

    print("case: batch sampler rows", flush=True)
    rng = np.random.default_rng(8)
    b, m = 6, 120
    tk = np.linspace(0, 2e-4, m)
    cx = rng.normal(0, 2 * math.pi * 5e3, (b, m))
    ctl = ss.generate_interpolation_sampler(tk, cx, per_simulation=True)

    def get_field(time_sample, sweep_parameters, field_sample):
        field_sample[0] = ctl(time_sample)
        field_sample[1] = 0.5 * ctl(time_sample + 1e-6)
        field_sample[2] = math.tau * 1e5
        field_sample[3] = 0

    sim = ss.Simulator(get_field, ss.SpinQuantumNumber.ONE)
    psi0 = ss.SpinQuantumNumber.ONE.plus_z
    res = sim.evaluate_batch(0.0, 2e-4, 2e-8, 1e-6, psi0, np.zeros((b, 1)))
    src, _, _ = ss.transpile(get_field)
    co = COracle(src, 3, sampler_names=("ctl_table",))
    co.set_samplers([(tk, cx, 3, True)])
    for s in range(b):
        t, u, st = _oracle_sim(co, [0.0], 0.0, 2e-4, 2e-8, 1e-6, psi0, sim_row=s)
        r = res[s]
        assert np.abs(r.time_evolution - u).max() <= TOL
        assert np.abs(r.state - st).max() <= TOL


def test_batch_graph_replays_match_eager():
    import torch

    import spinsim_amd as ss

    print("case: batch graph replay", flush=True)
    rng = np.random.default_rng(5)
    b = 16
    sweep = np.tile(np.array(F.RABI_SWEEP_HALF.sweep), (b, 1))
    sweep[:, 2] += rng.normal(0, 500, b)
    sim = ss.Simulator(F.RABI_SWEEP_HALF.src, ss.SpinQuantumNumber.HALF)
    t1 = rng.uniform(1e-4, 4e-4, b)
    plan = ss.BatchPlan(sim, 0.0, t1, 5e-8, 1e-6, ss.SpinQuantumNumber.HALF.plus_z, sweep)
    plan.run()
    torch.cuda.synchronize()
    eager_u = plan.u.clone()
    eager_state = plan.state.clone()
    plan.capture()
    for k in range(4):
        plan.u.zero_()
        plan.state.zero_()
        plan.replay()
        torch.cuda.synchronize()
        assert torch.equal(plan.u, eager_u), k
        assert torch.equal(plan.state, eager_state), k
        # unrelated eager work between replays
        x = torch.randn(1000, device="cuda")
        _ = (x * 2).sum().item()


RANDOM_FIELD_ONE = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = s[0] * cos(s[1] * t + s[2]); f[1] = s[3] * sin(s[4] * t);
  f[2] = s[5] + s[6] * cos(s[7] * t); f[3] = s[8] * cos(s[9] * t); }
"""
RANDOM_FIELD_HALF = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = s[0] * cos(s[1] * t + s[2]); f[1] = s[3] * sin(s[4] * t);
  f[2] = s[5] + s[6] * cos(s[7] * t); }
"""


@pytest.mark.parametrize("k", range(12))
def test_random_batch_matches_oracle(k):
    """Seeded random batches: per-simulation start, duration, integration and
    output steps and sweep rows; random lanes per step and rows per grab;
    float64 to 1e-9, float32-bulk to 1e-6 on the states."""
    import spinsim_amd as ss

    rng = np.random.default_rng(500 + k)
    spin = "one" if k % 2 else "half"
    method = ("magnus_cf4", "euler", "heun")[k % 3]
    precision = "fp32_bulk" if k % 4 == 3 else "fp64"
    rotating = bool(rng.integers(2)) if precision == "fp64" else True
    b = int(rng.integers(1, 70))
    tau = 2 * math.pi
    amp = tau * 10 ** rng.uniform(3, 4.5, b)
    sweep = np.stack([amp, tau * 10 ** rng.uniform(3, 5.5, b), rng.uniform(0, tau, b),
                      amp * rng.uniform(0, 1, b), tau * 10 ** rng.uniform(3, 5, b),
                      tau * 10 ** rng.uniform(4, 5.5, b), amp * rng.uniform(0, 0.3, b),
                      tau * 10 ** rng.uniform(2, 4, b), amp * rng.uniform(0, 0.5, b),
                      tau * 10 ** rng.uniform(3, 5, b)], 1)
    t0 = rng.uniform(-5e-5, 5e-5, b)
    dt = 10 ** rng.uniform(-8, -6.8, b)
    dt_out = dt * rng.integers(1, 30, b)
    t1 = t0 + dt_out * rng.integers(0, 400, b) + dt_out * rng.uniform(0, 0.99, b)
    d = 2 if spin == "half" else 3
    psi = rng.normal(size=(b, d)) + 1j * rng.normal(size=(b, d))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    lanes = int(rng.choice([1, 2, 3, 4, 8, 16, 32]))
    items = int(rng.choice([1, 2, 4, 16]))
    print(f"case {k}: spin={spin} {method} {precision} rotating={rotating} B={b} "
          f"lanes={lanes} items={items}", flush=True)
    src = RANDOM_FIELD_HALF if spin == "half" else RANDOM_FIELD_ONE
    sq = ss.SpinQuantumNumber.HALF if spin == "half" else ss.SpinQuantumNumber.ONE
    methods = {"magnus_cf4": "MAGNUS_CF4", "euler": "MIDPOINT_SAMPLE", "heun": "HALF_STEP"}
    sim = ss.Simulator(src, sq, use_rotating_frame=rotating, precision=precision,
                       integration_method=getattr(ss.IntegrationMethod, methods[method]))
    res = sim.evaluate_batch(t0, t1, dt, dt_out, psi, sweep, lanes_per_step=lanes,
                             items_per_grab=items)
    co = COracle(src, d, method, None, rotating)
    tol = TOL if precision == "fp64" else 1e-6
    spin_all = res.spin.cpu().numpy()
    worst = 0.0
    for s in range(b):
        r = res[s]
        t, u, st = _oracle_sim(co, sweep[s], t0[s], t1[s], dt[s], dt_out[s], psi[s])
        assert np.array_equal(r.time, t), s
        err = np.abs(r.state - st).max(initial=0)
        worst = max(worst, err)
        assert err <= tol, s
        if precision == "fp64":
            assert np.abs(r.time_evolution - u).max(initial=0) <= TOL, s
            a, e = res.offsets[s], res.offsets[s + 1]
            assert np.abs(spin_all[a:e] - co.get_spin(st)).max(initial=0) <= TOL, s
    print(f"  max state error {worst:.2e} over {b} simulations")
