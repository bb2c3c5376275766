"""Multi-axis interpolation samplers on the GPU: ss_interp_eval against the
oracle's one-axis samplers, and simulations whose field reads a multi-axis
table -- one simulation, and a batch with one (A, M) table per simulation --
against the oracle (same translated source)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu
TOL = 1e-9


def _close(got, want, tol):
    np.testing.assert_allclose(got, want, rtol=0, atol=tol, equal_nan=True)


def test_interp_eval_multi_axis():
    import spinsim_amd as ss

    print("case: multi-axis ss_interp_eval", flush=True)
    rng = np.random.default_rng(61)
    tn = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0, 1, 30)]))
    tn[5] = tn[6]
    for knots in (np.linspace(0.0, 1.0, 41), tn):
        tt = np.concatenate([rng.uniform(-0.2, 1.2, 2000), knots])
        for order in (0, 1, 3):
            for zb in (True, False):
                amp = rng.normal(size=(3, knots.size))
                got = ss.generate_interpolation_sampler(knots, amp, order, zb)(tt)
                rows = rng.normal(size=(2, 4, knots.size))
                sb = ss.generate_interpolation_sampler(knots, rows, order, zb, per_simulation=True)
                for a in range(3):
                    _close(got[:, a], O.InterpolationSampler(knots, amp[a], order, zb)(tt), 1e-12)
                for b in range(2):
                    g = sb(tt, row=b)
                    for a in range(4):
                        want = O.InterpolationSampler(knots, rows[b, a], order, zb)(tt)
                        _close(g[:, a], want, 1e-12)


def _inputs(rng, rows=None):
    knots = np.linspace(0.0, 2e-4, 81)
    shape = (3, knots.size) if rows is None else (rows, 3, knots.size)
    return knots, rng.normal(0, math.tau * 2e4, shape)


def _field(ctl):
    def get_field(time_sample, sweep_parameters, field_sample):
        v = ctl(time_sample)
        field_sample[0] = v[0]
        field_sample[1] = v[1]
        field_sample[2] = math.tau * 1e5 + 0.1 * v[2]
        field_sample[3] = math.tau * 2e3
    return get_field


@pytest.mark.parametrize("rotating", [True, False])
def test_multi_axis_field_matches_oracle(rotating):
    import spinsim_amd as ss

    print(f"case: multi-axis field, one simulation, rotating={rotating}", flush=True)
    rng = np.random.default_rng(67)
    knots, amp = _inputs(rng)
    ctl = ss.generate_interpolation_sampler(knots, amp)
    field = _field(ctl)
    psi0 = O.EIGENSTATES["one"]["plus_z"]
    sim = ss.Simulator(field, ss.SpinQuantumNumber.ONE, use_rotating_frame=rotating)
    res = sim.evaluate(0.0, 2e-4, 2e-8, 1e-6, psi0)
    src, _, _ = ss.transpile(field)
    co = COracle(src, 3, rotating=rotating, sampler_names=("unused",))
    co.set_samplers([(knots, amp, 3, True, True)])
    dt, n, nf, _ = O.timing(0.0, 2e-4, 2e-8, 1e-6)
    t, u = co.time_evolution([0.0], 0.0, dt, 1e-6, nf, n)
    st = co.get_state(psi0, u)
    assert np.array_equal(res.time, t)
    assert np.abs(res.time_evolution - u).max() <= TOL
    assert np.abs(res.state - st).max() <= TOL
    assert np.abs(res.state - psi0).max() > 1e-2


def test_multi_axis_rows_batch_matches_oracle():
    """A batch whose simulation s reads its own (A, M) table, row s of (B, A, M)."""
    import spinsim_amd as ss

    print("case: multi-axis field, per-simulation rows, batch", flush=True)
    rng = np.random.default_rng(71)
    b = 5
    knots, amp = _inputs(rng, rows=b)
    ctl = ss.generate_interpolation_sampler(knots, amp, per_simulation=True)
    field = _field(ctl)
    psi0 = O.EIGENSTATES["one"]["plus_z"]
    sim = ss.Simulator(field, ss.SpinQuantumNumber.ONE)
    dur = rng.uniform(5e-5, 2e-4, b)
    res = sim.evaluate_batch(0.0, dur, 2e-8, 1e-6, psi0, np.zeros((b, 1)))
    src, _, _ = ss.transpile(field)
    co = COracle(src, 3, sampler_names=("unused",))
    co.set_samplers([(knots, amp, 3, True, True)])
    for s in range(b):
        dt, n, nf, _ = O.timing(0.0, dur[s], 2e-8, 1e-6)
        t, u = co.time_evolution([0.0], 0.0, dt, 1e-6, nf, n, sim=s)
        st = co.get_state(psi0, u)
        r = res[s]
        assert np.abs(r.time_evolution - u).max(initial=0) <= TOL
        assert np.abs(r.state - st).max(initial=0) <= TOL
