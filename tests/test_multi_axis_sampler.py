"""Multi-axis interpolation samplers on the CPU: the oracle's restatement
against one one-axis sampler per axis, the host build of the product's
sampler against the oracle, the translation of Python field functions that
read them, and a CPU-device run of such a field against the oracle.

The reference raises for 2-D amplitudes ("Try generating a single
interpolator for each axis instead", spinsim/__init__.py:2062-2063); an
(A, M) amplitude here is A components on one time grid, component a equal to
the one-axis sampler of amplitude[a]."""
from __future__ import annotations

import math

import numpy as np
import pytest

import spinsim_amd as ss
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

from test_cpu_device import assert_close

TRIVIAL = """
__device__ void get_field(double t, const double* s, double* f, SS_CONTEXT)
{ f[0] = 0; f[1] = 0; f[2] = 1; }
"""


def _close(got, want, tol):
    """Within tol, NaN where the reference's own 0/0 gives NaN (t on a repeated knot)."""
    np.testing.assert_allclose(got, want, rtol=0, atol=tol, equal_nan=True)


def _grids(rng):
    tn = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0, 1, 30)]))
    tn[5] = tn[6]  # a repeated knot
    return np.linspace(0.0, 1.0, 41), tn


def _times(rng, knots):
    return np.concatenate([rng.uniform(-0.2, 1.2, 400), knots, [knots[0], knots[-1]]])


def test_oracle_multi_axis_equals_one_axis_samplers():
    """Component a of the C oracle's multi-axis table (one row per simulation
    too) equals the NumPy oracle's one-axis sampler of amplitude row a, every
    order and boundary, uniform and uneven grids."""
    rng = np.random.default_rng(41)
    co = COracle(TRIVIAL, 2, sampler_names=("unused",))
    for knots in _grids(rng):
        tt = _times(rng, knots)
        for order in (0, 1, 3):
            for zb in (True, False):
                amp = rng.normal(size=(3, knots.size))
                rows = rng.normal(size=(2, 4, knots.size))
                co.set_samplers([(knots, amp, order, zb, True), (knots, rows, order, zb, True)])
                got = co.interp_eval(0, tt, axes=3)
                for a in range(3):
                    want = O.InterpolationSampler(knots, amp[a], order, zb)(tt)
                    _close(got[:, a], want, 1e-14)
                # fewer components than the table holds: the leading ones
                assert np.array_equal(co.interp_eval(0, tt, axes=2), got[:, :2], equal_nan=True)
                for b in range(2):
                    got = co.interp_eval(1, tt, sim=b, axes=4)
                    for a in range(4):
                        want = O.InterpolationSampler(knots, rows[b, a], order, zb)(tt)
                        _close(got[:, a], want, 1e-14)


def test_host_multi_axis_sampler_matches_oracle():
    """The product's sampler (host build) on (A, M) and per-simulation
    (B, A, M) tables against the NumPy oracle's one-axis samplers."""
    rng = np.random.default_rng(43)
    for knots in _grids(rng):
        tt = _times(rng, knots)
        for order in (0, 1, 3):
            for zb in (True, False):
                amp = rng.normal(size=(3, knots.size))
                s = ss.generate_interpolation_sampler(knots, amp, order, zb)
                assert s.axes == 3 and s.rows == 1
                got = s.evaluate_host(tt)
                assert got.shape == (tt.size, 3)
                rows = rng.normal(size=(2, 3, knots.size))
                sb = ss.generate_interpolation_sampler(knots, rows, order, zb, per_simulation=True)
                for a in range(3):
                    want = O.InterpolationSampler(knots, amp[a], order, zb)(tt)
                    _close(got[:, a], want, 1e-12)
                    for b in range(2):
                        want = O.InterpolationSampler(knots, rows[b, a], order, zb)(tt)
                        _close(sb.evaluate_host(tt, row=b)[:, a], want, 1e-12)
                assert s.evaluate_host(0.5).shape == (3,)
    with pytest.raises(IndexError):
        s.evaluate_host(0.5, row=1)


def _multi_axis_inputs(rng):
    knots = np.linspace(0.0, 2e-4, 81)
    amp = rng.normal(0, math.tau * 2e4, (3, knots.size))
    return knots, amp


def test_transpile_multi_axis_reads():
    rng = np.random.default_rng(47)
    knots, amp = _multi_axis_inputs(rng)
    ctl = ss.generate_interpolation_sampler(knots, amp)

    def get_field(time_sample, sweep_parameters, field_sample):
        v = ctl(time_sample)
        field_sample[0] = v[0]
        field_sample[1] = v[1]
        field_sample[2] = math.tau * 1e5 + ctl(time_sample)[2]

    src, samplers, uses_ctx = ss.transpile(get_field)
    assert uses_ctx and samplers == [ctl]
    assert "double v[3] = {0};" in src
    assert "SS_SAMPLE_AXES(0, time_sample, v, 3);" in src
    assert "SS_SAMPLE_AXIS(0, (int)(2), time_sample)" in src

    def scalar_use(time_sample, sweep_parameters, field_sample):
        field_sample[0] = ctl(time_sample)

    def bad_axis(time_sample, sweep_parameters, field_sample):
        v = ctl(time_sample)
        field_sample[0] = v[3]

    def rebinds(time_sample, sweep_parameters, field_sample):
        v = ctl(time_sample)
        v = 2.0
        field_sample[0] = v

    for fn in (scalar_use, bad_axis, rebinds):
        with pytest.raises(ss.TranspileError):
            ss.transpile(fn)


def test_field_source_named_multi_axis_sampler():
    from spinsim_amd import _native
    from spinsim_amd.simulator import FieldSource

    ctl = ss.generate_interpolation_sampler([0.0, 1.0, 2.0], np.ones((2, 3)))
    fs = FieldSource("__device__ void get_field(double t, const double* s, double* f, SS_CONTEXT)"
                     "{ double v[2]; ctl(t, v); f[0] = v[0]; f[1] = v[1]; f[2] = 1; }",
                     samplers={"ctl": ctl})
    assert fs.source.startswith("#define ctl(t, out) SS_SAMPLE_AXES(0, (t), (out), 2)")
    code, _ = _native.compile_code(fs.source, 2, 0, 0, True, 24, True)
    assert code[:4] == b"\x7fELF"


def test_cpu_device_multi_axis_field():
    """A spin-one field whose x, y and z modulation come from one three-axis
    table: the CPU device against the oracle (same translated source), and
    against the same field read through three one-axis samplers."""
    rng = np.random.default_rng(53)
    knots, amp = _multi_axis_inputs(rng)
    ctl = ss.generate_interpolation_sampler(knots, amp)
    singles = [ss.generate_interpolation_sampler(knots, amp[a]) for a in range(3)]
    sx, sy, sz = singles

    def multi(time_sample, sweep_parameters, field_sample):
        v = ctl(time_sample)
        field_sample[0] = v[0]
        field_sample[1] = v[1]
        field_sample[2] = math.tau * 1e5 + 0.1 * v[2]
        field_sample[3] = math.tau * 2e3

    def single(time_sample, sweep_parameters, field_sample):
        field_sample[0] = sx(time_sample)
        field_sample[1] = sy(time_sample)
        field_sample[2] = math.tau * 1e5 + 0.1 * sz(time_sample)
        field_sample[3] = math.tau * 2e3

    psi0 = O.EIGENSTATES["one"]["plus_z"]
    args = (0.0, 2e-4, 2e-8, 1e-6, psi0)
    res = ss.Simulator(multi, ss.SpinQuantumNumber.ONE, device=ss.Device.CPU).evaluate(*args)
    res1 = ss.Simulator(single, ss.SpinQuantumNumber.ONE, device=ss.Device.CPU).evaluate(*args)
    src, _, _ = ss.transpile(multi)
    co = COracle(src, 3, sampler_names=("unused",))
    co.set_samplers([(knots, amp, 3, True, True)])
    dt, n, nf, _ = O.timing(0.0, 2e-4, 2e-8, 1e-6)
    t, u = co.time_evolution([0.0], 0.0, dt, 1e-6, nf, n)
    state = co.get_state(psi0, u)
    assert_close(res, (t, u, state, co.get_spin(state)))
    assert np.abs(res.state - res1.state).max() <= 1e-11
    assert np.abs(res.state - psi0).max() > 1e-2  # the controls act
