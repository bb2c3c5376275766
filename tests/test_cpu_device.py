"""The CPU devices (Device.CPU, CPU_SINGLE, PYTHON): the host build of the
product's integrator text, held to the float64 oracle's C restatement and the
committed golden fixtures at 1e-9 (absolute, on every operator entry, state
amplitude and spin component), like the HIP path.  These run without a GPU."""
from __future__ import annotations

import math
import time

import numpy as np
import pytest

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle
from test_golden import FIELDS, TRAJECTORIES, load

import spinsim_amd as ss
from spinsim_amd import host

TOL = 1e-9
FP32_TOL = 1e-6
METHODS = {"magnus_cf4": "MAGNUS_CF4", "euler": "MIDPOINT_SAMPLE", "heun": "HALF_STEP"}
EXPOS = {None: None, "analytic": "ANALYTIC", "lie_trotter": "LIE_TROTTER",
         "lie_trotter_8": "LIE_TROTTER_8"}


def _spin(label):
    return ss.SpinQuantumNumber.HALF if label == "half" else ss.SpinQuantumNumber.ONE


def cpu_sim(field, method="magnus_cf4", expo=None, rotating=True, squares=24,
            device=ss.Device.CPU, precision="fp64", threads=None):
    e = EXPOS[expo]
    return ss.Simulator(field.src if hasattr(field, "src") else field, _spin(field.spin)
                        if hasattr(field, "spin") else ss.SpinQuantumNumber.HALF,
                        device=device,
                        exponentiation_method=None if e is None else getattr(ss.ExponentiationMethod, e),
                        use_rotating_frame=rotating,
                        integration_method=getattr(ss.IntegrationMethod, METHODS[method]),
                        number_of_squares=squares, precision=precision, number_of_threads=threads)


def run_oracle(field, t0, t1, dt, dt_out, psi0, method="magnus_cf4", expo=None, rotating=True,
               squares=24, sweep=None):
    dim = 2 if field.spin == "half" else 3
    sweep = field.sweep if sweep is None else sweep
    co = COracle(field.src, dim, method, expo, rotating, squares)
    dt2, n, nf, _ = O.timing(t0, t1, dt, dt_out)
    t, u = co.time_evolution(sweep, float(t0), dt2, float(dt_out), nf, n)
    state = co.get_state(psi0, u)
    return t, u, state, co.get_spin(state)


def assert_close(res, ref, tol=TOL):
    t, u, state, spin = ref
    assert isinstance(res.state, np.ndarray)
    assert np.array_equal(res.time, t)
    assert np.abs(res.time_evolution - u).max(initial=0) <= tol
    assert np.abs(res.state - state).max(initial=0) <= tol
    assert np.abs(res.spin - spin).max(initial=0) <= tol


@pytest.mark.parametrize("name", TRAJECTORIES)
def test_trajectory_matches_fixture(name):
    meta, z = load(name)
    field = FIELDS[meta["field"]]
    sim = cpu_sim(field, meta["method"], meta["expo"], meta["rotating"])
    res = sim.evaluate(meta["t0"], meta["t1"], meta["dt"], meta["dt_out"], z["psi0"],
                       meta["sweep"])
    r = z["rows"]
    assert res.time.shape == (meta["n_rows"],)
    assert np.array_equal(res.time[r], z["time"])
    assert np.abs(res.time_evolution[r] - z["u"]).max() <= TOL
    assert np.abs(res.state[r] - z["state"]).max() <= TOL
    assert np.abs(res.spin[r] - z["spin"]).max() <= TOL


@pytest.mark.parametrize("name", ["larmor_half_1ms", "rabi_one_1ms", "methods_magnus_cf4_half_lab",
                                  "methods_magnus_cf4_one_rotating"])
def test_fp32_bulk_trajectory_matches_fixture(name):
    meta, z = load(name)
    sim = cpu_sim(FIELDS[meta["field"]], meta["method"], None, meta["rotating"],
                  precision="fp32_bulk")
    res = sim.evaluate(meta["t0"], meta["t1"], meta["dt"], meta["dt_out"], z["psi0"],
                       meta["sweep"])
    assert np.abs(res.state[z["rows"]] - z["state"]).max() <= FP32_TOL


CASES = []
for _field in (F.TILTED_HALF, F.RABI_SWEEP_HALF):
    for _expo in ("analytic", "lie_trotter"):
        for _m in METHODS:
            for _rot in (True, False):
                CASES.append((_field, _m, _expo, _rot))
for _field in (F.TILTED_ONE, F.RABI_SWEEP_ONE):
    for _m in METHODS:
        for _rot in (True, False):
            CASES.append((_field, _m, None, _rot))


@pytest.mark.parametrize("field,method,expo,rotating", CASES,
                         ids=[f"{c[0].name}-{c[1]}-{c[2]}-{'rot' if c[3] else 'lab'}" for c in CASES])
def test_methods_frames(field, method, expo, rotating):
    psi0 = O.EIGENSTATES[field.spin]["plus_x"]
    args = (1e-5, 1e-5 + 4e-4, 1e-7, 1e-6, psi0)
    res = cpu_sim(field, method, expo, rotating).evaluate(*args, field.sweep)
    assert_close(res, run_oracle(field, *args, method=method, expo=expo, rotating=rotating))


@pytest.mark.parametrize("method", list(METHODS))
def test_eight_generator_lie_trotter(method):
    psi0 = O.EIGENSTATES["one"]["plus_x"]
    args = (1e-5, 1e-5 + 4e-4, 1e-7, 1e-6, psi0)
    for rotating in (True, False):
        res = cpu_sim(F.TILTED_EIGHT, method, "lie_trotter_8", rotating).evaluate(
            *args, F.TILTED_EIGHT.sweep)
        assert_close(res, run_oracle(F.TILTED_EIGHT, *args, method=method, expo="lie_trotter_8",
                                     rotating=rotating))


@pytest.mark.parametrize("squares", [0, 1, 7, 22, 24, 30])
def test_number_of_squares(squares):
    psi0 = O.EIGENSTATES["one"]["plus_x"]
    args = (0, 1e-4, 1e-7, 1e-6, psi0)
    res = cpu_sim(F.TILTED_ONE, squares=squares).evaluate(*args, F.TILTED_ONE.sweep)
    assert_close(res, run_oracle(F.TILTED_ONE, *args, squares=squares))


def test_larmor_full_size_and_closed_form():
    """Spin-half Larmor precession, 1e6 steps (2e5 x 5), against the oracle
    and psi(t) = (exp(-i pi 1e3 t), exp(i pi 1e3 t)) / sqrt 2."""
    psi0 = O.EIGENSTATES["half"]["plus_x"]
    res = cpu_sim(F.LARMOR).evaluate(0, 100e-3, 100e-9, 500e-9, psi0)
    assert res.time.shape == (200000,)
    assert_close(res, run_oracle(F.LARMOR, 0, 100e-3, 100e-9, 500e-9, psi0))
    t = res.time
    exact = np.stack([np.exp(-1j * math.pi * 1e3 * t), np.exp(1j * math.pi * 1e3 * t)], 1)
    assert np.abs(res.state - exact / math.sqrt(2)).max() < 1e-9


def test_rabi_one_full_size():
    """Spin-one Rabi drive, Lie-Trotter, 1e6 steps, against the oracle."""
    psi0 = O.EIGENSTATES["one"]["plus_z"]
    res = cpu_sim(F.RABI_ONE).evaluate(0, 100e-3, 100e-9, 1e-6, psi0)
    assert res.time.shape == (100000,)
    assert_close(res, run_oracle(F.RABI_ONE, 0, 100e-3, 100e-9, 1e-6, psi0))


def test_devices_and_thread_counts_agree_bitwise():
    """CPU (all cores, or number_of_threads), CPU_SINGLE and PYTHON compute
    every coarse step with the same arithmetic: identical results."""
    psi0 = O.EIGENSTATES["one"]["plus_y"]
    args = (0, 3e-4, 5e-8, 1e-6, psi0, F.TILTED_ONE.sweep)
    runs = [cpu_sim(F.TILTED_ONE, device=d, threads=n).evaluate(*args)
            for d, n in ((ss.Device.CPU, None), (ss.Device.CPU, 3), (ss.Device.CPU_SINGLE, None),
                         (ss.Device.PYTHON, None))]
    for r in runs[1:]:
        assert np.array_equal(r.time_evolution, runs[0].time_evolution)
        assert np.array_equal(r.state, runs[0].state)
    sims = [cpu_sim(F.TILTED_ONE, device=ss.Device.CPU, threads=3),
            cpu_sim(F.TILTED_ONE, device=ss.Device.CPU_SINGLE),
            cpu_sim(F.TILTED_ONE, device=ss.Device.PYTHON)]
    assert [s._host_threads() for s in sims] == [3, 1, 1]


def test_default_device_is_cpu_without_hip(monkeypatch):
    """The reference picks CUDA when available, else CPU (:435-439); here
    ROCm when a HIP device is visible, else CPU."""
    import torch

    monkeypatch.setattr(torch.cuda, "is_available", lambda: False)
    sim = ss.Simulator(F.LARMOR.src, ss.SpinQuantumNumber.HALF)
    assert sim.device == ss.Device.CPU
    res = sim.evaluate(0, 1e-5, 1e-7, 1e-6, O.EIGENSTATES["half"]["plus_x"])
    assert res.state.shape == (10, 2)


def test_get_time_evolution_and_spin_calculator_in_place():
    """The reference's kernel signature (:854) fills numpy arrays in place;
    spin_calculator (:961) takes a state series."""
    sim = cpu_sim(F.TILTED_HALF)
    dt, n, nf, _ = O.timing(0, 2e-5, 1e-7, 1e-6)
    t = np.empty(n)
    u = np.empty((n, 2, 2), np.complex128)
    sim.get_time_evolution(F.TILTED_HALF.sweep, t, np.array([0.0, 2e-5]), dt, 1e-6, u)
    ref = run_oracle(F.TILTED_HALF, 0, 2e-5, 1e-7, 1e-6, O.EIGENSTATES["half"]["plus_x"])
    assert np.array_equal(t, ref[0]) and np.abs(u - ref[1]).max() <= TOL
    assert np.abs(sim.spin_calculator(ref[2]) - ref[3]).max() <= TOL


def test_edge_sizes():
    psi0 = O.EIGENSTATES["half"]["plus_y"]
    res = cpu_sim(F.TILTED_HALF).evaluate(0, 0.5e-6, 1e-7, 1e-6, psi0, F.TILTED_HALF.sweep)
    assert res.time.shape == (0,) and res.state.shape == (0, 2) and res.spin.shape == (0, 3)
    for args in ((0, 1e-6, 1e-7, 1e-6), (0, 5e-5, 1e-6, 1e-6), (0, 3e-5, 0.3e-6, 1e-6)):
        res = cpu_sim(F.TILTED_HALF).evaluate(*args, psi0, F.TILTED_HALF.sweep)
        assert_close(res, run_oracle(F.TILTED_HALF, *args, psi0))


def test_python_callback():
    def get_field_rabi(time_sample, sweep_parameters, field_sample):
        field_sample[0] = 2 * sweep_parameters[1] * math.cos(sweep_parameters[0] * time_sample)
        field_sample[1] = 0
        field_sample[2] = sweep_parameters[0] + sweep_parameters[2]

    sim = ss.Simulator(get_field_rabi, ss.SpinQuantumNumber.HALF, device=ss.Device.CPU)
    psi0 = O.EIGENSTATES["half"]["plus_z"]
    res = sim.evaluate(0, 2e-4, 1e-7, 1e-6, psi0, F.RABI_SWEEP_HALF.sweep)
    assert_close(res, run_oracle(F.RABI_SWEEP_HALF, 0, 2e-4, 1e-7, 1e-6, psi0))


def test_sampler_field_run_and_host_sampler():
    """A field read from a cubic interpolation table, and the sampler called
    on the host, against the oracle's sampler (every order and boundary)."""
    rng = np.random.default_rng(11)
    time_sampled = np.arange(0, 2e-3, 1e-5)
    noise = rng.normal(size=time_sampled.size)
    sampler = ss.generate_interpolation_sampler(time_sampled, noise)

    def get_field(time_sample, sweep_parameters, field_sample):
        field_sample[2] = math.tau * (2e4 + sweep_parameters[0] * sampler(time_sample))
        field_sample[0] = math.tau * 3e3
        field_sample[1] = 0
        field_sample[3] = math.tau * 500

    sim = ss.Simulator(get_field, ss.SpinQuantumNumber.ONE, device=ss.Device.CPU)
    psi0 = O.EIGENSTATES["one"]["plus_x"]
    res = sim.evaluate(0, 2e-3, 1e-7, 1e-6, psi0, [2e3])
    src, _, _ = ss.transpile(get_field)
    co = COracle(src, 3, sampler_names=("unused_name",))
    co.set_samplers([(time_sampled, noise, 3, True)])
    dt, n, nf, _ = O.timing(0, 2e-3, 1e-7, 1e-6)
    t, u = co.time_evolution([2e3], 0.0, dt, 1e-6, nf, n)
    state = co.get_state(psi0, u)
    assert_close(res, (t, u, state, co.get_spin(state)))
    tn = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0, 1, 30)]))
    tn[5] = tn[6]
    for knots in (np.linspace(0.0, 1.0, 41), tn):
        amp = rng.normal(size=knots.size)
        for order in (0, 1, 3):
            for zb in (True, False):
                s = ss.generate_interpolation_sampler(knots, amp, order, zb)
                ref = O.InterpolationSampler(knots, amp, order, zb)
                tt = np.concatenate([rng.uniform(-0.2, 1.2, 500), knots])
                got, want = s.evaluate_host(tt), ref(tt)
                assert np.abs(got - want).max() <= 1e-12


def test_gaussian_pulse_known_answer():
    """gaussian_example.py: error against the analytic pi-pulse below 1e-6 at
    40 steps, fourth-order convergence."""
    def cumulative_gaussian(t):
        return 0.5 * (1 + math.erf(t / math.sqrt(2.0))) - 0.5 * (1 + math.erf(-5.0 / math.sqrt(2.0)))

    tt = np.arange(-5.0, 5.1, 2.0)
    analytic = np.asarray([[math.cos(0.5 * math.pi * cumulative_gaussian(t)),
                            -1j * math.sin(0.5 * math.pi * cumulative_gaussian(t))] for t in tt])
    errors = []
    for step in (0.25, 0.1, 0.01):
        psi0 = np.asarray([1, 0], np.complex128)
        res = cpu_sim(F.GAUSSIAN).evaluate(-5.0, 7.0, step, 2.0, psi0, F.GAUSSIAN.sweep)
        assert_close(res, run_oracle(F.GAUSSIAN, -5.0, 7.0, step, 2.0, psi0))
        errors.append(np.sqrt(np.sum(np.abs(res.state - analytic) ** 2)) / 5)
    assert errors[0] < 1e-6 and errors[1] < errors[0] / 20 and errors[2] < 1e-11


INFINITE_AT_ZERO = F.Field("infinite_at_zero", """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = 1.0 / t; f[1] = 0; f[2] = SS_TAU * 1e5; f[3] = 0; }
""", None, "one")


def test_fp32_bulk_residual_squares_bounded():
    """The float32-bulk spin-one residual's squaring count is bounded for every
    input: inf and NaN norm bounds give 0 (the series then carries inf/NaN
    into the result, as float64 does), huge finite ones at most ~1,027; a
    field infinite at t = 0 (1/t sampled there by the half-step method)
    gives NaN operators promptly instead of ~2^31 squarings per step."""
    sim = cpu_sim(INFINITE_AT_ZERO, method="heun", precision="fp32_bulk")
    sq = sim._host_module.lib.ss_host_residual_squares
    assert sq(float("inf")) == 0 and sq(float("nan")) == 0 and sq(-1.0) == 0
    assert sq(0.17) == 0 and sq(0.18) == 1 and sq(0.34) == 1 and sq(0.35) == 2
    assert sq(1.7976931348623157e308) == 1027
    t = time.perf_counter()
    res = sim.evaluate(0.0, 4e-6, 1e-7, 1e-6, O.EIGENSTATES["one"]["plus_z"])
    assert time.perf_counter() - t < 30
    assert np.isnan(res.time_evolution[0]).any()
    assert np.isfinite(res.time_evolution[1:]).all()


def test_host_library_rejects_bad_arguments():
    with pytest.raises(host.HostError):
        host._check(host.host_lib().ss_host_state(None, None, None, None, 3, 4), "ss_host_state")
    with pytest.raises(ValueError):
        cpu_sim(F.TILTED_HALF)._host_module.time_evolution(
            [0.0], None, 0.0, 1e-7, 1e-6, np.empty(3), np.empty((2, 2, 2), np.complex128), 1)
