"""HIP path vs the FP64 oracle on the same inputs (tolerance 1e-9, absolute, on
every state amplitude, operator entry and spin component)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu

TOL = 1e-9

METHODS = {"magnus_cf4": "MAGNUS_CF4", "euler": "MIDPOINT_SAMPLE", "heun": "HALF_STEP"}


def _ss():
    import spinsim_amd as ss
    return ss


def _spin_enum(spin):
    ss = _ss()
    return ss.SpinQuantumNumber.HALF if spin == "half" else ss.SpinQuantumNumber.ONE


def run_product(field, t0, t1, dt, dt_out, psi0, method="magnus_cf4", expo=None, rotating=True,
                lanes=None, squares=24, sweep=None):
    ss = _ss()
    e = None if expo is None else {"analytic": ss.ExponentiationMethod.ANALYTIC,
                                   "lie_trotter": ss.ExponentiationMethod.LIE_TROTTER,
                                   "lie_trotter_8": ss.ExponentiationMethod.LIE_TROTTER_8}[expo]
    sim = ss.Simulator(field.src, _spin_enum(field.spin), exponentiation_method=e,
                       use_rotating_frame=rotating,
                       integration_method=getattr(ss.IntegrationMethod, METHODS[method]),
                       number_of_squares=squares)
    sweep = field.sweep if sweep is None else sweep
    res = sim.evaluate(t0, t1, dt, dt_out, psi0, sweep, lanes_per_step=lanes)
    return res


def run_oracle(field, t0, t1, dt, dt_out, psi0, method="magnus_cf4", expo=None, rotating=True,
               squares=24, sweep=None):
    dim = 2 if field.spin == "half" else 3
    sweep = field.sweep if sweep is None else sweep
    co = COracle(field.src, dim, method, expo, rotating, squares)
    dt2, n, nf, _ = O.timing(t0, t1, dt, dt_out)
    time, u = co.time_evolution(sweep, float(t0), dt2, float(dt_out), nf, n)
    state = co.get_state(psi0, u)
    spin = co.get_spin(state)
    return time, u, state, spin


def assert_close(res, ref, tol=TOL):
    time, u, state, spin = ref
    assert res.time.shape == time.shape
    assert np.array_equal(res.time, time)
    assert np.abs(res.time_evolution - u).max(initial=0) <= tol
    assert np.abs(res.state - state).max(initial=0) <= tol
    assert np.abs(res.spin - spin).max(initial=0) <= tol


def test_larmor_full_size_and_closed_form():
    """Spin-half Larmor precession, 1e6 steps; also the exact answer."""
    print("case: larmor full size", flush=True)
    psi0 = O.EIGENSTATES["half"]["plus_x"]
    res = run_product(F.LARMOR, 0, 100e-3, 100e-9, 500e-9, psi0)
    ref = run_oracle(F.LARMOR, 0, 100e-3, 100e-9, 500e-9, psi0)
    assert res.time.shape == (200000,)
    assert_close(res, ref)
    t = res.time
    exact = np.stack([np.exp(-1j * math.pi * 1000 * t), np.exp(1j * math.pi * 1000 * t)], 1) / math.sqrt(2)
    assert np.abs(res.state - exact).max() < 1e-9
    s_exact = np.stack([0.5 * np.cos(TAU1000 * t), 0.5 * np.sin(TAU1000 * t), 0 * t], 1)
    assert np.abs(res.spin - s_exact).max() < 1e-9


TAU1000 = 2 * math.pi * 1000


def test_larmor_twenty_million_rows():
    """A large single simulation: 2e7 coarse rows of one fine step each
    (Larmor, 0.1 s at 5 ns): 1.3 GB of operators, 39,063 scan blocks; every
    row against the C oracle, and the closed form."""
    print("case: larmor 2e7 rows", flush=True)
    psi0 = O.EIGENSTATES["half"]["plus_x"]
    res = run_product(F.LARMOR, 0, 0.1, 5e-9, 5e-9, psi0)
    time, u, state, spin = run_oracle(F.LARMOR, 0, 0.1, 5e-9, 5e-9, psi0)
    n = 20000000
    assert res.time.shape == (n,)
    assert np.array_equal(res.time, time)
    assert np.abs(res.time_evolution - u).max() <= TOL
    # A chain of N = 2e7 float64 products carries ~N eps = 4e-9 of rounding;
    # the oracle multiplies in sequence, the scan in a tree, so their states
    # may differ by that much (both stay within it of the closed form).
    chain = n * np.finfo(float).eps
    t = res.time
    exact = np.stack([np.exp(-1j * math.pi * 1000 * t), np.exp(1j * math.pi * 1000 * t)], 1)
    exact /= math.sqrt(2)
    e_gpu = np.abs(res.state - exact).max()
    e_ora = np.abs(state - exact).max()
    print(f"  vs closed form: HIP {e_gpu:.2e}, oracle {e_ora:.2e}; chain bound {chain:.2e}")
    assert np.abs(res.state - state).max() <= chain
    assert np.abs(res.spin - spin).max() <= chain
    assert e_gpu <= chain and e_ora <= chain


def test_rabi_one_full_size():
    """Spin-one Rabi drive, Lie-Trotter, 1e6 steps, plain FP64."""
    print("case: rabi_one full size", flush=True)
    psi0 = O.EIGENSTATES["one"]["plus_z"]
    res = run_product(F.RABI_ONE, 0, 100e-3, 100e-9, 1e-6, psi0)
    ref = run_oracle(F.RABI_ONE, 0, 100e-3, 100e-9, 1e-6, psi0)
    assert res.time.shape == (100000,)
    assert_close(res, ref)


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
    print(f"case: {field.name} {method} {expo} rotating={rotating}", flush=True)
    psi0 = O.EIGENSTATES[field.spin]["plus_x"]
    args = (field, 1e-5, 1e-5 + 4e-4, 1e-7, 1e-6, psi0)
    res = run_product(*args, method=method, expo=expo, rotating=rotating)
    ref = run_oracle(*args, method=method, expo=expo, rotating=rotating)
    assert_close(res, ref)


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("rotating", [True, False])
def test_eight_generator_lie_trotter(method, rotating):
    """ExponentiationMethod.LIE_TROTTER_8: all eight su(3) directions driven,
    both frames (U1/U2 and V1/V2 rotate at 2x and 1x the frame rate)."""
    print(f"case: lie_trotter_8 {method} rotating={rotating}", flush=True)
    psi0 = O.EIGENSTATES["one"]["plus_x"]
    args = (F.TILTED_EIGHT, 1e-5, 1e-5 + 4e-4, 1e-7, 1e-6, psi0)
    res = run_product(*args, method=method, expo="lie_trotter_8", rotating=rotating)
    ref = run_oracle(*args, method=method, expo="lie_trotter_8", rotating=rotating)
    assert_close(res, ref)


def test_eight_generator_python_callback():
    """A numba-style Python field writing all eight entries, translated to HIP-C."""
    ss = _ss()
    print("case: lie_trotter_8 python callback", flush=True)

    def get_field(time_sample, sweep_parameters, field_sample):
        field_sample[0] = math.tau * 3e4 * math.cos(math.tau * 2e4 * time_sample)
        field_sample[1] = 0
        field_sample[2] = math.tau * 2e4
        field_sample[3] = math.tau * 4e3
        field_sample[4] = math.tau * 8e3 * math.cos(math.tau * 1.5e4 * time_sample)
        field_sample[5] = math.tau * 3e3
        field_sample[6] = math.tau * 6e3 * math.sin(math.tau * 2.5e4 * time_sample)
        field_sample[7] = -math.tau * 2e3

    sim = ss.Simulator(get_field, ss.SpinQuantumNumber.ONE,
                       exponentiation_method=ss.ExponentiationMethod.LIE_TROTTER_8)
    psi0 = O.EIGENSTATES["one"]["plus_z"]
    res = sim.evaluate(0, 2e-4, 1e-7, 1e-6, psi0)
    src, _, _ = ss.transpile(get_field)
    co = COracle(src, 3, "magnus_cf4", "lie_trotter_8")
    dt, n, nf, _ = O.timing(0, 2e-4, 1e-7, 1e-6)
    _, u = co.time_evolution([0.0], 0.0, dt, 1e-6, nf, n)
    assert np.abs(res.time_evolution - u).max() <= TOL
    assert np.abs(res.state - co.get_state(psi0, u)).max() <= TOL


@pytest.mark.parametrize("lanes", [1, 2, 3, 4, 5, 7, 8, 10, 16, 64])
def test_lanes_per_step_partition(lanes):
    """The chunked shuffle-tree product matches the sequential oracle for any G."""
    print(f"case: lanes_per_step={lanes}", flush=True)
    psi0 = O.EIGENSTATES["one"]["plus_y"]
    args = (F.TILTED_ONE, 0, 2e-4, 5e-8, 1e-6, psi0)
    res = run_product(*args, lanes=lanes)
    ref = run_oracle(*args)
    assert_close(res, ref)


PHASED = [
    # (t1, dt, dt_out): n_fine 200 over 3 rows (powers of two, uneven split);
    # one row; one fine step per row; uneven rows per phase (n_fine 7);
    # n_fine 20 with several phases
    (3e-4, 5e-7, 1e-4),
    (1e-6, 1e-7, 1e-6),
    (1e-3, 1e-6, 1e-6),
    (7.0003e-3, 1e-7, 7e-7),
    (2e-4, 5e-8, 1e-6),
]


@pytest.mark.parametrize("t1,dt,dt_out", PHASED)
def test_phased_launch(t1, dt, dt_out):
    """The default launch (rows dealt to the resident waves in phases of
    decreasing lanes-per-step item length) matches the oracle, also on one
    wave per SIMD."""
    print(f"case: phased t1={t1} dt={dt} dt_out={dt_out}", flush=True)
    psi0 = O.EIGENSTATES["one"]["plus_y"]
    args = (F.TILTED_ONE, 0, t1, dt, dt_out, psi0)
    ref = run_oracle(*args)
    assert_close(run_product(*args, lanes=0), ref)
    assert_close(run_product(*args, lanes=-1), ref)
    assert_close(run_product(*args), ref)


@pytest.mark.parametrize("squares", [0, 1, 7, 22, 23, 24, 30])
def test_number_of_squares(squares):
    """Spin-one Lie-Trotter at several square counts, either side of 12
    hypercubes (22 squares: 11, squared in full; 23, 24 and 30: 12 and 15,
    with the derived last row)."""
    print(f"case: number_of_squares={squares}", flush=True)
    psi0 = O.EIGENSTATES["one"]["plus_x"]
    args = (F.TILTED_ONE, 0, 1e-4, 1e-7, 1e-6, psi0)
    res = run_product(*args, squares=squares)
    ref = run_oracle(*args, squares=squares)
    assert_close(res, ref)


STRONG_ONE = F.Field("strong_one", """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = SS_TAU * 1.2e5 * cos(SS_TAU * 3e4 * t); f[1] = SS_TAU * 9e4 * sin(SS_TAU * 5e4 * t);
  f[2] = SS_TAU * 1.5e5 * cos(SS_TAU * 2e4 * t); f[3] = SS_TAU * 2e5; }
""", None, "one")


@pytest.mark.parametrize("rotating", [True, False])
def test_strong_spin_one_steps(rotating):
    """Coarse 1 us steps of a strong spin-one field with a quadratic shift, so
    |g| per exponential reaches ~1-2: the regime where the reference's
    Lie-Trotter start departs most from special unitarity and the kernel's
    derived-row squares differ most from full squares (a few 1e-15 per
    exponential at 24 squares), held to 1e-9 over 4,000 steps."""
    print(f"case: strong spin-one steps rotating={rotating}", flush=True)
    psi0 = O.EIGENSTATES["one"]["plus_x"]
    args = (STRONG_ONE, 0, 4e-3, 1e-6, 1e-5, psi0)
    assert_close(run_product(*args, rotating=rotating), run_oracle(*args, rotating=rotating))


def test_edge_sizes():
    print("case: edge sizes", flush=True)
    psi0 = O.EIGENSTATES["half"]["plus_y"]
    # fewer than one output step: N = 0
    res = run_product(F.TILTED_HALF, 0, 0.5e-6, 1e-7, 1e-6, psi0)
    assert res.time.shape == (0,) and res.state.shape == (0, 2) and res.spin.shape == (0, 3)
    # exactly one output step
    res = run_product(F.TILTED_HALF, 0, 1e-6, 1e-7, 1e-6, psi0)
    ref = run_oracle(F.TILTED_HALF, 0, 1e-6, 1e-7, 1e-6, psi0)
    assert_close(res, ref)
    # one fine step per output step
    res = run_product(F.TILTED_HALF, 0, 5e-5, 1e-6, 1e-6, psi0)
    ref = run_oracle(F.TILTED_HALF, 0, 5e-5, 1e-6, 1e-6, psi0)
    assert_close(res, ref)
    # non-integer ratio: dt is re-derived (warning path) and the counts follow it
    res = run_product(F.TILTED_HALF, 0, 3e-5, 0.3e-6, 1e-6, psi0)
    ref = run_oracle(F.TILTED_HALF, 0, 3e-5, 0.3e-6, 1e-6, psi0)
    assert_close(res, ref)


def test_python_callback_matches_hip_source():
    """A numba-style Python get_field, translated to HIP-C, gives the same run."""
    print("case: python callback", flush=True)
    ss = _ss()

    def get_field_rabi(time_sample, sweep_parameters, field_sample):
        field_sample[0] = 2 * sweep_parameters[1] * math.cos(sweep_parameters[0] * time_sample)
        field_sample[1] = 0
        field_sample[2] = sweep_parameters[0] + sweep_parameters[2]

    sim = ss.Simulator(get_field_rabi, ss.SpinQuantumNumber.HALF)
    psi0 = O.EIGENSTATES["half"]["plus_z"]
    res = sim.evaluate(0, 2e-4, 1e-7, 1e-6, psi0, F.RABI_SWEEP_HALF.sweep)
    ref = run_oracle(F.RABI_SWEEP_HALF, 0, 2e-4, 1e-7, 1e-6, psi0)
    assert_close(res, ref)


def test_gaussian_pulse_known_answer():
    """gaussian_example.py: error vs the analytic pi-pulse below 1e-6 at 40 steps."""
    print("case: gaussian pulse", flush=True)

    def cumulative_gaussian(t):
        return 0.5 * (1 + math.erf(t / math.sqrt(2.0))) - 0.5 * (1 + math.erf(-5.0 / math.sqrt(2.0)))

    time = np.arange(-5.0, 5.1, 2.0)
    analytic = np.asarray([[math.cos(0.5 * math.pi * cumulative_gaussian(t)),
                            -1j * math.sin(0.5 * math.pi * cumulative_gaussian(t))] for t in time])
    errors = []
    for step in (0.25, 0.1, 0.01):
        res = run_product(F.GAUSSIAN, -5.0, 7.0, step, 2.0, np.asarray([1, 0], np.complex128))
        ref = run_oracle(F.GAUSSIAN, -5.0, 7.0, step, 2.0, np.asarray([1, 0], np.complex128))
        assert_close(res, ref)
        errors.append(np.sqrt(np.sum(np.abs(res.state - analytic) ** 2)) / 5)
    assert errors[0] < 1e-6          # 40 steps (about.examples.rst:567)
    assert errors[1] < errors[0] / 20   # fourth order: (0.25/0.1)^4 = 39
    assert errors[2] < 1e-11


def test_state_scan_random_unitaries():
    """ss_state_scan vs the sequential product on sizes around tile edges and
    look-backs over more than 256 tiles."""
    ss = _ss()
    import torch

    rng = np.random.default_rng(7)
    for dim in (2, 3):
        for n in (1, 2, 7, 8, 9, 255, 256, 257, 1023, 1024, 1025, 2049, 4097, 65536, 300001):
            print(f"case: state scan dim={dim} n={n}", flush=True)
            a = rng.normal(size=(n, dim, dim)) + 1j * rng.normal(size=(n, dim, dim))
            q, _ = np.linalg.qr(a)
            psi0 = O.EIGENSTATES["half" if dim == 2 else "one"]["plus_x"]
            co = COracle(F.LARMOR.src if dim == 2 else F.RABI_ONE.src, dim)
            ref = co.get_state(psi0, q)
            sim = ss.Simulator(F.LARMOR.src if dim == 2 else F.RABI_ONE.src,
                               ss.SpinQuantumNumber.HALF if dim == 2 else ss.SpinQuantumNumber.ONE)
            st = sim.state_scan(psi0, torch.as_tensor(q, device="cuda")).cpu().numpy()
            assert np.abs(st - ref).max() <= TOL
            spin = sim.spin_calculator(st)
            assert np.abs(spin - co.get_spin(ref)).max() <= TOL


def test_interpolation_sampler_device_eval():
    """ss_interp_eval vs the oracle's restatement of the sampler, every order/boundary."""
    ss = _ss()
    rng = np.random.default_rng(3)
    tu = np.linspace(0.0, 1.0, 41)
    tn = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0, 1, 30)]))
    tn[5] = tn[6]  # a repeated knot
    for time in (tu, tn):
        amp = rng.normal(size=time.size)
        for order in (0, 1, 3):
            for zb in (True, False):
                print(f"case: sampler order={order} zero_boundary={zb}", flush=True)
                s = ss.generate_interpolation_sampler(time, amp, order, zb)
                ref = O.InterpolationSampler(time, amp, order, zb)
                t = np.concatenate([rng.uniform(-0.2, 1.2, 2000), time, [-1e-300, 1.0 + 1e-16]])
                got, want = s(t), ref(t)
                assert np.array_equal(np.isnan(got), np.isnan(want))
                ok = ~np.isnan(want)
                assert np.abs(got[ok] - want[ok]).max() <= 1e-12


def test_sampler_field_run():
    """A field read from an interpolation table (noise on z, interpolator_example.py)."""
    print("case: sampler field", flush=True)
    ss = _ss()
    rng = np.random.default_rng(11)
    time_sampled = np.arange(0, 2e-3, 1e-5)
    noise = rng.normal(size=time_sampled.size)
    sampler = ss.generate_interpolation_sampler(time_sampled, noise)

    def get_field(time_sample, sweep_parameters, field_sample):
        field_sample[2] = math.tau * (2e4 + sweep_parameters[0] * sampler(time_sample))
        field_sample[0] = math.tau * 3e3
        field_sample[1] = 0
        field_sample[3] = math.tau * 500

    sim = ss.Simulator(get_field, ss.SpinQuantumNumber.ONE)
    psi0 = O.EIGENSTATES["one"]["plus_x"]
    res = sim.evaluate(0, 2e-3, 1e-7, 1e-6, psi0, [2e3])
    src, samplers, ctx = ss.transpile(get_field)
    co = COracle(src, 3, sampler_names=("unused_name",))
    co.set_samplers([(time_sampled, noise, 3, True)])
    dt, n, nf, _ = O.timing(0, 2e-3, 1e-7, 1e-6)
    t, u = co.time_evolution([2e3], 0.0, dt, 1e-6, nf, n)
    state = co.get_state(psi0, u)
    assert_close(res, (t, u, state, co.get_spin(state)))
