"""FP32-bulk kernels: error against the float64 oracle, stated per case.

Tolerances are the accuracy claims of the float32-bulk path: 1e-6 on every
state amplitude for the full 1e6-step spin-one workload (the headline bar),
1e-8 for short rotating-frame runs, looser where each step's rotation is
large (lab frame)."""
from __future__ import annotations

import numpy as np
import pytest

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu


def _run(field, t0, t1, dt, dt_out, psi0, rotating=True, precision="fp32_bulk", expo=None,
         method="MAGNUS_CF4", lanes=None):
    import spinsim_amd as ss

    spin = ss.SpinQuantumNumber.HALF if field.spin == "half" else ss.SpinQuantumNumber.ONE
    e = None if expo is None else getattr(ss.ExponentiationMethod, expo)
    sim = ss.Simulator(field.src, spin, use_rotating_frame=rotating, precision=precision,
                       exponentiation_method=e,
                       integration_method=getattr(ss.IntegrationMethod, method))
    return sim.evaluate(t0, t1, dt, dt_out, psi0, field.sweep, lanes_per_step=lanes)


def _oracle(field, t0, t1, dt, dt_out, psi0, rotating=True, method="magnus_cf4"):
    dim = 2 if field.spin == "half" else 3
    co = COracle(field.src, dim, method, None, rotating)
    dt2, n, nf, _ = O.timing(t0, t1, dt, dt_out)
    _, u = co.time_evolution(field.sweep, float(t0), dt2, float(dt_out), nf, n)
    return u, co.get_state(psi0, u)


@pytest.mark.parametrize("field", [F.TILTED_HALF, F.TILTED_ONE, F.RABI_SWEEP_HALF,
                                   F.RABI_SWEEP_ONE], ids=lambda f: f.name)
@pytest.mark.parametrize("method", ["MAGNUS_CF4", "MIDPOINT_SAMPLE", "HALF_STEP"])
def test_fp32_bulk_rotating_frame(field, method):
    print(f"case: fp32_bulk {field.name} {method}", flush=True)
    psi0 = O.EIGENSTATES[field.spin]["plus_x"]
    m = {"MAGNUS_CF4": "magnus_cf4", "MIDPOINT_SAMPLE": "euler", "HALF_STEP": "heun"}[method]
    res = _run(field, 0.0, 4e-4, 1e-7, 1e-6, psi0, method=method)
    u, st = _oracle(field, 0.0, 4e-4, 1e-7, 1e-6, psi0, method=m)
    err = np.abs(res.state - st).max()
    print(f"  state err {err:.3e}  U err {np.abs(res.time_evolution - u).max():.3e}")
    assert err <= 1e-8


@pytest.mark.parametrize("field", [F.TILTED_HALF, F.TILTED_ONE], ids=lambda f: f.name)
def test_fp32_bulk_lab_frame(field):
    print(f"case: fp32_bulk lab {field.name}", flush=True)
    psi0 = O.EIGENSTATES[field.spin]["plus_y"]
    res = _run(field, 0.0, 4e-4, 1e-7, 1e-6, psi0, rotating=False)
    _, st = _oracle(field, 0.0, 4e-4, 1e-7, 1e-6, psi0, rotating=False)
    err = np.abs(res.state - st).max()
    print(f"  state err {err:.3e}")
    assert err <= 1e-6


def _unitarity(u):
    """max |U U^+ - I| over the coarse operators."""
    d = u.shape[1]
    dev = np.abs(np.einsum("nij,nkj->nik", u, u.conj()) - np.eye(d)).max(initial=0)
    print(f"  max |U U^+ - I| = {dev:.3e}")
    return dev


def test_fp32_bulk_large_steps_scaling_and_squaring():
    """Lab-frame 700 kHz bias: |A| = 0.44 per step, beyond the Taylor range."""
    print("case: fp32_bulk large steps", flush=True)
    psi0 = O.EIGENSTATES["one"]["plus_x"]
    res = _run(F.RABI_ONE, 0.0, 2e-5, 1e-7, 1e-6, psi0, rotating=False)
    _, st = _oracle(F.RABI_ONE, 0.0, 2e-5, 1e-7, 1e-6, psi0, rotating=False)
    err = np.abs(res.state - st).max()
    print(f"  state err {err:.3e}")
    assert err <= 1e-5
    assert _unitarity(res.time_evolution) <= 1e-14


def test_fp32_bulk_batch_operators_reunitarised():
    """The batched kernel stores re-unitarised coarse operators too."""
    import spinsim_amd as ss

    print("case: fp32_bulk batch unitarity", flush=True)
    rng = np.random.default_rng(5)
    b = 16
    sweep = np.tile(np.array(F.RABI_SWEEP_ONE.sweep), (b, 1))
    sweep[:, 1] *= rng.uniform(0.5, 20.0, b)
    sim = ss.Simulator(F.RABI_SWEEP_ONE.src, ss.SpinQuantumNumber.ONE, precision="fp32_bulk",
                       use_rotating_frame=False)
    res = sim.evaluate_batch(0.0, rng.uniform(1e-5, 5e-5, b), 1e-7, 1e-6,
                             O.EIGENSTATES["one"]["plus_x"], sweep)
    assert _unitarity(res.time_evolution.cpu().numpy()) <= 1e-14


def test_fp32_bulk_spin_half_large_angles():
    """Spin half, lab frame, 700 kHz bias at dt = 500 ns: theta = |g|/2 = 1.1 per
    step, past the float32 series of the closed-form residual (float64 sincos)."""
    print("case: fp32_bulk spin-half large angles", flush=True)
    psi0 = O.EIGENSTATES["half"]["plus_x"]
    res = _run(F.RABI_SWEEP_HALF, 0.0, 1e-4, 5e-7, 1e-6, psi0, rotating=False)
    _, st = _oracle(F.RABI_SWEEP_HALF, 0.0, 1e-4, 5e-7, 1e-6, psi0, rotating=False)
    err = np.abs(res.state - st).max()
    print(f"  state err {err:.3e}")
    assert err <= 1e-5


def test_fp32_bulk_full_rabi_one_within_1e6():
    """The headline workload (1e6 CF4 steps): state error <= 1e-6 vs float64."""
    print("case: fp32_bulk full rabi_one", flush=True)
    psi0 = O.EIGENSTATES["one"]["plus_z"]
    res = _run(F.RABI_ONE, 0.0, 100e-3, 100e-9, 1e-6, psi0)
    u, st = _oracle(F.RABI_ONE, 0.0, 100e-3, 100e-9, 1e-6, psi0)
    err = np.abs(res.state - st).max()
    print(f"  state err {err:.3e}")
    assert err <= 1e-6
    # every coarse operator is re-unitarised before it is stored
    assert _unitarity(res.time_evolution) <= 1e-14


def test_precision_error_helper():
    import spinsim_amd as ss

    print("case: precision_error", flush=True)
    sim = ss.Simulator(F.TILTED_ONE.src, ss.SpinQuantumNumber.ONE, precision="fp32_bulk")
    err = ss.precision_error(sim, 0.0, 4e-4, 1e-7, 1e-6, O.EIGENSTATES["one"]["plus_x"])
    print(f"  {err}")
    assert 0 < err["state"] <= 1e-8 and err["time_evolution"] <= 1e-8 and err["unitarity"] <= 1e-14
