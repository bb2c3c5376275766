"""The HIP path against the committed golden fixtures (tests/golden/*.npz):
float64 kernels to 1e-9 on every stored operator entry, state amplitude and
spin component; the float32-bulk kernels to 1e-6 on the states; adjoint
gradients to 1e-8 relative."""
from __future__ import annotations

import math

import numpy as np
import pytest

import fields as F
from test_golden import FIELDS, TRAJECTORIES, load

pytestmark = pytest.mark.gpu
TOL = 1e-9
FP32_TOL = 1e-6
INTEGRATION = {"magnus_cf4": "MAGNUS_CF4", "euler": "MIDPOINT_SAMPLE", "heun": "HALF_STEP"}
EXPONENTIATION = {None: None, "analytic": "ANALYTIC", "lie_trotter": "LIE_TROTTER",
                  "lie_trotter_8": "LIE_TROTTER_8"}


def _simulator(meta, field, precision="fp64"):
    import spinsim_amd as ss

    spin = ss.SpinQuantumNumber.HALF if meta["spin"] == "half" else ss.SpinQuantumNumber.ONE
    expo = EXPONENTIATION[meta["expo"]]
    return ss.Simulator(field.src, spin,
                        exponentiation_method=None if expo is None
                        else getattr(ss.ExponentiationMethod, expo),
                        use_rotating_frame=meta["rotating"],
                        integration_method=getattr(ss.IntegrationMethod,
                                                   INTEGRATION[meta["method"]]),
                        precision=precision)


@pytest.mark.parametrize("name", TRAJECTORIES)
def test_trajectory_matches_fixture(name):
    meta, z = load(name)
    print(f"case: golden {name}", flush=True)
    sim = _simulator(meta, FIELDS[meta["field"]])
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
    print(f"case: golden fp32_bulk {name}", flush=True)
    sim = _simulator(meta, FIELDS[meta["field"]], precision="fp32_bulk")
    res = sim.evaluate(meta["t0"], meta["t1"], meta["dt"], meta["dt_out"], z["psi0"],
                       meta["sweep"])
    err = np.abs(res.state[z["rows"]] - z["state"]).max()
    print(f"  state error {err:.2e}")
    assert err <= FP32_TOL


@pytest.mark.parametrize("precision,tol", [("fp64", TOL), ("fp32_bulk", FP32_TOL)])
def test_batch_matches_fixture(precision, tol):
    import spinsim_amd as ss

    meta, z = load("batch_rabi_half_4")
    print(f"case: golden batch {precision}", flush=True)
    sim = ss.Simulator(F.RABI_SWEEP_HALF.src, ss.SpinQuantumNumber.HALF, precision=precision)
    res = sim.evaluate_batch(z["t0"], z["t1"], meta["dt"], meta["dt_out"], z["psi0"], z["sweep"])
    assert np.array_equal(np.asarray(res.offsets), z["offsets"])
    r = z["rows"]
    assert np.array_equal(res.time.cpu().numpy()[r], z["time"])
    assert np.abs(res.state.cpu().numpy()[r] - z["state"]).max() <= tol
    assert np.abs(res.spin.cpu().numpy()[r] - z["spin"]).max() <= tol
    if precision == "fp64":
        assert np.abs(res.time_evolution.cpu().numpy()[r] - z["u"]).max() <= tol


def _z_bias(time_sample, sweep_parameters, field_sample):
    field_sample[2] = 2 * math.pi * 1e5


@pytest.mark.parametrize("precision,psi_tol,grad_tol", [("fp64", TOL, 1e-8),
                                                        ("fp32_bulk", FP32_TOL, 1e-6)])
def test_adjoint_matches_fixture(precision, psi_tol, grad_tol):
    import torch

    import spinsim_amd as ss

    meta, z = load("adjoint_one_2")
    print(f"case: golden adjoint {precision}", flush=True)
    cs = ss.ControlledSimulator(ss.SpinQuantumNumber.ONE, [ss.Control("x", z["tk"]),
                                                           ss.Control("y", z["tk"])], _z_bias,
                                precision=precision)
    amps = torch.tensor(z["amp"], device="cuda", requires_grad=True)
    psi_t = cs.final_state(amps, meta["t0"], meta["t1"], meta["dt"], meta["dt_out"], z["psi0"])
    (1 - psi_t[:, -1].abs() ** 2).sum().backward()
    assert np.abs(psi_t.detach().cpu().numpy() - z["psi_t"]).max() <= psi_tol
    grad = amps.grad.cpu().numpy()
    err = np.abs(grad - z["grad"]).max() / np.abs(z["grad"]).max()
    print(f"  gradient relative error {err:.2e}")
    assert err <= grad_tol
