"""The committed golden fixtures (tests/golden/*.npz, written from the float64
NumPy oracle by tests/golden/make_golden.py), pinned against what the
mathematics fixes -- Larmor's closed form, an adaptive ODE solve (DOP853,
rtol = atol = 1e-13), central finite differences, unitarity and norm -- and
reproduced by both oracle restatements.  The GPU tests hold the HIP path to
the same files."""
from __future__ import annotations

import importlib.util
import json
import math
import os

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import fields as F
from oracle import adjoint_reference as AR
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle
from test_oracle_pinning import hamiltonian

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = {f.name: f for f in (F.LARMOR, F.RABI_ONE, F.RABI_SWEEP_HALF, F.TILTED_HALF,
                              F.TILTED_ONE, F.TILTED_EIGHT)}
NAMES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))
TRAJECTORIES = [n for n in NAMES if n != "batch_rabi_half_4" and n != "adjoint_one_2"]


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        arrays = {k: z[k] for k in z.files}
    return json.loads(str(arrays.pop("meta"))), arrays


def _generator():
    spec = importlib.util.spec_from_file_location("make_golden",
                                                  os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _ode(rhs, psi0, t_eval):
    sol = solve_ivp(rhs, (0.0, float(t_eval[-1])), np.asarray(psi0, complex), method="DOP853",
                    t_eval=t_eval, rtol=1e-13, atol=1e-13)
    return sol.y.T


def _field_rhs(field, sweep, lie):
    def rhs(t, y):
        f = np.zeros((lie, 1))
        field.fn(np.array([t]), np.asarray(sweep), f)
        return -1j * hamiltonian(f[:, 0] if field.spin == "one" else f[:3, 0], field.spin) @ y
    return rhs


def test_every_fixture_is_committed():
    expected = {"larmor_half_1ms", "rabi_one_1ms", "eight_generator_one", "batch_rabi_half_4",
                "adjoint_one_2"}
    expected |= {f"methods_{m}_{s}_{fr}" for m in ("magnus_cf4", "euler", "heun")
                 for s in ("half", "one") for fr in ("rotating", "lab")}
    assert set(NAMES) == expected


@pytest.fixture(scope="module")
def regenerated():
    return dict(_generator().fixtures())


@pytest.mark.parametrize("name", NAMES)
def test_numpy_oracle_reproduces_fixture(name, regenerated):
    meta, arrays = load(name)
    new = regenerated[name]
    assert json.loads(str(new.pop("meta"))) == meta
    assert set(new) == set(arrays)
    for k, v in arrays.items():
        if v.dtype.kind in "iu":
            assert np.array_equal(new[k], v), k
        else:
            assert np.abs(new[k] - v).max(initial=0) <= 1e-12 * max(1.0, np.abs(v).max()), k


def _c_oracle_run(field, meta, psi0, sweep, t0, t1):
    dim = 2 if field.spin == "half" else 3
    co = COracle(field.src, dim, meta["method"], meta["expo"], meta["rotating"])
    dt, n, nf, _ = O.timing(t0, t1, meta["dt"], meta["dt_out"])
    time, u = co.time_evolution(sweep, float(t0), dt, float(meta["dt_out"]), nf, n)
    state = co.get_state(psi0, u)
    return time, u, state, co.get_spin(state)


@pytest.mark.parametrize("name", TRAJECTORIES)
def test_c_oracle_reproduces_trajectory(name):
    meta, z = load(name)
    field = FIELDS[meta["field"]]
    time, u, state, spin = _c_oracle_run(field, meta, z["psi0"], meta["sweep"], meta["t0"],
                                         meta["t1"])
    assert time.size == meta["n_rows"]
    r = z["rows"]
    assert np.array_equal(time[r], z["time"])
    for got, want in ((u[r], z["u"]), (state[r], z["state"]), (spin[r], z["spin"])):
        assert np.abs(got - want).max() <= 1e-10


def test_c_oracle_reproduces_batch():
    meta, z = load("batch_rabi_half_4")
    field = FIELDS[meta["field"]]
    for s in range(len(z["t1"])):
        a, b = z["offsets"][s], z["offsets"][s + 1]
        time, u, state, spin = _c_oracle_run(field, meta, z["psi0"][s], z["sweep"][s],
                                             z["t0"][s], z["t1"][s])
        assert time.size == b - a
        m = (z["rows"] >= a) & (z["rows"] < b)
        r = z["rows"][m] - a
        assert np.array_equal(time[r], z["time"][m])
        assert np.abs(state[r] - z["state"][m]).max() <= 1e-10
        assert np.abs(spin[r] - z["spin"][m]).max() <= 1e-10
        assert np.abs(u[r] - z["u"][m]).max() <= 1e-10


def test_larmor_fixture_is_the_closed_form():
    """psi(t) = (e^{-i pi 1000 t}, e^{i pi 1000 t}) / sqrt2, S = (cos, sin, 0) / 2,
    U = diag(e^{-i pi 1000 dt_out}, e^{i pi 1000 dt_out})."""
    meta, z = load("larmor_half_1ms")
    t = z["time"]
    assert np.array_equal(t, meta["dt_out"] * z["rows"].astype(float))
    exact = np.stack([np.exp(-1j * math.pi * 1000 * t), np.exp(1j * math.pi * 1000 * t)], 1)
    assert np.abs(z["state"] - exact / math.sqrt(2)).max() < 1e-12
    w = 2 * math.pi * 1000
    assert np.abs(z["spin"] - np.stack([np.cos(w * t), np.sin(w * t), 0 * t], 1) / 2).max() < 1e-12
    phase = np.exp(-1j * math.pi * 1000 * meta["dt_out"])
    assert np.abs(z["u"] - np.diag([phase, phase.conjugate()])).max() < 1e-12


# Bound on the fixture's distance from the ODE solution per method: fourth
# order at dt = 100 ns (or 20 ns) sits at 1e-10 - 1e-9, the Lie-Trotter floor
# over 1e4 steps at ~6e-7; the second-order methods at ~5e-5.  A wrong sample
# time, weight or frame term moves any of them by orders of magnitude.
ODE_TOL = {"magnus_cf4": 2e-9, "euler": 1e-4, "heun": 1e-4}


@pytest.mark.parametrize("name", [n for n in TRAJECTORIES if n != "larmor_half_1ms"])
def test_trajectory_fixture_matches_ode(name):
    meta, z = load(name)
    field = FIELDS[meta["field"]]
    lie = 8 if meta["expo"] == "lie_trotter_8" else 4
    ref = _ode(_field_rhs(field, meta["sweep"], lie), z["psi0"], z["time"])
    tol = 1e-6 if name == "rabi_one_1ms" else ODE_TOL[meta["method"]]
    assert np.abs(z["state"] - ref).max() < tol


@pytest.mark.parametrize("name", TRAJECTORIES + ["batch_rabi_half_4"])
def test_fixture_invariants(name):
    """Unitary step operators, unit-norm states, |S| = 1/2 for spin half."""
    meta, z = load(name)
    u = z["u"]
    eye = np.eye(u.shape[1])
    assert np.abs(np.conj(np.swapaxes(u, 1, 2)) @ u - eye).max() < 1e-12
    assert np.abs(np.linalg.norm(z["state"], axis=1) - 1).max() < 1e-12
    if meta["spin"] == "half":
        assert np.abs(np.linalg.norm(z["spin"], axis=1) - 0.5).max() < 1e-12


def test_batch_fixture_matches_ode():
    """Each simulation's first 100 us against the ODE solution.  CF4's own
    error from the 1.4 MHz counter-rotating term (0.44 rad per 50 ns step)
    grows linearly to 2e-8 - 6e-8 there, with the Rabi amplitude."""
    meta, z = load("batch_rabi_half_4")
    field = FIELDS[meta["field"]]
    for s in range(len(z["t1"])):
        a, b = z["offsets"][s], z["offsets"][s + 1]
        m = (z["rows"] >= a) & (z["rows"] < b) & (z["time"] <= 1e-4)
        ref = _ode(_field_rhs(field, z["sweep"][s], 4), z["psi0"][s], z["time"][m])
        assert np.abs(z["state"][m] - ref).max() < 1.5e-7


def _controlled(z, s, amp=None):
    amp = z["amp"][s] if amp is None else amp

    def z_bias(t, sweep, f):
        f[2] = 2 * math.pi * 1e5
    samplers = [(k, O.InterpolationSampler(z["tk"], amp[k], 3, True), 1.0) for k in range(2)]
    return AR.controlled_field(z_bias, samplers)


def test_adjoint_fixture_final_state_matches_ode():
    meta, z = load("adjoint_one_2")
    for s in range(len(z["amp"])):
        field = _controlled(z, s)

        def rhs(t, y):
            f = np.zeros((4, 1))
            field(np.array([t]), (0.0,), f)
            return -1j * hamiltonian(f[:, 0], "one") @ y
        ref = _ode(rhs, z["psi0"], np.array([meta["t1"]]))[-1]
        assert np.abs(z["psi_t"][s] - ref).max() < 2e-9


def test_adjoint_fixture_gradient_matches_finite_differences():
    """Central differences of the NumPy oracle's forward pass, 4 coordinates
    per simulation (the fixture's gradient comes from an expm_frechet walk)."""
    meta, z = load("adjoint_one_2")
    h = 1e-4 * 2 * math.pi * 5e3
    for s in range(len(z["amp"])):
        scale = np.abs(z["grad"][s]).max()
        for (k, j) in [(0, 3), (1, 10), (0, 17), (1, 0)]:
            loss = []
            for sign in (1, -1):
                amp = z["amp"][s].copy()
                amp[k, j] += sign * h
                _, u, _, _ = O.get_time_evolution(_controlled(z, s, amp), "one", 0.0, meta["t1"],
                                                  meta["dt"], meta["dt_out"])
                psi_t = u[-1] @ O.get_state(z["psi0"], u)[-1]
                loss.append(1 - abs(psi_t[-1]) ** 2)
            fd = (loss[0] - loss[1]) / (2 * h)
            assert abs(fd - z["grad"][s, k, j]) < 1e-6 * scale, (s, k, j)
