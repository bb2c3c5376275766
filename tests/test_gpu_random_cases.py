"""Seeded random configurations against the float64 C oracle: spin, integration
and exponentiation method, frame, number of squares, integration and output
steps (including ratios that are not whole and outputs shorter than a step),
start time, duration, sweep values and the launch shape of the time-evolution
kernel.  Every operator entry, state amplitude and spin component within 1e-9,
time stamps bit-exact."""
from __future__ import annotations

import math
import warnings

import numpy as np
import pytest

from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu
TOL = 1e-9
TAU = 2 * math.pi

# Generic fields: every entry depends on time and on the sweep values.
FIELD_HALF = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = s[0] * cos(s[1] * t + s[2]); f[1] = s[3] * sin(s[4] * t);
  f[2] = s[5] + s[6] * cos(s[7] * t); }
"""
FIELD_ONE = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = s[0] * cos(s[1] * t + s[2]); f[1] = s[3] * sin(s[4] * t);
  f[2] = s[5] + s[6] * cos(s[7] * t); f[3] = s[8] * cos(s[9] * t); }
"""
FIELD_EIGHT = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = s[0] * cos(s[1] * t + s[2]); f[1] = s[3] * sin(s[4] * t);
  f[2] = s[5] + s[6] * cos(s[7] * t); f[3] = s[8] * cos(s[9] * t);
  f[4] = s[8] * sin(s[1] * t); f[5] = 0.5 * s[3]; f[6] = s[0] * cos(s[4] * t); f[7] = -s[6]; }
"""
METHODS = {"magnus_cf4": "MAGNUS_CF4", "euler": "MIDPOINT_SAMPLE", "heun": "HALF_STEP"}
N_CASES = 36


def _case(k):
    rng = np.random.default_rng(1000 + k)
    spin = ("half", "one", "one")[k % 3]
    if spin == "half":
        expo = ("analytic", "lie_trotter")[rng.integers(2)]
    else:
        expo = ("lie_trotter", "lie_trotter_8")[int(rng.uniform() < 0.3)]
    method = list(METHODS)[rng.integers(3)]
    rotating = bool(rng.integers(2))
    squares = int(rng.choice([24, 24, 24, 0, 1, 7, 10]))
    dt = float(10 ** rng.uniform(-8.3, -6.7))
    ratio = int(rng.integers(1, 41))
    dt_out = dt * ratio
    if rng.uniform() < 0.2:
        dt_out *= rng.uniform(0.9, 1.1)          # not a whole number of steps
    if rng.uniform() < 0.1:
        dt_out = dt * rng.uniform(0.3, 0.6)      # output shorter than a step
    t0 = float(rng.uniform(-1e-4, 1e-4))
    n_out = int(rng.integers(0, 3000))
    t1 = t0 + dt_out * (n_out + rng.uniform(0, 0.999))
    amp = TAU * 10 ** rng.uniform(3, 5)
    sweep = [amp, TAU * 10 ** rng.uniform(3, 6), rng.uniform(0, TAU), amp * rng.uniform(0, 1),
             TAU * 10 ** rng.uniform(3, 5), TAU * 10 ** rng.uniform(3, 6),
             amp * rng.uniform(0, 0.3), TAU * 10 ** rng.uniform(2, 4),
             amp * rng.uniform(0, 0.5), TAU * 10 ** rng.uniform(3, 5)]
    lanes = [None, 0, -1, -2, int(rng.integers(1, 65))][rng.integers(5)]
    d = 2 if spin == "half" else 3
    psi0 = rng.normal(size=d) + 1j * rng.normal(size=d)
    psi0 /= np.linalg.norm(psi0)
    return dict(spin=spin, expo=expo, method=method, rotating=rotating, squares=squares, dt=dt,
                dt_out=dt_out, t0=t0, t1=t1, sweep=sweep, lanes=lanes, psi0=psi0)


@pytest.mark.parametrize("k", range(N_CASES))
def test_random_configuration_matches_oracle(k):
    import spinsim_amd as ss

    c = _case(k)
    print(f"case {k}: " + ", ".join(f"{n}={c[n]}" for n in ("spin", "expo", "method", "rotating",
                                                             "squares", "dt", "dt_out", "t0",
                                                             "t1", "lanes")), flush=True)
    src = (FIELD_HALF if c["spin"] == "half" else
           FIELD_EIGHT if c["expo"] == "lie_trotter_8" else FIELD_ONE)
    sq = ss.SpinQuantumNumber.HALF if c["spin"] == "half" else ss.SpinQuantumNumber.ONE
    sim = ss.Simulator(src, sq, exponentiation_method=getattr(ss.ExponentiationMethod,
                                                              c["expo"].upper()),
                       use_rotating_frame=c["rotating"],
                       integration_method=getattr(ss.IntegrationMethod, METHODS[c["method"]]),
                       number_of_squares=c["squares"])
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")   # the time-step warning, when the ratio is not whole
        res = sim.evaluate(c["t0"], c["t1"], c["dt"], c["dt_out"], c["psi0"], c["sweep"],
                           lanes_per_step=c["lanes"])
    dt, n, nf, _ = O.timing(c["t0"], c["t1"], c["dt"], c["dt_out"])
    co = COracle(src, sq.dimension, c["method"], c["expo"], c["rotating"], c["squares"])
    time, u = co.time_evolution(c["sweep"], float(c["t0"]), dt, float(c["dt_out"]), nf, n)
    state = co.get_state(c["psi0"], u)
    assert res.time.shape == (n,)
    assert np.array_equal(res.time, time)
    assert np.abs(res.time_evolution - u).max(initial=0) <= TOL
    assert np.abs(res.state - state).max(initial=0) <= TOL
    assert np.abs(res.spin - co.get_spin(state)).max(initial=0) <= TOL

