"""precision="fp32" on the GPU (float32-bulk step operators plus the field
function's sin, cos and exp in float32 on float64-reduced arguments) against
the float64 oracle, the error stated per case."""
from __future__ import annotations

import math

import numpy as np
import pytest

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("field,t1,dt,dt_out", [
    (F.RABI_SWEEP_HALF, 20e-3, 5e-8, 1e-6),   # a carrier commensurate with the step
    (F.RABI_ONE, 100e-3, 1e-7, 1e-6),         # the full spin-one Rabi workload
], ids=["rabi_half_20ms", "rabi_one_100ms"])
def test_fp32_against_oracle(field, t1, dt, dt_out):
    import spinsim_amd as ss

    print(f"case: fp32 {field.name} t1={t1}", flush=True)
    spin = ss.SpinQuantumNumber.HALF if field.spin == "half" else ss.SpinQuantumNumber.ONE
    psi0 = O.EIGENSTATES[field.spin]["plus_z"]
    res = ss.Simulator(field.src, spin, precision="fp32").evaluate(0.0, t1, dt, dt_out, psi0,
                                                                  field.sweep)
    bulk = ss.Simulator(field.src, spin, precision="fp32_bulk").evaluate(0.0, t1, dt, dt_out,
                                                                        psi0, field.sweep)
    co = COracle(field.src, spin.dimension)
    dt2, n, nf, _ = O.timing(0.0, t1, dt, dt_out)
    _, u = co.time_evolution(field.sweep, 0.0, dt2, dt_out, nf, n)
    ref = co.get_state(psi0, u)
    err = float(np.abs(res.state - ref).max())
    err_bulk = float(np.abs(bulk.state - ref).max())
    print(f"  state error fp32 {err:.3e}, fp32_bulk {err_bulk:.3e}, "
          f"fp32 - fp32_bulk {float(np.abs(res.state - bulk.state).max()):.3e}")
    assert err <= 1e-6
    assert not np.array_equal(res.state, bulk.state)


def test_fp32_long_noisy_field():
    """Gaussian pulses (exp), a cosine carrier and a cubic noise table, 1e6
    steps: fp32 against the oracle."""
    import spinsim_amd as ss

    print("case: fp32 pulses + noise", flush=True)
    rng = np.random.default_rng(20261203)
    amp, centre, width = rng.uniform(1e3, 1e4, 4), rng.uniform(0, 0.01, 4), rng.uniform(5e-6, 5e-5, 4)
    grid = np.arange(10001) * 1e-6
    noise = rng.normal(0, 2 * math.pi * 50, grid.size)
    sampler = ss.generate_interpolation_sampler(grid, noise)
    sweep = np.concatenate([amp, centre, width])

    def get_field(time_sample, sweep_parameters, field_sample):
        z = 7e5
        for k in range(4):
            d = time_sample - sweep_parameters[4 + k]
            z += sweep_parameters[k] * math.exp(-d * d / (2 * sweep_parameters[8 + k] ** 2))
        field_sample[2] = 2 * math.pi * z + sampler(time_sample)
        field_sample[0] = 2 * math.pi * 1e3 * math.cos(2 * math.pi * 7e5 * time_sample)
        field_sample[1] = 0

    psi0 = ss.SpinQuantumNumber.HALF.plus_z
    res = ss.Simulator(get_field, ss.SpinQuantumNumber.HALF, precision="fp32").evaluate(
        0.0, 0.01, 1e-8, 1e-6, psi0, sweep)
    src, _, _ = ss.transpile(get_field)
    co = COracle(src, 2, sampler_names=("unused",))
    co.set_samplers([(grid, noise, 3, True)])
    dt2, n, nf, _ = O.timing(0.0, 0.01, 1e-8, 1e-6)
    _, u = co.time_evolution(sweep, 0.0, dt2, 1e-6, nf, n)
    ref = co.get_state(psi0, u)
    err = float(np.abs(res.state - ref).max())
    print(f"  state error {err:.3e}")
    assert err <= 1e-6
