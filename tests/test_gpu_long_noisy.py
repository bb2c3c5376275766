"""The long noisy spin-half run (1e7 CF4 steps): plain float64 against the C
oracle at 1e-9, float32-bulk error and unitarity drift stated."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu


def test_long_noisy_run():
    import torch

    import bench
    import spinsim_amd as ss

    print("case: long noisy run 1e7 steps", flush=True)
    sweep, grid, noise = bench.long_noisy_inputs()
    sampler = ss.generate_interpolation_sampler(grid, noise)
    field = bench.make_long_noisy_field(sampler)
    psi0 = ss.SpinQuantumNumber.HALF.plus_z
    res = {}
    for prec in ("fp64", "fp32_bulk"):
        sim = ss.Simulator(field, ss.SpinQuantumNumber.HALF, precision=prec)
        res[prec] = sim.evaluate(0.0, 0.1, 10e-9, 1e-6, psi0, sweep)
    src, _, _ = ss.transpile(field)
    co = COracle(src, 2, sampler_names=("noise_table",))
    co.set_samplers([(grid, noise, 3, True)])
    dt, n, nf, _ = O.timing(0.0, 0.1, 10e-9, 1e-6)
    assert (n, nf) == (100000, 100)
    t, u = co.time_evolution(sweep, 0.0, dt, 1e-6, nf, n)
    st = co.get_state(psi0, u)
    r = res["fp64"]
    assert np.array_equal(r.time, t)
    err64 = np.abs(r.state - st).max()
    assert np.abs(r.time_evolution - u).max() <= 1e-9
    assert err64 <= 1e-9
    err32 = np.abs(res["fp32_bulk"].state - st).max()
    uu = res["fp32_bulk"].time_evolution
    drift32 = np.abs(np.einsum("nji,njk->nik", uu.conj(), uu) - np.eye(2)).max()
    uu = r.time_evolution
    drift64 = np.abs(np.einsum("nji,njk->nik", uu.conj(), uu) - np.eye(2)).max()
    print(f"  fp64 state err {err64:.3e} drift {drift64:.3e}; fp32_bulk state err {err32:.3e} "
          f"drift {drift32:.3e}")
    assert err32 <= 1e-6
    assert drift32 <= 1e-9 and drift64 <= 1e-12
