"""precision="fp32": float32-bulk step operators plus the field function's
sin, cos and exp in float32 on arguments reduced in float64 -- the host build
of those functions against float64 numpy, and CPU-device runs against the
float64 oracle (the error stated per case)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import fields as F
import spinsim_amd as ss
from oracle import spinsim_oracle as O
from spinsim_amd import _native, host

from test_cpu_device import run_oracle


def test_float32_transcendentals_accuracy():
    """|sin32 - sin|, |cos32 - cos| <= 3e-11 absolute for |x| up to 1e10 (the
    float64 reduction and table keep the phase exact; float32 carries only
    cos b - 1 and sin b / b - 1 of the small remainder b), exp32 within 5e-12
    relative over [-708, 709] (float32 carries b^2 q of exp b = 1 + b + b^2 q);
    0 below -746 as float64 gives; non-finite and out-of-range arguments take
    the float64 functions."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-1e7, 1e7, 200000), rng.uniform(-10, 10, 50000),
                        np.arange(-2000, 2000) * math.tau, rng.uniform(-1e10, 1e10, 5000)])
    huge = rng.uniform(-1e14, 1e14, 5000)  # the reduction quotient itself rounds here
    for fn, ref in (("sin", np.sin), ("cos", np.cos)):
        err = np.abs(host.f32_eval(fn, x) - ref(x))
        assert err.max() <= 3e-11, (fn, err.max())
        assert np.median(err) <= 1e-12
        assert np.abs(host.f32_eval(fn, huge) - ref(huge)).max() <= 1e-9
    xe = np.concatenate([rng.uniform(-708, 709, 200000), rng.uniform(-5, 5, 50000)])
    rel = np.abs(host.f32_eval("exp", xe) / np.exp(xe) - 1)
    assert rel.max() <= 5e-12
    assert (host.f32_eval("exp", [-746.5, -1e9, -1e300]) == 0).all()
    for fn, ref, extra in (("sin", np.sin, [1e16, -3e15]), ("cos", np.cos, [1e16, -3e15]),
                           ("exp", np.exp, [800.0, -800.0, -720.0])):
        special = [math.nan, math.inf, -math.inf] + extra
        with np.errstate(all="ignore"):
            np.testing.assert_allclose(host.f32_eval(fn, special), ref(np.array(special)),
                                       rtol=1e-15, atol=0, equal_nan=True)


@pytest.mark.parametrize("field,dt,bound", [(F.RABI_SWEEP_HALF, 5e-8, 1e-6),
                                            (F.RABI_ONE, 1e-7, 1e-6)], ids=lambda v: getattr(v, "name", str(v)))
def test_cpu_device_fp32_against_oracle(field, dt, bound):
    """2 ms of the Rabi drives (cosine carriers through cos32) in precision fp32
    on the CPU device: within the bound of the float64 oracle, and not equal
    to the float32-bulk run (the float32 transcendentals are in use)."""
    spin = ss.SpinQuantumNumber.HALF if field.spin == "half" else ss.SpinQuantumNumber.ONE
    psi0 = O.EIGENSTATES[field.spin]["plus_z"]
    args = (0.0, 2e-3, dt, 1e-6, psi0, field.sweep)
    res = ss.Simulator(field.src, spin, device=ss.Device.CPU, precision="fp32").evaluate(*args)
    bulk = ss.Simulator(field.src, spin, device=ss.Device.CPU,
                        precision="fp32_bulk").evaluate(*args)
    _, _, ref, _ = run_oracle(field, 0.0, 2e-3, dt, 1e-6, psi0)
    err = float(np.abs(res.state - ref).max())
    print(f"{field.name}: fp32 state error {err:.3e}, "
          f"fp32_bulk {float(np.abs(bulk.state - ref).max()):.3e}")
    assert err <= bound
    assert not np.array_equal(res.state, bulk.state)


def test_fp32_compiles_for_gfx950():
    for dim, src in ((2, F.RABI_SWEEP_HALF.src), (3, F.RABI_ONE.src)):
        code, _ = _native.compile_code(src, dim, 0, 1 if dim == 3 else 0, True, 24, False, 3)
        assert code[:4] == b"\x7fELF"
