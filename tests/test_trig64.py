"""The field callbacks' table-driven float64 sin and cos (every precision but
fp32; ss_sin64 / ss_cos64 of ss_prelude.h), host build of the same text the
GPU modules compile, against numpy's float64 functions."""
from __future__ import annotations

import math

import numpy as np

from spinsim_amd import host


def test_float64_trig_accuracy():
    """Within 2.3e-16 absolute (about one ulp of 1) for |x| up to 1e14: the
    Cody-Waite reduction to a multiple of 2 pi / 256 keeps the phase to
    ~1e-18, the table entries are correctly rounded and the b^6 series leave
    out less than 2e-20; most points are exact."""
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-1e7, 1e7, 200000), rng.uniform(-10, 10, 50000),
                        np.arange(-2000, 2000) * (math.tau / 256),  # table nodes
                        np.arange(-2000, 2000) * (math.tau / 256) + math.tau / 512,  # cell edges
                        rng.uniform(-1e10, 1e10, 20000), rng.uniform(-1e14, 1e14, 20000),
                        np.array([0.0, -0.0, 1e-300, -5e-324, math.pi, math.pi / 2])])
    for fn, ref in (("sin", np.sin), ("cos", np.cos)):
        err = np.abs(host.f64_eval(fn, x) - ref(x))
        assert err.max() <= 2.3e-16, (fn, float(err.max()))
        assert np.mean(err == 0) > 0.5


def test_float64_trig_out_of_range_takes_library_functions():
    special = [math.nan, math.inf, -math.inf, 1e15, -3e16, 1e300]
    for fn, ref in (("sin", np.sin), ("cos", np.cos)):
        with np.errstate(all="ignore"):
            np.testing.assert_array_equal(host.f64_eval(fn, special), ref(np.array(special)))


def test_float64_trig_symmetry():
    """sin(-x) = -sin(x) and cos(-x) = cos(x) to the same bound (the reduction
    rounds k to nearest, symmetric about 0)."""
    x = np.random.default_rng(12).uniform(0, 1e6, 50000)
    assert np.abs(host.f64_eval("sin", -x) + host.f64_eval("sin", x)).max() <= 2.3e-16
    assert np.abs(host.f64_eval("cos", -x) - host.f64_eval("cos", x)).max() <= 2.3e-16
