"""The drop-in API as the reference's example scripts use it (example.py,
interpolator_example.py): numpy results, lazy spin, sweeps, host-callable
samplers, the kernel-level get_time_evolution, and device outputs."""
from __future__ import annotations

import math

import numpy as np
import pytest

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu


def get_field_larmor(time_sample, sweep_parameters, field_sample):
    field_sample[0] = 0
    field_sample[1] = 0
    field_sample[2] = math.tau * 1000


def get_field_rabi(time_sample, sweep_parameters, field_sample):
    field_sample[0] = 2 * sweep_parameters[1] * math.cos(sweep_parameters[0] * time_sample)
    field_sample[1] = 0
    field_sample[2] = sweep_parameters[0]
    field_sample[3] = 0


RABI_SRC = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = 2 * s[1] * cos(s[0] * t); f[1] = 0; f[2] = s[0]; f[3] = 0; }
"""


def test_example_scripts():
    import spinsim_amd as spinsim

    print("case: example.py usage", flush=True)
    sim = spinsim.Simulator(get_field_larmor, spinsim.SpinQuantumNumber.HALF)
    res = sim.evaluate(0e-3, 100e-3, 100e-9, 500e-9, spinsim.SpinQuantumNumber.HALF.plus_x)
    assert isinstance(res.time, np.ndarray) and res.time.shape == (200000,)
    assert res.time_evolution.dtype == np.complex128 and res.time_evolution.shape == (200000, 2, 2)
    assert "spin" not in res.__dict__
    spin = res.spin
    assert "spin" in res.__dict__ and spin.shape == (200000, 3)
    w = math.tau * 1000
    assert np.abs(spin[:, 0] - 0.5 * np.cos(w * res.time)).max() < 1e-9

    sim = spinsim.Simulator(get_field_rabi, spinsim.SpinQuantumNumber.ONE)
    co = COracle(RABI_SRC, 3)
    dt, n, nf, _ = O.timing(0e-3, 100e-3, 100e-9, 500e-9)
    for sweep in ([math.tau * 20e3, math.tau * 1e3], [math.tau * 40e3, math.tau * 1e3]):
        res = sim.evaluate(0e-3, 100e-3, 100e-9, 500e-9, spinsim.SpinQuantumNumber.ONE.plus_z,
                           sweep)
        t, u = co.time_evolution(sweep, 0.0, dt, 500e-9, nf, n)
        st = co.get_state(O.EIGENSTATES["one"]["plus_z"], u)
        assert np.abs(res.state - st).max() <= 1e-9
        assert np.abs(res.spin - co.get_spin(st)).max() <= 1e-9


def test_interpolator_example_usage():
    import spinsim_amd as spinsim

    print("case: interpolator_example.py usage", flush=True)
    rng = np.random.default_rng(0)
    time_sampled = np.arange(0, 100e-3, 1e-3)
    larmor_sampled = rng.normal(size=time_sampled.size)
    larmor_interpolator = spinsim.generate_interpolation_sampler(time_sampled, larmor_sampled)

    def get_field_larmor_interpolator(time_sample, user_parameters, field_sample):
        noise_amplitude = user_parameters[0]
        field_sample[2] = math.tau * (200 + noise_amplitude * larmor_interpolator(time_sample))
        field_sample[0] = 0
        field_sample[1] = 0
        field_sample[3] = 0

    sim = spinsim.Simulator(get_field_larmor_interpolator, spinsim.SpinQuantumNumber.ONE)
    quiet = sim.evaluate(0, 100e-3, 1e-6, 10e-6, spinsim.SpinQuantumNumber.ONE.plus_x, [0])
    noisy = sim.evaluate(0, 100e-3, 1e-6, 10e-6, spinsim.SpinQuantumNumber.ONE.plus_x, [50])
    assert quiet.spin.shape == noisy.spin.shape == (10000, 3)
    # quiet: Larmor at 200 Hz, spin-one plus_x -> <Jx> = cos(2 pi 200 t)
    assert np.abs(quiet.spin[:, 0] - np.cos(math.tau * 200 * quiet.time)).max() < 1e-8
    # host-callable sampler (the reference's Device.CPU sampler used for plotting)
    plot = spinsim.generate_interpolation_sampler(time_sampled, larmor_sampled,
                                                  device=spinsim.Device.CPU)
    time_plot = np.arange(0, 100e-3, 100e-6)
    ref = O.InterpolationSampler(time_sampled, larmor_sampled)
    assert np.abs(plot(time_plot) - ref(time_plot)).max() < 1e-12
    assert abs(plot(0.0505) - ref(0.0505)) < 1e-12


def test_kernel_level_entry_points_and_device_outputs():
    import torch

    import spinsim_amd as spinsim

    print("case: kernel-level entry points", flush=True)
    sim = spinsim.Simulator(F.TILTED_ONE.src, spinsim.SpinQuantumNumber.ONE)
    dt, n, nf, _ = O.timing(0, 3e-4, 1e-7, 1e-6)
    time_coarse = np.empty(n)
    u = np.empty((n, 3, 3), np.complex128)
    sim.get_time_evolution(np.array([0.0]), time_coarse, np.array([0.0, 3e-4]), dt, 1e-6, u)
    co = COracle(F.TILTED_ONE.src, 3)
    t_ref, u_ref = co.time_evolution([0.0], 0.0, dt, 1e-6, nf, n)
    assert np.array_equal(time_coarse, t_ref) and np.abs(u - u_ref).max() <= 1e-9
    res = sim.evaluate(0, 3e-4, 1e-7, 1e-6, O.EIGENSTATES["one"]["plus_y"], output="torch")
    assert isinstance(res.state, torch.Tensor) and res.state.is_cuda
    st_ref = co.get_state(O.EIGENSTATES["one"]["plus_y"], u_ref)
    assert np.abs(res.state.cpu().numpy() - st_ref).max() <= 1e-9
    assert isinstance(res.spin, torch.Tensor)
    assert np.abs(sim.spin_calculator(st_ref) - co.get_spin(st_ref)).max() <= 1e-12
    info = sim.module_info()
    assert info["vgprs"] > 0 and info["scratch_bytes"] == 0


def test_numpy_results_staged_through_pinned_memory():
    """evaluate() numpy results (pinned staging, copies overlapping the scan)
    equal the device results bit for bit, and later calls never overwrite
    earlier results; above the staging limit the plain copy path gives the
    same arrays."""
    import torch

    import spinsim_amd as ss
    from spinsim_amd import simulator as S

    print("case: pinned numpy results", flush=True)
    sim = ss.Simulator(F.TILTED_ONE.src, ss.SpinQuantumNumber.ONE)
    args = (0.0, 3e-4, 1e-7, 1e-6)
    dev = sim.evaluate(*args, O.EIGENSTATES["one"]["plus_x"], output="torch")
    a = sim.evaluate(*args, O.EIGENSTATES["one"]["plus_x"])
    keep = a.state.copy()
    for name in ("time", "time_evolution", "state"):
        assert np.array_equal(getattr(a, name), getattr(dev, name).cpu().numpy()), name
    b = sim.evaluate(*args, O.EIGENSTATES["one"]["minus_z"])
    torch.cuda.synchronize()
    assert np.array_equal(a.state, keep)
    assert not np.array_equal(a.state, b.state)
    assert np.array_equal(a.spin, dev.spin.cpu().numpy())
    limit = S.PINNED_RESULT_BYTES
    try:
        S.PINNED_RESULT_BYTES = 0
        c = sim.evaluate(*args, O.EIGENSTATES["one"]["plus_x"])
    finally:
        S.PINNED_RESULT_BYTES = limit
    for name in ("time", "time_evolution", "state"):
        assert np.array_equal(getattr(a, name), getattr(c, name)), name
    empty = sim.evaluate(0.0, 0.0, 1e-7, 1e-6, O.EIGENSTATES["one"]["plus_x"])
    assert empty.state.shape == (0, 3)


def test_default_device_is_rocm_with_a_hip_device():
    """With a HIP device visible the default device is ROCm (the GPU tests
    that construct default simulators run the HIP kernels, never the host
    build)."""
    import spinsim_amd as ss

    sim = ss.Simulator(F.LARMOR.src, ss.SpinQuantumNumber.HALF)
    assert sim.device == ss.Device.ROCM and not sim._host
