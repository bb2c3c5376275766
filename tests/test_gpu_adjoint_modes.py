"""Adjoint gradient reduction modes, and the largest control count.

* deterministic=True: every gradient entry summed in row order -- two runs
  give identical bits, and agree with the atomic mode to float64 rounding;
  a coarse step spanning more knots than the 8-knot window is refused.
* Four controls (x, y, z, q on spin one; the largest LDS window launch)
  against the float64 CPU reference gradient and central differences, in
  both modes.
* A control whose axis the module was not compiled for is refused.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import adjoint_reference as AR

pytestmark = pytest.mark.gpu
TAU = 2 * math.pi


def _static_z_one(t, s, f):
    f[2] = TAU * 1e5
    f[3] = TAU * 2e3


def _psi_grad(psi):
    g = np.zeros_like(psi)
    g[..., -1] = -2 * psi[..., -1]
    return g


def _grad(cs, amp, t1, psi0, deterministic):
    import torch

    a = torch.tensor(amp, device="cuda", requires_grad=True)
    psi = cs.final_state(a, 0.0, t1, 20e-9, 1e-6, psi0, deterministic=deterministic)
    (1 - psi[:, -1].abs() ** 2).sum().backward()
    return a.grad.cpu().numpy(), psi.detach().cpu().numpy()


def test_deterministic_gradients_repeat_bit_for_bit():
    import spinsim_amd as ss

    print("case: deterministic adjoint, repeated runs", flush=True)
    rng = np.random.default_rng(101)
    b, m, t1 = 64, 201, 100e-6
    tk = np.linspace(0.0, t1, m)
    amp = rng.normal(0, TAU * 5e4, (b, 2, m))
    cs = ss.ControlledSimulator(ss.SpinQuantumNumber.ONE, [ss.Control("x", tk),
                                                           ss.Control("y", tk)], _static_z_one)
    psi0 = ss.SpinQuantumNumber.ONE.plus_z
    g1, _ = _grad(cs, amp, t1, psi0, True)
    g2, _ = _grad(cs, amp, t1, psi0, True)
    ga, _ = _grad(cs, amp, t1, psi0, False)
    assert np.array_equal(g1, g2)
    scale = np.abs(ga).max()
    err = np.abs(g1 - ga).max() / scale
    print(f"  max |grad| {scale:.3e}; deterministic vs atomic {err:.3e} relative")
    assert err <= 1e-13


def test_deterministic_refuses_windows_it_cannot_hold():
    import spinsim_amd as ss

    print("case: deterministic adjoint, knots denser than the window", flush=True)
    rng = np.random.default_rng(103)
    t1 = 20e-6
    tk = np.linspace(0.0, t1, 401)  # 20 knots per 1 us coarse step
    amp = rng.normal(0, TAU * 5e4, (2, 1, tk.size))
    cs = ss.ControlledSimulator(ss.SpinQuantumNumber.ONE, [ss.Control("x", tk)], _static_z_one)
    psi0 = ss.SpinQuantumNumber.ONE.plus_z
    with pytest.raises(RuntimeError, match="8-knot gradient window"):
        _grad(cs, amp, t1, psi0, True)
    g, psi = _grad(cs, amp, t1, psi0, False)  # the atomic mode takes any knot spacing
    assert np.isfinite(g).all() and np.abs(g).max() > 0
    # 20 intervals per coarse step: the moment sums flush at every knot
    for s in range(2):
        psi_ref, g_ref = AR.final_state_and_gradient(_static_z_one, "one", [(0, tk, amp[s, 0], 3, True, 1.0)],
                                                     0.0, t1, 20e-9, 1e-6, psi0, _psi_grad)
        assert np.abs(psi[s] - psi_ref).max() < 1e-9
        err = np.abs(g[s, 0] - g_ref[0]).max() / np.abs(g_ref[0]).max()
        print(f"  sim {s}: dense knots, relative error vs CPU reference {err:.3e}")
        assert err < 1e-8


@pytest.mark.parametrize("deterministic", [False, True])
def test_four_controls_spin_one(deterministic):
    """x, y, z and q controls at once (4 LDS windows per lane)."""
    import torch

    import spinsim_amd as ss

    print(f"case: four controls, deterministic={deterministic}", flush=True)
    rng = np.random.default_rng(107)
    t1 = 10e-6
    tk = np.linspace(-1e-6, t1 + 1e-6, 13)
    amp = rng.normal(0, TAU * 5e4, (2, 4, tk.size))
    scales = (1.0, 0.8, 0.3, 0.5)
    controls = [ss.Control(ax, tk, scale=sc) for ax, sc in zip("xyzq", scales)]
    cs = ss.ControlledSimulator(ss.SpinQuantumNumber.ONE, controls, _static_z_one)
    psi0 = ss.SpinQuantumNumber.ONE.plus_x
    grad, psi = _grad(cs, amp, t1, psi0, deterministic)
    for s in range(2):
        ref_controls = [(ax, tk, amp[s, k], 3, True, sc)
                        for k, (ax, sc) in enumerate(zip((0, 1, 2, 3), scales))]
        psi_ref, g_ref = AR.final_state_and_gradient(_static_z_one, "one", ref_controls, 0.0, t1,
                                                     20e-9, 1e-6, psi0, _psi_grad)
        assert np.abs(psi[s] - psi_ref).max() < 1e-9
        g_ref = np.stack(g_ref)
        err = np.abs(grad[s] - g_ref).max() / np.abs(g_ref).max()
        print(f"  sim {s}: relative error vs CPU reference {err:.3e}")
        assert err < 1e-8
    # central differences on a few coordinates of every control
    h = 1e-4 * TAU * 5e4
    coords = [(s, k, int(j)) for s in range(2) for k in range(4) for j in rng.choice(tk.size, 2)]
    rows = []
    for s, k, j in coords:
        for sign in (1, -1):
            a = amp[s].copy()
            a[k, j] += sign * h
            rows.append(a)
    with torch.no_grad():
        psi_p = cs.final_state(torch.tensor(np.stack(rows), device="cuda"), 0.0, t1, 20e-9, 1e-6,
                               psi0).cpu().numpy()
    loss = 1 - np.abs(psi_p[:, -1]) ** 2
    fd = (loss[0::2] - loss[1::2]) / (2 * h)
    ad = np.array([grad[s, k, j] for s, k, j in coords])
    rel = np.abs(fd - ad).max() / np.abs(ad).max()
    print(f"  {len(coords)} coordinates vs central differences: {rel:.3e}")
    assert rel < 1e-5


def test_control_axis_outside_compiled_set_is_refused():
    """A module compiled for x and y controls (SS_ADJ_AXES = 3) refuses a
    descriptor naming z at launch: every row returns early with status bit 4."""
    import torch

    import spinsim_amd as ss

    print("case: control axis outside the compiled set", flush=True)
    rng = np.random.default_rng(109)
    t1 = 5e-6
    tk = np.linspace(0.0, t1, 11)
    amp = rng.normal(0, TAU * 5e4, (2, 2, tk.size))
    cs = ss.ControlledSimulator(ss.SpinQuantumNumber.ONE, [ss.Control("x", tk), ss.Control("y", tk)],
                                _static_z_one)
    psi0 = ss.SpinQuantumNumber.ONE.plus_z
    g, _ = _grad(cs, amp, t1, psi0, False)  # compiles the module for axes {x, y}
    assert np.isfinite(g).all()
    cs.controls[1].axis = 2  # the descriptors now name z; the module stays compiled for x, y
    try:
        with pytest.raises(RuntimeError, match="SS_ADJ_AXES"):
            _grad(cs, amp, t1, psi0, False)
    finally:
        cs.controls[1].axis = 1
    torch.cuda.synchronize()


@pytest.mark.parametrize("mixed", [False, True], ids=["all_small", "small_and_large"])
def test_small_norm_series_against_reference(mixed):
    """Steps whose generator norm bound is below 0.004 on every lane of a wave
    take one piece with shorter series (Taylor order 5, Frechet order 4);
    waves holding a larger step take the general path.  Weak controls on
    every simulation (both waves small), and weak controls on the first
    seven of twelve with strong ones after (10 rows a simulation, one thread
    a row: rows 0-63 small, rows 64-119 with large steps), against the
    float64 CPU reference gradient."""
    print(f"case: small-norm adjoint series, mixed={mixed}", flush=True)
    rng = np.random.default_rng(113)
    b, t1 = 12, 10e-6
    tk = np.linspace(0.0, t1, 21)
    sig = np.full(b, TAU * 5e3)
    if mixed:
        sig[7:] = TAU * 2e5  # nu ~ 0.01-0.05 per step: rows 64-119 (wave 1) take the general path
    amp = rng.normal(0, 1, (b, 2, tk.size)) * sig[:, None, None]
    import spinsim_amd as ss

    cs = ss.ControlledSimulator(ss.SpinQuantumNumber.ONE, [ss.Control("x", tk), ss.Control("y", tk)],
                                _static_z_one)
    psi0 = ss.SpinQuantumNumber.ONE.plus_z
    grad, psi = _grad(cs, amp, t1, psi0, False)
    for s in range(b):
        psi_ref, g_ref = AR.final_state_and_gradient(
            _static_z_one, "one", [(0, tk, amp[s, 0], 3, True, 1.0), (1, tk, amp[s, 1], 3, True, 1.0)],
            0.0, t1, 20e-9, 1e-6, psi0, _psi_grad)
        assert np.abs(psi[s] - psi_ref).max() < 1e-9
        g_ref = np.stack(g_ref)
        err = np.abs(grad[s] - g_ref).max() / np.abs(g_ref).max()
        assert err < 1e-8, (s, err)
    print(f"  {b} simulations within 1e-8 of the CPU reference gradient")
