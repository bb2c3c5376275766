"""Adjoint gradients vs an exact float64 CPU reference and vs central finite
differences (state-transfer loss L = 1 - |psi(T)[last]|^2)."""
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


def _static_z_half(t, s, f):
    f[2] = TAU * 1e5


def _psi_grad(psi):
    g = np.zeros_like(psi)
    g[..., -1] = -2 * psi[..., -1]
    return g


@pytest.mark.parametrize("spin,rotating", [("one", True), ("one", False), ("half", True),
                                           ("half", False)])
def test_adjoint_matches_cpu_reference(spin, rotating):
    import torch

    import spinsim_amd as ss

    print(f"case: adjoint vs CPU reference spin={spin} rotating={rotating}", flush=True)
    rng = np.random.default_rng(17)
    b, m = 2, 11
    t1 = 10e-6
    tk = np.linspace(0.0, t1, m)
    amp = rng.normal(0, TAU * 8e4, (b, 2, m))
    static = _static_z_one if spin == "one" else _static_z_half
    sq = ss.SpinQuantumNumber.ONE if spin == "one" else ss.SpinQuantumNumber.HALF
    cs = ss.ControlledSimulator(sq, [ss.Control("x", tk), ss.Control("y", tk)], static,
                                use_rotating_frame=rotating)
    amps = torch.tensor(amp, device="cuda", requires_grad=True)
    psi0 = sq.plus_z
    psi_t = cs.final_state(amps, 0.0, t1, 20e-9, 1e-6, psi0)
    loss = (1 - psi_t[:, -1].abs() ** 2).sum()
    loss.backward()
    grad = amps.grad.cpu().numpy()
    for s in range(b):
        controls = [(0, tk, amp[s, 0], 3, True, 1.0), (1, tk, amp[s, 1], 3, True, 1.0)]
        psi_ref, g_ref = AR.final_state_and_gradient(static, spin, controls, 0.0, t1, 20e-9,
                                                     1e-6, psi0, _psi_grad, rotating=rotating)
        assert np.abs(psi_t[s].detach().cpu().numpy() - psi_ref).max() < 1e-9
        g_ref = np.stack(g_ref)
        scale = np.abs(g_ref).max()
        err = np.abs(grad[s] - g_ref).max() / scale
        print(f"  sim {s}: max |grad| {scale:.3e}, relative error {err:.3e}")
        assert err < 1e-8


def test_adjoint_matches_finite_differences():
    """Central differences of the GPU forward, all perturbations in one batch."""
    import torch

    import spinsim_amd as ss

    print("case: adjoint vs finite differences", flush=True)
    rng = np.random.default_rng(23)
    b, m, t1 = 3, 41, 40e-6
    tk = np.linspace(0.0, t1, m)
    sigma = TAU * 5e4
    amp = rng.normal(0, sigma, (b, 2, m))
    cs = ss.ControlledSimulator(ss.SpinQuantumNumber.ONE, [ss.Control("x", tk),
                                                           ss.Control("y", tk)], _static_z_one)
    psi0 = ss.SpinQuantumNumber.ONE.plus_z
    amps = torch.tensor(amp, device="cuda", requires_grad=True)
    psi_t = cs.final_state(amps, 0.0, t1, 20e-9, 1e-6, psi0)
    (1 - psi_t[:, -1].abs() ** 2).sum().backward()
    grad = amps.grad.cpu().numpy()
    coords = [(s, k, j) for s in range(b) for k in range(2) for j in rng.choice(m, 6, False)]
    h = 1e-4 * sigma
    rows = []
    for (s, k, j) in coords:
        for sign in (1, -1):
            a = amp[s].copy()
            a[k, j] += sign * h
            rows.append(a)
    batch = torch.tensor(np.stack(rows), device="cuda")
    with torch.no_grad():
        psi_p = cs.final_state(batch, 0.0, t1, 20e-9, 1e-6, psi0).cpu().numpy()
    loss_p = 1 - np.abs(psi_p[:, -1]) ** 2
    fd = (loss_p[0::2] - loss_p[1::2]) / (2 * h)
    ad = np.array([grad[s, k, j] for (s, k, j) in coords])
    rel = np.abs(fd - ad).max() / np.abs(ad).max()
    print(f"  {len(coords)} coordinates: max |grad| {np.abs(ad).max():.3e}, rel err {rel:.3e}")
    assert rel < 1e-5


def test_adjoint_zero_order_and_linear_controls_and_q_axis():
    import torch

    import spinsim_amd as ss

    print("case: adjoint zoh/linear/q", flush=True)
    rng = np.random.default_rng(29)
    t1 = 8e-6
    tk = np.linspace(-1e-6, t1 + 1e-6, 9)
    amp = rng.normal(0, TAU * 6e4, (1, 3, 9))
    controls = [ss.Control("x", tk, order=0), ss.Control("y", tk, order=1, zero_boundary=False),
                ss.Control("q", tk, order=3, scale=0.5)]
    cs = ss.ControlledSimulator(ss.SpinQuantumNumber.ONE, controls, _static_z_one)
    psi0 = ss.SpinQuantumNumber.ONE.plus_x
    amps = torch.tensor(amp, device="cuda", requires_grad=True)
    psi_t = cs.final_state(amps, 0.0, t1, 20e-9, 1e-6, psi0)
    (1 - psi_t[:, -1].abs() ** 2).sum().backward()
    grad = amps.grad.cpu().numpy()[0]
    ref_controls = [(0, tk, amp[0, 0], 0, True, 1.0), (1, tk, amp[0, 1], 1, False, 1.0),
                    (3, tk, amp[0, 2], 3, True, 0.5)]
    psi_ref, g_ref = AR.final_state_and_gradient(_static_z_one, "one", ref_controls, 0.0, t1,
                                                 20e-9, 1e-6, psi0, _psi_grad)
    assert np.abs(psi_t[0].detach().cpu().numpy() - psi_ref).max() < 1e-9
    g_ref = np.stack(g_ref)
    err = np.abs(grad - g_ref).max() / np.abs(g_ref).max()
    print(f"  relative error {err:.3e}")
    assert err < 1e-8


@pytest.mark.parametrize("k", range(8))
def test_random_adjoint_configuration(k):
    """Seeded random controls (1-3, the first x or y, the others on any axis,
    orders 0/1/3, either boundary
    mode, scales, non-uniform knots), spins, methods and frames: gradients
    against central differences of the float64 forward (every method; a z
    control also moves the rotating frequency, which the gradient follows)
    and, for Magnus CF4, against the CPU reference gradient."""
    import torch

    import spinsim_amd as ss

    rng = np.random.default_rng(900 + k)
    spin = ("one", "half")[k % 2]
    method = ("MAGNUS_CF4", "MIDPOINT_SAMPLE", "HALF_STEP")[(k // 2) % 3]
    rotating = bool(rng.integers(2))
    t1 = float(rng.uniform(4e-6, 12e-6))
    dt = float(rng.choice([10e-9, 20e-9, 40e-9]))
    axes = ["x", "y", "z"] + (["q"] if spin == "one" else [])
    n_ctl = int(rng.integers(1, 4))
    controls, ref_controls, widths = [], [], []
    for c in range(n_ctl):
        m = int(rng.integers(5, 25))
        tk = np.sort(rng.uniform(-1e-6, t1 + 1e-6, m))
        order = int(rng.choice([0, 1, 3]))
        zb = bool(rng.integers(2))
        scale = float(rng.uniform(0.5, 1.5))
        # the first control is transverse, so the loss depends on the amplitudes
        ax = axes[rng.integers(2 if c == 0 else len(axes))]
        controls.append(ss.Control(ax, tk, order=order, zero_boundary=zb, scale=scale))
        widths.append(m)
        ref_controls.append((controls[-1].axis, tk, order, zb, scale))
    print(f"case {k}: spin={spin} {method} rotating={rotating} dt={dt} t1={t1:.2e} controls="
          + ", ".join(f"({c.axis},{c.order},{int(c.zero_boundary)},{len(c.time)})" for c in controls),
          flush=True)
    sq = ss.SpinQuantumNumber.ONE if spin == "one" else ss.SpinQuantumNumber.HALF
    static = _static_z_one if spin == "one" else _static_z_half
    cs = ss.ControlledSimulator(sq, controls, static, use_rotating_frame=rotating,
                                integration_method=getattr(ss.IntegrationMethod, method))
    b = 2
    amp = [rng.normal(0, TAU * 4e4, (b, w)) for w in widths]
    psi0 = sq.plus_x
    amps = [torch.tensor(a, device="cuda", requires_grad=True) for a in amp]
    psi_t = cs.final_state(amps, 0.0, t1, dt, 1e-6, psi0)
    (1 - psi_t[:, -1].abs() ** 2).sum().backward()
    grad = [a.grad.cpu().numpy() for a in amps]
    scale = max(np.abs(g).max() for g in grad)
    # central differences, all perturbations in one batch
    h = 1e-4 * TAU * 4e4
    coords = [(s, c, int(j)) for s in range(b) for c in range(n_ctl)
              for j in rng.choice(widths[c], min(3, widths[c]), False)]
    batch = []
    for (s, c, j) in coords:
        for sign in (1, -1):
            row = [amp[cc][s].copy() for cc in range(n_ctl)]
            row[c][j] += sign * h
            batch.append(row)
    with torch.no_grad():
        pert = [torch.tensor(np.stack([r[cc] for r in batch]), device="cuda")
                for cc in range(n_ctl)]
        psi_p = cs.final_state(pert, 0.0, t1, dt, 1e-6, psi0).cpu().numpy()
    loss_p = 1 - np.abs(psi_p[:, -1]) ** 2
    fd = (loss_p[0::2] - loss_p[1::2]) / (2 * h)
    ad = np.array([grad[c][s, j] for (s, c, j) in coords])
    rel = np.abs(fd - ad).max() / scale
    print(f"  max |grad| {scale:.3e}, vs central differences {rel:.2e}")
    assert rel < 1e-7
    if method == "MAGNUS_CF4":
        for s in range(b):
            ctl = [(ax, tk, amp[c][s], order, zb, sc)
                   for c, (ax, tk, order, zb, sc) in enumerate(ref_controls)]
            psi_ref, g_ref = AR.final_state_and_gradient(static, spin, ctl, 0.0, t1, dt, 1e-6,
                                                         psi0, _psi_grad, rotating=rotating)
            assert np.abs(psi_t[s].detach().cpu().numpy() - psi_ref).max() < 1e-9
            err = max(np.abs(grad[c][s] - g_ref[c]).max() for c in range(n_ctl)) / scale
            print(f"  sim {s}: vs CPU reference gradient {err:.2e}")
            assert err < 1e-8
