"""Pin the CPU reference gradient against central differences of the oracle's
own forward pass (independent of every derivative formula)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import adjoint_reference as AR
from oracle import spinsim_oracle as O

TAU = 2 * math.pi


def _static(t, s, f):
    f[2] = TAU * 1e5
    f[3] = TAU * 3e3


@pytest.mark.parametrize("spin,rotating", [("one", True), ("half", False)])
def test_reference_gradient_matches_finite_differences(spin, rotating):
    rng = np.random.default_rng(3)
    t1, m = 6e-6, 7
    tk = np.linspace(0, t1, m)
    amp = rng.normal(0, TAU * 8e4, (2, m))
    controls = [(0, tk, amp[0], 3, True, 1.0), (1, tk, amp[1], 1, True, 0.7)]
    psi0 = O.EIGENSTATES[spin]["plus_z"]

    def loss(psi):
        return 1 - abs(psi[-1]) ** 2

    def loss_grad(psi):
        g = np.zeros_like(psi)
        g[-1] = -2 * psi[-1]
        return g

    psi_t, grads = AR.final_state_and_gradient(_static if spin == "one" else None, spin, controls,
                                               0.0, t1, 50e-9, 1e-6, psi0, loss_grad,
                                               rotating=rotating)
    h = 1e-3 * TAU * 8e4
    for c in range(2):
        for j in (0, 3, 6):
            vals = []
            for sign in (1, -1):
                a = [list(x) for x in controls]
                pert = amp[c].copy()
                pert[j] += sign * h
                a[c][2] = pert
                field = AR.controlled_field(_static if spin == "one" else None,
                                            [(x[0], O.InterpolationSampler(x[1], x[2], x[3], x[4]),
                                              x[5]) for x in a])
                _, u, _, _ = O.get_time_evolution(field, spin, 0.0, t1, 50e-9, 1e-6,
                                                  use_rotating_frame=rotating)
                st = O.get_state(psi0, u)
                vals.append(loss(u[-1] @ st[-1]))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - grads[c][j]) <= 1e-6 * max(abs(fd), 1e-12) + 1e-14, (c, j, fd, grads[c][j])


@pytest.mark.parametrize("spin", ["one", "half"])
def test_reference_gradient_through_the_rotating_frequency(spin):
    """A z-axis control moves each coarse step's rotating frequency (the z field
    at the step's midpoint).  At a 250 ns step the CF4 result depends on the
    frame at the 1e-6 level, so central differences separate the full gradient
    (agreement ~2e-9 of max |grad|) from one that holds the frequency fixed
    (1.5e-6 to 2.8e-6 here)."""
    rng = np.random.default_rng(5)
    t1, m, dt = 6e-6, 7, 250e-9
    tk = np.linspace(0, t1, m)
    amp = rng.normal(0, TAU * 8e4, (2, m))
    controls = [(0, tk, amp[0], 3, True, 1.0), (2, tk, amp[1], 1, True, 0.8)]
    psi0 = O.EIGENSTATES[spin]["plus_x"]
    static = _static if spin == "one" else None

    def loss(psi):
        return 1 - abs(psi[-1]) ** 2

    def loss_grad(psi):
        g = np.zeros_like(psi)
        g[-1] = -2 * psi[-1]
        return g

    _, grads = AR.final_state_and_gradient(static, spin, controls, 0.0, t1, dt, 1e-6, psi0,
                                           loss_grad)
    scale = max(np.abs(g).max() for g in grads)
    h = 2e-4 * TAU * 8e4
    for c in range(2):
        for j in range(m):
            vals = []
            for sign in (1, -1):
                a = [list(x) for x in controls]
                pert = amp[c].copy()
                pert[j] += sign * h
                a[c][2] = pert
                field = AR.controlled_field(static, [(x[0], O.InterpolationSampler(*x[1:5]), x[5])
                                                     for x in a])
                _, u, _, _ = O.get_time_evolution(field, spin, 0.0, t1, dt, 1e-6)
                vals.append(loss(u[-1] @ O.get_state(psi0, u)[-1]))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - grads[c][j]) < 1e-7 * scale, (c, j, fd, grads[c][j])
