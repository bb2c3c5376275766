"""Float64 CPU reference gradient for the adjoint -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's checks use it.  For one simulation it follows the
oracle's CF4 forward pass (spinsim_oracle.get_time_evolution with its step
record: same sample times, rotating frame and CF4 combinations as
spinsim/__init__.py:623-686, :796-852), forms every step exponential exactly
with scipy.linalg.expm, walks the costate backward, and takes each step's
derivative with scipy.linalg.expm_frechet.  It shares no code with the HIP
adjoint (which uses a series for the Frechet derivative and walks states
backward).

The rotating frequency of coarse step i is the z field at the step's midpoint
(halved for spin half), so a z-axis control also moves the frame: every
frame-transformed sample of the step (x + iy turned by exp(-i m rot t_rel),
z shifted by -m rot, m = 2 for spin half, 1 for spin one) and the frame exit
exp(-+i rot dt_out) on the first and last components.  Per unit of the
midpoint z field that is

    sum over samples of  t_rel (dL/df_x f_y - dL/df_y f_x) - dL/df_z
      + (dt_out / m) (Im(conj(lam_0) psi_0) - Im(conj(lam_last) psi_last))

with f the frame-transformed sample and psi, lam the state and costate after
the exit; it is scattered onto the z controls at the midpoint, so the
gradient is that of the computed forward map.

Control fields: field[axis] += scale * sampler_k(t), sampler_k an oracle
InterpolationSampler; gradients with respect to the sampler amplitudes use
that the sampler is linear in them (the weights are sampler evaluations of
unit amplitude vectors).
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import expm, expm_frechet

from . import spinsim_oracle as O

SQ2 = np.sqrt(2)
JH = [np.array([[0, 1], [1, 0]]) / 2, np.array([[0, -1j], [1j, 0]]) / 2,
      np.array([[1, 0], [0, -1]]) / 2]
J1 = [np.array([[0, 1, 0], [1, 0, 1], [0, 1, 0]]) / SQ2,
      np.array([[0, -1j, 0], [1j, 0, -1j], [0, 1j, 0]]) / SQ2,
      np.diag([1.0, 0, -1]), np.diag([1.0, -2, 1]) / 3]


def _generator(g, spin):
    ops = JH if spin == O.HALF else J1
    return -1j * sum(g[k] * ops[k] for k in range(len(ops)))


def controlled_field(static_fn, controls):
    """Vectorised field: static part plus scale * sampler on each control axis.

    controls: list of (axis, sampler, scale)."""
    def fn(t, sweep, f):
        if static_fn is not None:
            static_fn(t, sweep, f)
        for axis, sampler, scale in controls:
            f[axis] = f[axis] + scale * sampler(np.atleast_1d(t))
    return fn


def final_state_and_gradient(static_fn, spin, controls, t0, t1, dt, dt_out, psi0, loss_grad,
                             sweep=(0.0,), rotating=True):
    """psi(T) and dL/d(amplitudes) for one simulation (CF4 only).

    controls: list of (axis, time, amplitude, order, zero_boundary, scale).
    loss_grad(psi_T) -> dL/dRe(psi) + i dL/dIm(psi) (torch's convention).
    Returns (psi_T, [grad per control]).
    """
    samplers = [(c[0], O.InterpolationSampler(c[1], c[2], c[3], c[4]), c[5]) for c in controls]
    field = controlled_field(static_fn, samplers)
    record = []
    _, u, dt2, nf = O.get_time_evolution(field, spin, t0, t1, dt, dt_out, sweep,
                                         use_rotating_frame=rotating, record=record)
    n = u.shape[0]
    d = 2 if spin == O.HALF else 3
    dt_out = float(dt_out)
    # exact forward of the whole run, interval by interval
    psi = np.asarray(psi0, complex)
    states = []  # (i, j, which): state before exponential
    steps = []   # (A, g, frame info)
    exits = []
    for i in range(n):
        rot = record[i * 0]["rot"][i] if record else 0.0
        for j in range(nf):
            rec = record[j]
            for which in (0, 1):
                g = (rec["g1"] if which == 0 else rec["g2"])[:, i]
                a = _generator(g, spin)
                states.append(psi)
                steps.append((i, j, which, a, g))
                psi = expm(a) @ psi
        if rotating:
            w = np.exp(1j * record[0]["rot"][i] * dt_out)
            x = np.ones(d, complex)
            x[0], x[-1] = np.conj(w), w
            psi = x * psi
            exits.append((len(steps), x))
    psi_t = psi
    lam = np.asarray(loss_grad(psi_t), complex)
    grads = [np.zeros(len(c[1])) for c in controls]
    # unit-amplitude samplers for the weights
    units = []
    for (axis, time, amp, order, zb, scale) in controls:
        units.append([O.InterpolationSampler(time, np.eye(len(time))[m], order, zb)
                      for m in range(len(time))])
    exit_at = {k: x for k, x in exits}
    exit_row = {k: r for r, (k, _) in enumerate(exits)}
    m_frame = 2.0 if spin == O.HALF else 1.0
    dz_mid = np.zeros(n)  # dL/d(z field at the midpoint of coarse step i)
    lie = 3 if spin == O.HALF else 4
    ops = JH if spin == O.HALF else J1
    w0 = (1.5 + O.SQRT3) / 6
    w1 = (1.5 - O.SQRT3) / 6
    dg = {}
    for idx in range(len(steps) - 1, -1, -1):
        if idx + 1 in exit_at:
            psi_x = states[idx + 1] if idx + 1 < len(states) else psi_t
            dz_mid[exit_row[idx + 1]] += dt_out / m_frame * (
                np.imag(np.conj(lam[0]) * psi_x[0]) - np.imag(np.conj(lam[-1]) * psi_x[-1]))
            lam = np.conj(exit_at[idx + 1]) * lam
        i, j, which, a, g = steps[idx]
        psi_b = states[idx]
        dgk = np.array([np.real(np.vdot(lam, expm_frechet(a, -1j * ops[k], compute_expm=False)
                                        @ psi_b)) for k in range(lie)])
        dg[(i, j, which)] = dgk
        lam = expm(a).conj().T @ lam
    for i in range(n):
        rot = record[0]["rot"][i]
        for j in range(nf):
            rec = record[j]
            d1, d2 = dg[(i, j, 0)], dg[(i, j, 1)]
            dfa = dt2 * (w0 * d1 + w1 * d2)
            dfb = dt2 * (w1 * d1 + w0 * d2)
            for df, tkey, wkey, s in ((dfa, "ta", "wa", "a"), (dfb, "tb", "wb", "b")):
                lab = df.copy()
                if rotating:
                    f = rec["f" + s][:, i]
                    dz_mid[i] += rec["r" + s][i] * (df[0] * f[1] - df[1] * f[0]) - df[2]
                    w = rec[wkey][i]
                    wf = w * w if spin == O.HALF else w
                    gw = (df[0] + 1j * df[1]) * wf
                    lab[0], lab[1] = gw.real, gw.imag
                t = rec[tkey][i]
                for c, (axis, time, amp, order, zb, scale) in enumerate(controls):
                    v = scale * (lab[axis] if axis < 3 else lab[lie - 1])
                    for m, us in enumerate(units[c]):
                        wgt = us(t)
                        if wgt != 0:
                            grads[c][m] += v * wgt
        if rotating:
            t_mid = record[0]["tc"][i] + dt_out / 2
            for c, (axis, time, amp, order, zb, scale) in enumerate(controls):
                if axis == 2:
                    for m, us in enumerate(units[c]):
                        grads[c][m] += scale * dz_mid[i] * us(np.atleast_1d(t_mid))[0]
    return psi_t, grads
