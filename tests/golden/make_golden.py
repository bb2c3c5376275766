"""Write the golden fixtures of this directory from the float64 NumPy oracle.

    python tests/golden/make_golden.py

The reference holds no vectors of its own, so these are the oracle's outputs
on short, seeded versions of the reference's workloads, kept small (a few
hundred rows each).  Each fixture is an .npz of plain arrays (no pickles) with
a JSON string ``meta`` naming the field (a key of tests/fields.py), spin,
integration and exponentiation method, frame, times and initial state.  The
tests pin every fixture against what the mathematics fixes (closed forms, an
adaptive ODE solve, finite differences) and against the C oracle, then hold
the HIP path to it.

Fixtures:
  larmor_half_1ms        spin-half Larmor precession, the first 1 ms of the
                         100 ms run (2 pi 1 kHz on z, dt 100 ns, output
                         500 ns, +x, analytic, rotating frame)
  rabi_one_1ms           spin-one Rabi drive with quadratic shift, the first
                         1 ms of the 100 ms run (dt 100 ns, output 1 us, +z,
                         Lie-Trotter)
  methods_*              a strongly time-dependent field on every axis, each
                         integration method, both frames, 200 us
  eight_generator_one    the eight-generator Lie-Trotter exponential, 60 us
  batch_rabi_half_4      four swept spin-half Rabi simulations drawn like the
                         4096-simulation ensemble (amplitude U[0.5, 5] kHz,
                         detuning N(0, 200 Hz), 700 kHz carrier, dt 50 ns,
                         output 1 us), durations U[1, 50] ms scaled by 1/50
  adjoint_one_2          two spin-one state-transfer pulses drawn like the
                         1024-pulse differentiation workload (cubic x and y
                         controls N(0, 2 pi 5 kHz), z = 2 pi 100 kHz, dt 20 ns),
                         20 us with 21 control samples: final state and the
                         gradient of 1 - |psi(T)[-1]|^2
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))                 # tests/ (fields)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))  # repo root (oracle)

import fields as F  # noqa: E402
from oracle import adjoint_reference as AR  # noqa: E402
from oracle import spinsim_oracle as O  # noqa: E402

TAU = 2 * math.pi
FIELDS = {f.name: f for f in (F.LARMOR, F.RABI_ONE, F.RABI_SWEEP_HALF, F.TILTED_HALF,
                              F.TILTED_ONE, F.TILTED_EIGHT)}


def _rows(n, every):
    """Every `every`-th output row and the last one."""
    return np.unique(np.concatenate([np.arange(0, n, every), [n - 1]])).astype(np.int64)


def _meta(**kw):
    return np.array(json.dumps(kw, sort_keys=True))


def trajectory(name, field, t0, t1, dt, dt_out, psi0, every, method="magnus_cf4", expo=None,
               rotating=True, sweep=None):
    sweep = list(field.sweep if sweep is None else sweep)
    time, u, state, spin = O.evaluate(field.fn, field.spin, t0, t1, dt, dt_out, psi0, sweep,
                                      exponentiation=expo, integration=method,
                                      use_rotating_frame=rotating)
    rows = _rows(time.size, every)
    psi0 = np.asarray(psi0, np.complex128)
    return name, dict(
        meta=_meta(kind="trajectory", field=field.name, spin=field.spin, method=method,
                   expo=expo, rotating=rotating, t0=t0, t1=t1, dt=dt, dt_out=dt_out,
                   sweep=sweep, n_rows=int(time.size)),
        psi0=psi0, rows=rows, time=time[rows], u=u[rows], state=state[rows], spin=spin[rows])


def batch():
    rng = np.random.default_rng(4096)
    b = 4
    sweep = np.stack([np.full(b, TAU * 7e5), TAU * rng.uniform(500, 5000, b),
                      TAU * rng.normal(0, 200, b)], 1)
    t1 = rng.uniform(1e-3, 50e-3, b) / 50
    dt, dt_out = 50e-9, 1e-6
    psi0 = np.tile(np.array([1, 0], np.complex128), (b, 1))
    counts, rows, state, spin, u, time = [0], [], [], [], [], []
    for s in range(b):
        tt, uu, st, sp = O.evaluate(F.RABI_SWEEP_HALF.fn, "half", 0.0, t1[s], dt, dt_out, psi0[s],
                                    sweep[s])
        r = _rows(tt.size, 25)
        rows.append(r + counts[-1])
        counts.append(counts[-1] + tt.size)
        for acc, x in ((state, st), (spin, sp), (u, uu), (time, tt)):
            acc.append(x[r])
    return "batch_rabi_half_4", dict(
        meta=_meta(kind="batch", field=F.RABI_SWEEP_HALF.name, spin="half", method="magnus_cf4",
                   expo=None, rotating=True, dt=dt, dt_out=dt_out),
        sweep=sweep, t0=np.zeros(b), t1=t1, psi0=psi0, offsets=np.array(counts, np.int64),
        rows=np.concatenate(rows), time=np.concatenate(time), u=np.concatenate(u),
        state=np.concatenate(state), spin=np.concatenate(spin))


def z_bias(t, sweep, f):
    f[2] = TAU * 1e5


def _last_level_grad(psi):
    """d(1 - |psi[-1]|^2) / dRe(psi) + i d/dIm(psi)."""
    g = np.zeros_like(psi)
    g[..., -1] = -2 * psi[..., -1]
    return g


def adjoint():
    rng = np.random.default_rng(1024)
    b, m, t1, dt, dt_out = 2, 21, 20e-6, 20e-9, 1e-6
    tk = np.linspace(0.0, t1, m)
    amp = rng.normal(0, TAU * 5e3, (b, 2, m))
    psi0 = np.array([1, 0, 0], np.complex128)
    psi_t, grad = [], []
    for s in range(b):
        controls = [(0, tk, amp[s, 0], 3, True, 1.0), (1, tk, amp[s, 1], 3, True, 1.0)]
        p, g = AR.final_state_and_gradient(z_bias, "one", controls, 0.0, t1, dt, dt_out, psi0,
                                           _last_level_grad)
        psi_t.append(p)
        grad.append(np.stack(g))
    return "adjoint_one_2", dict(
        meta=_meta(kind="adjoint", spin="one", static="z = 2 pi 100 kHz", controls=["x", "y"],
                   order=3, zero_boundary=True, t0=0.0, t1=t1, dt=dt, dt_out=dt_out,
                   loss="1 - |psi(T)[-1]|^2"),
        tk=tk, amp=amp, psi0=psi0, psi_t=np.stack(psi_t), grad=np.stack(grad))


def fixtures():
    plus_x = O.EIGENSTATES["half"]["plus_x"]
    yield trajectory("larmor_half_1ms", F.LARMOR, 0.0, 1e-3, 100e-9, 500e-9, plus_x, 100,
                     expo="analytic")
    yield trajectory("rabi_one_1ms", F.RABI_ONE, 0.0, 1e-3, 100e-9, 1e-6,
                     O.EIGENSTATES["one"]["plus_z"], 50, expo="lie_trotter")
    for method in ("magnus_cf4", "euler", "heun"):
        for rotating in (True, False):
            frame = "rotating" if rotating else "lab"
            yield trajectory(f"methods_{method}_half_{frame}", F.TILTED_HALF, 0.0, 2e-4, 100e-9,
                             2e-6, O.EIGENSTATES["half"]["plus_y"], 10, method=method,
                             expo="analytic" if rotating else "lie_trotter", rotating=rotating)
            yield trajectory(f"methods_{method}_one_{frame}", F.TILTED_ONE, 0.0, 2e-4, 100e-9,
                             2e-6, O.EIGENSTATES["one"]["plus_x"], 10, method=method,
                             rotating=rotating)
    yield trajectory("eight_generator_one", F.TILTED_EIGHT, 0.0, 6e-5, 20e-9, 1e-6,
                     O.EIGENSTATES["one"]["plus_x"], 5, expo="lie_trotter_8")
    yield batch()
    yield adjoint()


def main():
    for name, arrays in fixtures():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrays)
        print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
