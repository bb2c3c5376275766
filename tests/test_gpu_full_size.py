"""Full-size runs of the two batched workloads against the oracle.

* The 4096-simulation spin-half Rabi ensemble (durations 1-50 ms, 50 ns
  steps, 1 us outputs: 1.06e8 coarse rows, a 6.8 GB operator array, so rows
  late in the CSR range sit beyond 2^32 bytes), float64, float32-bulk and fp32,
  in one persistent launch and one batch scan; float64 and fp32 also with the
  operators kept as first columns (compact).  64 simulations -- the first,
  the last, the longest, the shortest and the rest spread over the range --
  are checked against the C oracle (float64 1e-9, float32 modes 1e-6), and
  the last simulation's operators and Bloch vectors too.
* The differentiation workload: 1024 spin-one pulses of 2 ms at 20 ns steps,
  x and y from 2000 cubic control knots each (the 2000-knot gradient
  windows), loss 1 - |psi(T)[-1]|^2.  psi(T) of 8 sampled pulses against
  the oracle and the adjoint gradient at 64 sampled coordinates against
  central differences of the float64 forward.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

pytestmark = pytest.mark.gpu
TAU = 2 * math.pi

RABI_SWEEP = """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = 2 * s[1] * cos(s[0] * t); f[1] = 0; f[2] = s[0] + s[2]; }
"""


def _ensemble(n_sims=4096):
    rng = np.random.default_rng(20261202)
    sweep = np.stack([np.full(n_sims, TAU * 7e5), TAU * rng.uniform(500, 5000, n_sims),
                      TAU * rng.normal(0, 200, n_sims)], 1)
    dur = rng.uniform(1e-3, 50e-3, n_sims)
    return sweep, dur


def _checked(dur, count=64):
    n = len(dur)
    picks = {0, n - 1, int(np.argmax(dur)), int(np.argmin(dur))}
    for s in np.linspace(0, n - 1, count).astype(int):
        if len(picks) >= count:
            break
        picks.add(int(s))
    return sorted(picks)


@pytest.mark.parametrize("precision,tol,compact", [("fp64", 1e-9, False), ("fp32_bulk", 1e-6, False),
                                                   ("fp32", 1e-6, False), ("fp64", 1e-9, True),
                                                   ("fp32", 1e-6, True)],
                         ids=["fp64", "fp32_bulk", "fp32", "fp64_compact", "fp32_compact"])
def test_ensemble_4096_full_size(precision, tol, compact):
    import torch

    import spinsim_amd as ss
    from spinsim_amd.batch import su2_from_columns

    print(f"case: 4096-simulation ensemble, {precision}, compact={compact}", flush=True)
    sweep, dur = _ensemble()
    psi0 = ss.SpinQuantumNumber.HALF.plus_z
    sim = ss.Simulator(RABI_SWEEP, ss.SpinQuantumNumber.HALF, precision=precision)
    plan = ss.BatchPlan(sim, 0.0, dur, 50e-9, 1e-6, psi0, sweep, with_spin=True, compact=compact)
    row_bytes = 32 if compact else 64
    assert plan.total > 1e8 and plan.u.numel() * 16 == plan.total * row_bytes
    if not compact:
        assert plan.u.numel() * 16 > 2 ** 32
    plan.run()
    torch.cuda.synchronize()
    co = COracle(RABI_SWEEP, 2)
    worst = 0.0
    for s in _checked(dur):
        dt, n, nf, _ = O.timing(0.0, dur[s], 50e-9, 1e-6)
        t_ref, u_ref = co.time_evolution(sweep[s], 0.0, dt, 1e-6, nf, n)
        st_ref = co.get_state(psi0, u_ref)
        a, b = int(plan.offsets_host[s]), int(plan.offsets_host[s + 1])
        assert b - a == n
        st = plan.state[a:b].cpu().numpy()
        err = float(np.abs(st - st_ref).max())
        worst = max(worst, err)
        assert err <= tol, f"simulation {s}: state error {err:.3e}"
        if s == len(dur) - 1:  # the last rows of the arrays (beyond 4 GiB of a full U)
            assert a * row_bytes > 2 ** 32 or compact
            assert np.array_equal(plan.time[a:b].cpu().numpy(), t_ref)
            u = (su2_from_columns(plan.u[a:b]) if compact else plan.u[a:b]).cpu().numpy()
            u_err = float(np.abs(u - u_ref).max())
            assert u_err <= tol, f"last simulation: operator error {u_err:.3e}"
            spin = plan.spin_out[a:b].cpu().numpy()
            assert np.abs(spin - co.get_spin(st_ref)).max() <= tol
    print(f"  {plan.total} rows, {plan.fine_steps} steps; worst sampled state error {worst:.3e}")


def _z_bias(time_sample, sweep_parameters, field_sample):
    field_sample[2] = TAU * 1e5


CONTROLLED_ONE = """
__device__ void get_field(double t, const double* s, double* f, SS_CONTEXT)
{ f[0] = ctl_x(t); f[1] = ctl_y(t); f[2] = SS_TAU * 1e5; f[3] = 0; }
"""


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_adjoint_1024_full_size(precision):
    import torch

    import spinsim_amd as ss

    print(f"case: 1024-pulse adjoint, {precision}", flush=True)
    rng = np.random.default_rng(20261202)
    n_sims, t_end, m, sigma = 1024, 2e-3, 2000, TAU * 5e3
    tk = np.linspace(0.0, t_end, m)
    amp = rng.normal(0, sigma, (n_sims, 2, m))
    psi0 = ss.SpinQuantumNumber.ONE.plus_z
    controls = [ss.Control("x", tk), ss.Control("y", tk)]
    cs = ss.ControlledSimulator(ss.SpinQuantumNumber.ONE, controls, _z_bias, precision=precision)
    a = torch.tensor(amp, device="cuda", requires_grad=True)
    psi = cs.final_state(a, 0.0, t_end, 20e-9, 1e-6, psi0)
    (1 - psi[:, -1].abs() ** 2).sum().backward()
    grad = a.grad.cpu().numpy()
    psi = psi.detach().cpu().numpy()
    assert np.isfinite(grad).all()

    co = COracle(CONTROLLED_ONE, 3, sampler_names=("ctl_x", "ctl_y"))
    co.set_samplers([(tk, amp[:, 0], 3, True), (tk, amp[:, 1], 3, True)])
    psi_tol = 1e-9 if precision == "fp64" else 1e-6
    for s in sorted({0, n_sims - 1, *rng.choice(n_sims, 6, False).tolist()}):
        _, u = co.time_evolution([0.0], 0.0, 20e-9, 1e-6, 50, 2000, sim=int(s))
        st = co.get_state(psi0, u)
        err = float(np.abs(u[-1] @ st[-1] - psi[s]).max())
        assert err <= psi_tol, f"pulse {s}: psi(T) error {err:.3e}"

    # central differences of the float64 forward, all perturbations in one batch;
    # knots 0 and m - 1 and the ones next to them included
    ref = cs if precision == "fp64" else ss.ControlledSimulator(ss.SpinQuantumNumber.ONE,
                                                                controls, _z_bias)
    h = 1e-4 * sigma
    coords, rows = [], []
    sims = [0, n_sims - 1, *rng.choice(np.arange(1, n_sims - 1), 6, False).tolist()]
    knots = [0, 1, m - 2, m - 1]
    for s in sims:
        for q in range(8):
            k = q % 2
            j = knots[q // 2] if s in (0, n_sims - 1) else int(rng.integers(m))
            coords.append((s, k, j))
            for sign in (1, -1):
                r = amp[s].copy()
                r[k, j] += sign * h
                rows.append(r)
    with torch.no_grad():
        psi_p = ref.final_state(torch.tensor(np.stack(rows), device="cuda"), 0.0, t_end, 20e-9,
                                1e-6, psi0).cpu().numpy()
    loss = 1 - np.abs(psi_p[:, -1]) ** 2
    fd = (loss[0::2] - loss[1::2]) / (2 * h)
    ad = np.array([grad[s, k, j] for s, k, j in coords])
    rel = float(np.abs(fd - ad).max() / np.abs(fd).max())
    print(f"  {len(coords)} coordinates: max |fd| {np.abs(fd).max():.3e}, relative error {rel:.3e}")
    assert rel <= (1e-6 if precision == "fp64" else 1e-5)
