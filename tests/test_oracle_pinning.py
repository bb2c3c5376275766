"""Pin the oracle against what the mathematics fixes, independently of its own
formulas: scipy's expm, an adaptive ODE solution, closed forms, known
eigenstate spins and convergence orders.  Both restatements (NumPy and C)
are checked, and against each other."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.linalg import expm

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import COracle

SQ2 = math.sqrt(2)
# Spin operators as documented by the reference (spinsim/__init__.py:1262-1395).
JH = [np.array([[0, 1], [1, 0]]) / 2, np.array([[0, -1j], [1j, 0]]) / 2,
      np.array([[1, 0], [0, -1]]) / 2]
J1 = [np.array([[0, 1, 0], [1, 0, 1], [0, 1, 0]]) / SQ2,
      np.array([[0, -1j, 0], [1j, 0, -1j], [0, 1j, 0]]) / SQ2,
      np.diag([1.0, 0, -1]), np.diag([1.0, -2, 1]) / 3]


# The further su(3) directions of the eight-generator exponential (:1458-1484).
J8 = J1 + [np.array([[0, 0, 1], [0, 0, 0], [1, 0, 0]]),
           np.array([[0, 0, -1j], [0, 0, 0], [1j, 0, 0]]),
           np.array([[0, 1, 0], [1, 0, -1], [0, -1, 0]]) / SQ2,
           np.array([[0, -1j, 0], [1j, 0, 1j], [0, -1j, 0]]) / SQ2]


def hamiltonian(f, spin):
    ops = JH if spin == "half" else (J8 if len(f) > 4 else J1)
    return sum(f[k] * ops[k] for k in range(min(len(ops), len(f))))


def to_mat(entries):
    return np.array([[entries[r][c] for c in range(len(entries))] for r in range(len(entries))])


# ---------------------------------------------------------------- exponentials
G_SAMPLES = [np.array(v, float) for v in (
    [0.3, -0.2, 0.5, 0.1], [1e-9, 2e-9, -3e-9, 4e-9], [0.0, 0.0, 0.7, 0.2], [0.0, 0.0, 0.0, 0.0],
    [2.5, 1.0, -3.0, 1.5], [0.0, 1.2, 0.0, -0.4], [-0.8, 0.0, 0.0, 0.0])]


@pytest.mark.parametrize("g", G_SAMPLES)
def test_exponentials_match_expm(g):
    cols = [np.array([v]) for v in g]
    ref_h = expm(-1j * hamiltonian(g[:3], "half"))
    e = to_mat(O.matrix_exponential_analytic(cols))[:, :, 0]
    assert np.abs(e - ref_h).max() < 1e-14
    e = to_mat(O.matrix_exponential_lie_trotter_half(cols))[:, :, 0]
    assert np.abs(e - ref_h).max() < 1e-12
    ref_1 = expm(-1j * hamiltonian(g, "one"))
    e = to_mat(O.matrix_exponential_lie_trotter_one(cols))[:, :, 0]
    assert np.abs(e - ref_1).max() < 1e-12


G8_SAMPLES = [np.array(v, float) for v in (
    [0.3, -0.2, 0.5, 0.1, 0.4, -0.3, 0.2, 0.6], [1e-9, 2e-9, -3e-9, 4e-9, 5e-9, -1e-9, 2e-9, 3e-9],
    [0.0, 0.0, 0.0, 0.0, 0.7, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -0.9],
    [2.5, 1.0, -3.0, 1.5, -2.0, 0.5, 1.7, -1.1], [0.0] * 8)]


@pytest.mark.parametrize("g", G8_SAMPLES)
def test_eight_generator_exponential_matches_expm(g):
    """The generators as the reference documents them, checked independently:
    hermitian, traceless, orthogonal under the trace form; the exponential
    equals scipy's expm of -i sum g_k lambda_k."""
    for k, a in enumerate(J8):
        assert np.allclose(a, a.conj().T) and abs(np.trace(a)) < 1e-15
        for m, b in enumerate(J8):
            if m != k:
                assert abs(np.trace(a @ b)) < 1e-15
    cols = [np.array([v]) for v in g]
    e = to_mat(O.matrix_exponential_lie_trotter_eight(cols))[:, :, 0]
    assert np.abs(e - expm(-1j * hamiltonian(g, "one"))).max() < 1e-12
    if not g[4:].any():
        e4 = to_mat(O.matrix_exponential_lie_trotter_one(cols[:4]))[:, :, 0]
        assert np.abs(e - e4).max() < 1e-13


def test_eight_generator_squares_converge():
    """Fewer squares leave the symmetric splitting error, O(P^-2) at scale
    1/P = 4^-h: each hypercube divides it by about 16."""
    g = np.array([0.9, -0.4, 1.3, 0.7, 0.8, -0.6, 0.5, 1.1])
    cols = [np.array([v]) for v in g]
    ref = expm(-1j * hamiltonian(g, "one"))
    errs = [np.abs(to_mat(O.matrix_exponential_lie_trotter_eight(cols, s))[:, :, 0] - ref).max()
            for s in (0, 2, 4, 6)]
    assert errs[0] > 1e-3
    for a, b in zip(errs, errs[1:]):
        assert a / 20 < b < a / 10


def test_lie_trotter_squares_converge():
    """Fewer squares give the Strang error O(4^-h^2...); more approach expm."""
    g = np.array([0.9, -0.4, 1.3, 0.7])
    cols = [np.array([v]) for v in g]
    ref = expm(-1j * hamiltonian(g, "one"))
    errs = [np.abs(to_mat(O.matrix_exponential_lie_trotter_one(cols, s))[:, :, 0] - ref).max()
            for s in (0, 2, 4, 6)]
    assert errs[0] > 1e-3
    for a, b in zip(errs, errs[1:]):
        assert b < a / 10  # each hypercube divides the splitting error by 16


def _square_derived_row(m):
    """(1 + m)^2 - 1 with the last row taken from special unitarity
    (spin half a10 = -conj a01, a11 = conj a00; spin one a2 = conj((-a02, -a12,
    a00 + a11) + a0 x a1)), the form the HIP squaring uses."""
    d = len(m)
    b = [[m[r][c] + (2 if r == c else 0) for c in range(d)] for r in range(d)]
    t = [[sum(m[r][k] * b[k][c] for k in range(d)) for c in range(d)] for r in range(d - 1)]
    if d == 2:
        return [t[0], [-np.conj(t[0][1]), np.conj(t[0][0])]]
    a0, a1 = t
    return [a0, a1, [np.conj(-a0[2] + a0[1] * a1[2] - a0[2] * a1[1]),
                     np.conj(-a1[2] + a0[2] * a1[0] - a0[0] * a1[2]),
                     np.conj(a0[0] + a1[1] + a0[0] * a1[1] - a0[1] * a1[0])]]


@pytest.mark.parametrize("fn,lie,q_zero,squares,bound", [
    ("matrix_exponential_lie_trotter_half", 3, False, 1, 1e-14),
    ("matrix_exponential_lie_trotter_half", 3, False, 7, 1e-14),
    ("matrix_exponential_lie_trotter_eight", 8, False, 1, 1e-14),
    ("matrix_exponential_lie_trotter_eight", 8, False, 7, 1e-14),
    ("matrix_exponential_lie_trotter_one", 4, True, 1, 1e-14),
    ("matrix_exponential_lie_trotter_one", 4, False, 24, 1e-14),
])
def test_lie_trotter_starts_special_unitary(monkeypatch, fn, lie, q_zero, squares, bound):
    """The Lie-Trotter starts are special unitary, so squaring with the last row
    derived from the others equals the full square to rounding: spin half and
    the eight generators at any number of squares, spin one when q = 0, and
    spin one at 24 squares for |g| <= 1."""
    rng = np.random.default_rng(11)
    g = [rng.normal(0, 0.5, 200) for _ in range(lie)]
    if q_zero:
        g[3] = np.zeros(200)
    full = np.array(getattr(O, fn)(g, squares))
    monkeypatch.setattr(O, "matrix_square_m1", _square_derived_row)
    derived = np.array(getattr(O, fn)(g, squares))
    assert np.abs(full - derived).max() < bound


def test_spin_one_lie_trotter_start_quirk():
    """The reference's spin-one start (:1780-1792) puts exp(2i(q - p)) on entry
    (0, 2), where the symmetric product exp(-iD/2) exp(-iR) exp(-iD/2) has
    exp(-2i(q + p)); with q != 0 the start is then not special unitary, and at
    few squares the full square and the derived-row square differ visibly."""
    g = [np.array([0.3]), np.array([-0.2]), np.array([0.4]), np.array([0.5])]
    start = np.array(O.matrix_exponential_lie_trotter_one(g, 0))[:, :, 0]
    a = math.hypot(0.3, -0.2)
    p, q = math.atan2(-0.2, 0.3), 0.5 / 6
    assert abs(start[0, 2] - (-math.sin(a / 2) ** 2 * np.exp(2j * (q - p)))) < 1e-15
    d = np.diag(np.exp(-1j * np.array([0.2 + q, -2 * q, -0.2 + q])))
    rot = expm(-1j * (0.3 * hamiltonian([1, 0, 0, 0], "one") - 0.2 * hamiltonian([0, 1, 0, 0], "one")))
    strang = d @ rot @ d
    off = np.abs(start - strang)
    assert off[0, 2] > 1e-3 and np.delete(off.ravel(), 2).max() < 1e-14
    full = np.array(O.matrix_exponential_lie_trotter_one([v * 16 for v in g], 2))
    saved = O.matrix_square_m1
    try:
        O.matrix_square_m1 = _square_derived_row
        derived = np.array(O.matrix_exponential_lie_trotter_one([v * 16 for v in g], 2))
    finally:
        O.matrix_square_m1 = saved
    assert np.abs(full - derived).max() > 1e-6


@pytest.mark.parametrize("spin,expo", [("half", "analytic"), ("half", "lie_trotter"),
                                       ("one", "lie_trotter")])
def test_c_oracle_single_step_is_expm(spin, expo):
    """One midpoint step in the lab frame of a constant field is exp(-i dt f.J)."""
    f = [0.31e6, -0.17e6, 0.52e6, 0.11e6]
    dim = 2 if spin == "half" else 3
    src = ("__device__ void get_field(double t, const double* s, double* f)"
           f"{{ f[0] = {f[0]!r}; f[1] = {f[1]!r}; f[2] = {f[2]!r}; f[3] = {f[3]!r}; }}")
    co = COracle(src, dim, "euler", expo, False)
    _, u = co.time_evolution([0.0], 0.0, 1e-6, 1e-6, 1, 3)
    ref = expm(-1j * 1e-6 * hamiltonian(f if dim == 3 else f[:3], spin))
    assert np.abs(u - ref).max() < 1e-12


# ---------------------------------------------------------------- closed forms
def test_larmor_closed_form_numpy_and_c():
    """Larmor precession: psi(t) = (e^{-i pi 1000 t}, e^{i pi 1000 t})/sqrt2."""
    psi0 = O.EIGENSTATES["half"]["plus_x"]
    t, u, state, spin = O.evaluate(F.LARMOR.fn, "half", 0, 2e-3, 100e-9, 500e-9, psi0)
    exact = np.stack([np.exp(-1j * math.pi * 1000 * t), np.exp(1j * math.pi * 1000 * t)], 1) / SQ2
    assert t.shape == (4000,)
    assert np.abs(state - exact).max() < 1e-12
    w = 2 * math.pi * 1000
    assert np.abs(spin - np.stack([0.5 * np.cos(w * t), 0.5 * np.sin(w * t), 0 * t], 1)).max() < 1e-12
    co = COracle(F.LARMOR.src, 2)
    dt, n, nf, _ = O.timing(0, 100e-3, 100e-9, 500e-9)
    assert (n, nf) == (200000, 5)
    tc, uc = co.time_evolution([0.0], 0.0, dt, 500e-9, nf, n)
    sc = co.get_state(psi0, uc)
    exact = np.stack([np.exp(-1j * math.pi * 1000 * tc), np.exp(1j * math.pi * 1000 * tc)], 1) / SQ2
    assert np.abs(sc - exact).max() < 1e-9
    # the lab frame gives the same answer
    tl, ul, sl, _ = O.evaluate(F.LARMOR.fn, "half", 0, 2e-3, 100e-9, 500e-9, psi0,
                               use_rotating_frame=False)
    assert np.abs(sl - state).max() < 1e-11


def _cumulative_gaussian(t):
    return 0.5 * (1 + math.erf(t / math.sqrt(2.0))) - 0.5 * (1 + math.erf(-5.0 / math.sqrt(2.0)))


def test_gaussian_pulse_known_answer():
    """gaussian_example.py:22-41; documented error < 1e-6 at 40 steps."""
    time = np.arange(-5.0, 5.1, 2.0)
    analytic = np.asarray([[math.cos(0.5 * math.pi * _cumulative_gaussian(t)),
                            -1j * math.sin(0.5 * math.pi * _cumulative_gaussian(t))] for t in time])
    errs = []
    for step in (0.5, 0.25, 0.125):
        _, _, state, _ = O.evaluate(F.GAUSSIAN.fn, "half", -5.0, 7.0, step, 2.0,
                                    np.array([1, 0], np.complex128))
        errs.append(np.sqrt(np.sum(np.abs(state - analytic) ** 2)) / 5)
    assert 1e-8 < errs[1] < 1e-6       # 40 steps
    assert 12 < errs[0] / errs[1] < 20  # fourth order
    assert 12 < errs[1] / errs[2] < 20


# ---------------------------------------------------------------- vs an ODE solver
def _ode_reference(field, spin, t0, t1, dt_out, psi0, lie=4):
    n = int(round((t1 - t0) / dt_out))

    def rhs(t, y):
        f = np.zeros((lie, 1))
        field.fn(np.array([t]), np.asarray(field.sweep), f)
        return -1j * hamiltonian(f[:, 0] if spin == "one" else f[:3, 0], spin) @ y

    ts = t0 + dt_out * np.arange(n)
    sol = solve_ivp(rhs, (t0, ts[-1]), psi0.astype(complex), method="DOP853", t_eval=ts,
                    rtol=1e-13, atol=1e-13)
    return ts, sol.y.T


@pytest.mark.parametrize("field,expo", [(F.TILTED_HALF, "analytic"),
                                        (F.TILTED_HALF, "lie_trotter"),
                                        (F.TILTED_ONE, None), (F.RABI_SWEEP_ONE, None),
                                        (F.TILTED_EIGHT, "lie_trotter_8")])
@pytest.mark.parametrize("rotating", [True, False])
def test_cf4_matches_ode_solution(field, expo, rotating):
    """Also pins the eight-generator frame rule (U1/U2 turn at twice the
    frame rate, V1/V2 at the frame rate, :586-596) through the rotating run."""
    psi0 = O.EIGENSTATES[field.spin]["plus_x"]
    ts, ref = _ode_reference(field, field.spin, 0.0, 6e-5, 1e-6, psi0, 8 if expo == "lie_trotter_8" else 4)
    _, _, state, _ = O.evaluate(field.fn, field.spin, 0.0, 6e-5, 10e-9, 1e-6, psi0, field.sweep,
                                exponentiation=expo, use_rotating_frame=rotating)
    assert np.abs(state[: ts.size] - ref).max() < 2e-8


@pytest.mark.parametrize("method,order", [("magnus_cf4", 4), ("euler", 2), ("heun", 2)])
def test_convergence_order(method, order):
    field = F.TILTED_ONE
    psi0 = O.EIGENSTATES["one"]["plus_y"]
    ts, ref = _ode_reference(field, "one", 0.0, 2e-5, 2e-6, psi0)
    errs = []
    for dt in (2e-7, 1e-7):
        _, _, state, _ = O.evaluate(field.fn, "one", 0.0, 2e-5, dt, 2e-6, psi0,
                                    integration=method, use_rotating_frame=False)
        errs.append(np.abs(state[: ts.size] - ref).max())
    ratio = errs[0] / errs[1]
    assert 2 ** order * 0.7 < ratio < 2 ** order * 1.4


# ---------------------------------------------------------------- state and spin
def test_get_state_offsets_and_order():
    # commuting diagonal steps: state[i] = exp(i sum_{k<i} theta_k) psi0
    rng = np.random.default_rng(0)
    th = rng.uniform(-1, 1, (50, 2))
    u = np.zeros((50, 2, 2), complex)
    u[:, 0, 0] = np.exp(1j * th[:, 0])
    u[:, 1, 1] = np.exp(1j * th[:, 1])
    psi0 = np.array([0.6, 0.8j])
    st = O.get_state(psi0, u)
    cum = np.vstack([[0, 0], np.cumsum(th, 0)[:-1]])
    assert np.abs(st - psi0 * np.exp(1j * cum)).max() < 1e-13
    # non-commuting: state[2] = H X psi0, not X H psi0
    x = np.array([[0, 1], [1, 0]], complex)
    h = np.array([[1, 1], [1, -1]], complex) / SQ2
    st = O.get_state(np.array([1, 0], complex), np.stack([x, h, x]))
    assert np.allclose(st[2], [1 / SQ2, -1 / SQ2])
    # C restatement agrees
    co = COracle(F.LARMOR.src, 2)
    assert np.abs(co.get_state(np.array([1, 0], complex), np.stack([x, h, x])) - st).max() == 0


@pytest.mark.parametrize("spin,expected", [
    ("half", {"plus_x": (0.5, 0, 0), "minus_x": (-0.5, 0, 0), "plus_y": (0, 0.5, 0),
              "minus_y": (0, -0.5, 0), "plus_z": (0, 0, 0.5), "minus_z": (0, 0, -0.5)}),
    ("one", {"plus_x": (1, 0, 0), "minus_x": (-1, 0, 0), "plus_y": (0, 1, 0),
             "plus_z": (0, 0, 1), "minus_z": (0, 0, -1), "zero_x": (0, 0, 0),
             "zero_y": (0, 0, 0), "zero_z": (0, 0, 0)})])
def test_spin_of_eigenstates(spin, expected):
    """<J> of the reference's eigenstates; also <psi|J|psi> computed directly."""
    ops = JH if spin == "half" else J1[:3]
    names = list(expected)
    states = np.stack([O.EIGENSTATES[spin][k] for k in names])
    s = O.get_spin(states)
    for row, name in zip(s, names):
        assert np.allclose(row, expected[name], atol=1e-15), name
    rng = np.random.default_rng(5)
    d = states.shape[1]
    psi = rng.normal(size=(20, d)) + 1j * rng.normal(size=(20, d))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    direct = np.stack([[np.vdot(p, op @ p).real for op in ops] for p in psi])
    assert np.abs(O.get_spin(psi) - direct).max() < 1e-14
    co = COracle(F.LARMOR.src if spin == "half" else F.RABI_ONE.src, d)
    assert np.abs(co.get_spin(psi) - direct).max() < 1e-14


# ---------------------------------------------------------------- step arithmetic
def test_timing_known_counts():
    assert O.timing(0, 100e-3, 100e-9, 500e-9)[1:] == (200000, 5, False)
    assert O.timing(0, 100e-3, 100e-9, 1e-6)[1:] == (100000, 10, False)
    assert O.timing(0, 100e-3, 10e-9, 1e-6)[1:] == (100000, 100, False)
    assert O.timing(0, 2e-3, 20e-9, 1e-6)[1:] == (2000, 50, False)
    dt, n, nf, warned = O.timing(0, 3e-5, 0.3e-6, 1e-6)
    assert warned and nf == 3 and n == 30 and dt == 1e-6 / 3
    dt, n, nf, warned = O.timing(0, 1e-5, 5e-6, 1e-6)  # dt > dt_out: one step per output
    assert nf == 1 and dt == 1e-6
    assert O.timing(0, 0.5e-6, 1e-7, 1e-6)[1] == 0


# ---------------------------------------------------------------- interpolation
def test_sampler_reproduces_polynomials():
    t = np.linspace(0, 2, 21)
    lin = 3 * t - 1
    quad = t ** 2 - 0.5 * t
    ts = np.linspace(0.1, 1.9, 301)  # interior intervals
    for order in (1, 3):
        s = O.InterpolationSampler(t, lin, order)
        assert np.abs(s(ts) - (3 * ts - 1)).max() < 1e-12
    s = O.InterpolationSampler(t, quad, 3)
    assert np.abs(s(ts) - (ts ** 2 - 0.5 * ts)).max() < 1e-12
    s = O.InterpolationSampler(t, quad, 0)
    off = np.linspace(0.105, 1.895, 180)  # off the knots: hold the left sample
    assert np.abs(s(off) - quad[np.floor(off / 0.1).astype(int)]).max() < 1e-15


def test_sampler_matches_scipy_hermite():
    from scipy.interpolate import CubicHermiteSpline

    rng = np.random.default_rng(2)
    t = np.sort(rng.uniform(0, 1, 15))
    a = rng.normal(size=15)
    sec = np.diff(a) / np.diff(t)
    m = np.empty(15)
    m[0], m[-1] = sec[0], sec[-1]
    m[1:-1] = 0.5 * (sec[:-1] + sec[1:])
    ref = CubicHermiteSpline(t, a, m)
    s = O.InterpolationSampler(t, a, 3)
    ts = rng.uniform(t[0], t[-1], 500)
    assert np.abs(s(ts) - ref(ts)).max() < 1e-12


def test_sampler_boundaries():
    t = np.array([0.0, 1.0, 2.0])
    a = np.array([2.0, -1.0, 4.0])
    for order in (0, 1, 3):
        z = O.InterpolationSampler(t, a, order, True)
        c = O.InterpolationSampler(t, a, order, False)
        assert z(-0.5) == 0 and z(2.5) == 0
        assert c(-0.5) == 2.0 and c(2.5) == 4.0
    assert O.InterpolationSampler(t, a, 3)(2.0) == 4.0


# ---------------------------------------------------------------- two restatements
@pytest.mark.parametrize("field", [F.TILTED_HALF, F.TILTED_ONE, F.RABI_SWEEP_HALF, F.TILTED_EIGHT])
@pytest.mark.parametrize("method", ["magnus_cf4", "euler", "heun"])
@pytest.mark.parametrize("rotating", [True, False])
def test_numpy_and_c_restatements_agree(field, method, rotating):
    dim = 2 if field.spin == "half" else 3
    expos = (["analytic", "lie_trotter"] if dim == 2 else
             ["lie_trotter_8"] if field is F.TILTED_EIGHT else [None])
    for expo in expos:
        t, u, dt, nf = O.get_time_evolution(field.fn, field.spin, 3e-6, 3e-6 + 5e-5, 1e-7, 1e-6,
                                            field.sweep, exponentiation=expo, integration=method,
                                            use_rotating_frame=rotating)
        co = COracle(field.src, dim, method, expo, rotating)
        tc, uc = co.time_evolution(field.sweep, 3e-6, dt, 1e-6, nf, t.size)
        assert np.array_equal(t, tc)
        assert np.abs(u - uc).max() < 1e-13
