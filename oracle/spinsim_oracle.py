"""FP64 CPU oracle for spinsim's integration path.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
``spinsim_amd`` never does, and never falls back to it.

It restates, in vectorised NumPy float64, the algorithm of the reference
``spinsim/__init__.py`` (one file; all line numbers below refer to it).  Each
coarse (output) step is one NumPy lane; the fine-step loop is sequential, as in
the reference's per-thread loop.  Arithmetic is kept in the reference's order
(time accumulation ``t += dt``, the ``(t + off) - t_c`` then ``+ t_c`` sample
time, residual "m1" squarings, the left-multiplication order), so that the
only differences from the reference are libm ulps and NumPy's complex
division.

Each function is pinned by checks independent of its own formulas, in
``tests/test_oracle_pinning.py``: scipy ``expm`` (exponentials), an adaptive
ODE solution and convergence orders (sampling, CF4 and frame), closed forms
(Larmor, the Gaussian pi-pulse of gaussian_example.py), eigenstate spins,
polynomial reproduction and scipy's CubicHermiteSpline (sampler), known step
counts.  Parity unpinned: the bit-level float behaviour of numba/libdevice,
which no test can reach and which sits far below the 1e-9 tolerance.
"""
from __future__ import annotations

import math
import warnings

import numpy as np

SQRT2 = math.sqrt(2)  # spinsim/__init__.py:12
SQRT3 = math.sqrt(3)  # spinsim/__init__.py:13

HALF, ONE = "half", "one"
MAGNUS_CF4, EULER, HEUN = "magnus_cf4", "euler", "heun"
ANALYTIC, LIE_TROTTER, LIE_TROTTER_8 = "analytic", "lie_trotter", "lie_trotter_8"

# Eigenstates, spinsim/__init__.py:36-56
EIGENSTATES = {
    HALF: {
        "plus_x": np.array([1, 1], np.complex128) / math.sqrt(2),
        "minus_x": np.array([-1, 1], np.complex128) / math.sqrt(2),
        "plus_y": np.array([1, 1j], np.complex128) / math.sqrt(2),
        "minus_y": np.array([1, -1j], np.complex128) / math.sqrt(2),
        "plus_z": np.array([1, 0], np.complex128),
        "minus_z": np.array([0, 1], np.complex128),
    },
    ONE: {
        "plus_x": np.array([1, math.sqrt(2), 1], np.complex128) / 2,
        "zero_x": np.array([-1, 0, 1], np.complex128) / math.sqrt(2),
        "minus_x": np.array([1, -math.sqrt(2), 1], np.complex128) / 2,
        "plus_y": np.array([-1, -1j * math.sqrt(2), 1], np.complex128) / 2,
        "zero_y": np.array([1, 0, 1], np.complex128) / math.sqrt(2),
        "minus_y": np.array([1, -1j * math.sqrt(2), 1], np.complex128) / 2,
        "plus_z": np.array([1, 0, 0], np.complex128),
        "zero_z": np.array([0, 1, 0], np.complex128),
        "minus_z": np.array([0, 0, 1], np.complex128),
    },
}


# --------------------------------------------------------------------------
# Host arithmetic of evaluate(): spinsim/__init__.py:1025-1034 and :832
# --------------------------------------------------------------------------
def timing(time_start, time_end, time_step_integration, time_step_output):
    """Return (dt, N, n_fine, warned) exactly as evaluate() derives them.

    dt is re-derived as ``dt_out / round(max(dt_out/dt, 1))`` (:1026), the
    warning fires on ``|dt_out/dt_old - round(dt_out/dt_old)| > 1e-6``
    (:1027), ``N = int((t1 - t0)/dt_out)`` truncates (:1034), and each coarse
    step takes ``floor(dt_out/dt + 0.5)`` fine steps (:832).
    """
    dt_old = float(time_step_integration)
    dt_out = float(time_step_output)
    dt = dt_out / round(max(dt_out / dt_old, 1))
    warned = math.fabs(dt_out / dt_old - round(dt_out / dt_old)) > 1e-6
    ends = np.asarray([time_start, time_end], np.float64)
    n_coarse = int((ends[1] - ends[0]) / dt_out)
    n_fine = math.floor(dt_out / dt + 0.5)
    return dt, n_coarse, n_fine, warned


# --------------------------------------------------------------------------
# Small-matrix helpers (entry lists of 1-D complex lanes), :1552-1615, :1679-1761
# --------------------------------------------------------------------------
def expm1i(b):
    """exp(ib) - 1 without cancellation, :1556-1557."""
    return -2 * (np.sin(b / 2) ** 2) + 1j * np.sin(b)


def cos_exp_m1(c, e):
    """cos(c) exp(ie) - 1, :1560-1561."""
    return (expm1i(c + e) + expm1i(-c + e)) / 2


def identity(dim, n):
    out = [[np.zeros(n, np.complex128) for _ in range(dim)] for _ in range(dim)]
    for k in range(dim):
        out[k][k] = np.ones(n, np.complex128)
    return out


def matrix_multiply(left, right):
    """(L R)[r,c] = sum_k L[r,k] R[k,c], summed in k order, :1583-1587 / :1707-1716."""
    dim = len(left)
    out = [[None] * dim for _ in range(dim)]
    for r in range(dim):
        for c in range(dim):
            acc = left[r][0] * right[0][c]
            for k in range(1, dim):
                acc = acc + left[r][k] * right[k][c]
            out[r][c] = acc
    return out


def matrix_square_m1(op):
    """Residual square (2 + a) a for A = 1 + a, :1589-1593 / :1718-1727."""
    dim = len(op)
    out = [[None] * dim for _ in range(dim)]
    for r in range(dim):
        for c in range(dim):
            acc = None
            for k in range(dim):
                term = ((2 + op[r][r]) * op[r][c]) if k == r else (op[r][k] * op[k][c])
                acc = term if acc is None else acc + term
            out[r][c] = acc
    return out


def _hypercubes(number_of_squares):
    """number_of_hypercubes and trotter_precision, :1547-1550."""
    hyper = math.ceil(number_of_squares / 2)
    if hyper < 0:
        hyper = 0
    return hyper, 4 ** hyper


# --------------------------------------------------------------------------
# Matrix exponentials of g (list of lie_dimension float lanes)
# --------------------------------------------------------------------------
def matrix_exponential_analytic(g):
    """Spin-half exp(-i g.J), J = sigma/2, :1617-1641."""
    x, y, z = (np.asarray(v, np.float64).copy() for v in g[:3])
    r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    pos = r > 0
    rs = np.where(pos, r, 1.0)
    x = np.where(pos, x / rs, 0.0)
    y = np.where(pos, y / rs, 0.0)
    z = np.where(pos, z / rs, 0.0)
    c = cos_exp_m1(r / 2, 0)
    s = np.sin(r / 2)
    one = np.ones_like(r)
    e00 = np.where(pos, c - 1j * z * s + 1, one)
    e10 = np.where(pos, (y - 1j * x) * s, 0)
    e01 = np.where(pos, -(y + 1j * x) * s, 0)
    e11 = np.where(pos, c + 1j * z * s + 1, one)
    return [[e00.astype(np.complex128), e01.astype(np.complex128)],
            [e10.astype(np.complex128), e11.astype(np.complex128)]]


def matrix_exponential_lie_trotter_half(g, number_of_squares=24):
    """Spin-half Strang/Lie-Trotter exponential with residual squaring, :1643-1673."""
    hyper, precision = _hypercubes(number_of_squares)
    gx, gy, gz = (np.asarray(v, np.float64) for v in g[:3])
    a = np.sqrt(gx * gx + gy * gy)
    pos = a > 0
    ep = np.where(pos, (gx + 1j * gy) / np.where(pos, a, 1.0), 1 + 0j)
    a = a / precision
    sa = -1j * np.sin(a / 2)
    z = gz / (2 * precision)
    res = [[cos_exp_m1(a / 2, -z), sa / ep],
           [sa * ep, cos_exp_m1(a / 2, z)]]
    for _ in range(hyper):
        res = matrix_square_m1(matrix_square_m1(res))
    res[0][0] = res[0][0] + 1
    res[1][1] = res[1][1] + 1
    return res


def matrix_exponential_lie_trotter_one(g, number_of_squares=24):
    """Spin-one exponential of -i(x Jx + y Jy + z Jz + q Q), :1767-1807."""
    hyper, precision = _hypercubes(number_of_squares)
    gx, gy, gz, gq = (np.asarray(v, np.float64) for v in g[:4])
    a = np.sqrt(gx * gx + gy * gy)
    p = np.where(a > 0, np.arctan2(gy, gx), 0.0)
    a = a / precision
    big_sa = np.sin(a / 2)
    sa = -1j * np.sin(a) / SQRT2
    z = gz / (2 * precision)
    q = gq / (6 * precision)
    save = cos_exp_m1(a / 2, -z - q)
    r00 = save * (save + 2)
    r10 = sa * np.exp(1j * (q + p - z))
    r20 = -(big_sa ** 2) * np.exp(2 * 1j * (p - q))
    r01 = sa * np.exp(1j * (q - p - z))
    r11 = cos_exp_m1(a, 4 * q)
    r21 = sa * np.exp(1j * (q + p + z))
    r02 = -(big_sa ** 2) * np.exp(2 * 1j * (q - p))
    r12 = sa * np.exp(1j * (q - p + z))
    save = cos_exp_m1(a / 2, z - q)
    r22 = save * (save + 2)
    res = [[r00, r01, r02], [r10, r11, r12], [r20, r21, r22]]
    res = [[np.asarray(v, np.complex128) for v in row] for row in res]
    for _ in range(hyper):
        res = matrix_square_m1(matrix_square_m1(res))
    for k in range(3):
        res[k][k] = res[k][k] + 1
    return res


def _expm1i(b):
    """exp(ib) - 1 without cancellation."""
    return -2 * np.sin(b / 2) ** 2 + 1j * np.sin(b)


def _rotation_one_m1(angle, ep, sign2):
    """exp(-i angle (cos p Jx + sin p Jy)) - I for spin one, ep = exp(ip);
    with sign2 = -1 the same conjugated by diag(1, 1, -1) (the V1/V2 pair)."""
    sh = np.sin(angle / 2)
    sa = -1j * np.sin(angle) / SQRT2
    epc = np.conj(ep)
    d = -(sh ** 2) + 0j
    return [[d, sa * epc, -(sh ** 2) * epc ** 2 * sign2],
            [sa * ep, -2 * sh ** 2 + 0j, sa * epc * sign2],
            [-(sh ** 2) * ep ** 2 * sign2, sa * ep * sign2, d]]


def _residual_product(f, r):
    """(I + f)(I + r) - I = f + r + f r."""
    fr = matrix_multiply(f, r)
    return [[f[i][k] + r[i][k] + fr[i][k] for k in range(3)] for i in range(3)]


def matrix_exponential_lie_trotter_eight(g, number_of_squares=24):
    """Spin-one exponential of -i(x Jx + y Jy + z Jz + q Q + u1 U1 + u2 U2 +
    v1 V1 + v2 V2) with the generators of the reference's documentation,
    :1422-1484.  The reference's own matrix_exponential_lie_trotter_8
    (:1810-1925) is not a usable specification (it reuses a stale variable and
    fills entries that must be zero), so this is the documented formula done
    as a symmetric Lie-Trotter product at scale 4**-hypercubes,
        T = D(1/2) V(1/2) U(1/2) J(1) U(1/2) V(1/2) D(1/2),
    each factor an exact exponential kept as its residual T - I, followed by
    the same residual squarings as the other Lie-Trotter exponentials."""
    hyper, precision = _hypercubes(number_of_squares)
    gx, gy, gz, gq, u1, u2, v1, v2 = (np.asarray(v, np.float64) for v in g[:8])
    s = 1.0 / precision

    def phase(a, b):
        n = np.sqrt(a * a + b * b)
        safe = np.where(n > 0, n, 1.0)
        return n, np.where(n > 0, (a + 1j * b) / safe, 1.0 + 0j)

    a_j, ep_j = phase(gx, gy)
    a_u, ep_u = phase(u1, u2)
    a_v, ep_v = phase(v1, v2)
    zero = np.zeros_like(gx) + 0j
    f_j = _rotation_one_m1(a_j * s, ep_j, 1.0)
    f_v = _rotation_one_m1(a_v * s / 2, ep_v, -1.0)
    th = a_u * s / 2
    cu = -2 * np.sin(th / 2) ** 2 + 0j
    su = -1j * np.sin(th)
    f_u = [[cu, zero, su * np.conj(ep_u)], [zero, zero, zero], [su * ep_u, zero, cu]]
    h = s / 2
    f_d = [[_expm1i(-h * (gz + gq / 3)), zero, zero],
           [zero, _expm1i(h * 2 * gq / 3), zero],
           [zero, zero, _expm1i(-h * (-gz + gq / 3))]]
    res = f_d
    for f in (f_v, f_u, f_j, f_u, f_v, f_d):
        res = _residual_product(f, res)
    res = [[np.asarray(v, np.complex128) for v in row] for row in res]
    for _ in range(hyper):
        res = matrix_square_m1(matrix_square_m1(res))
    for k in range(3):
        res[k][k] = res[k][k] + 1
    return res


# --------------------------------------------------------------------------
# Rotating frame, :582-618 (frame entry) and :838-852 (frame exit)
# --------------------------------------------------------------------------
def transform_frame(field, rotating_wave, winding, dim, rotating):
    if not rotating:
        return  # transform_frame_lab, :615-617
    if dim == 2:
        # transform_frame_spin_half_rotating, :607-612
        x = (field[0] + 1j * field[1]) / (winding ** 2)
        field[0] = x.real
        field[1] = x.imag
        field[2] = field[2] - 2 * rotating_wave
    else:
        # transform_frame_spin_one_rotating, :599-604
        x = (field[0] + 1j * field[1]) / winding
        field[0] = x.real
        field[1] = x.imag
        field[2] = field[2] - rotating_wave
        if len(field) > 4:
            # transform_frame_spin_one_rotating_8, :586-596: U rotates twice as fast
            x = (field[4] + 1j * field[5]) * np.conj(winding) * np.conj(winding)
            field[4] = x.real
            field[5] = x.imag
            x = (field[6] + 1j * field[7]) * np.conj(winding)
            field[6] = x.real
            field[7] = x.imag


# --------------------------------------------------------------------------
# Time-evolution kernel, :796-896, vectorised over coarse index
# --------------------------------------------------------------------------
def resolve_methods(spin, exponentiation=None, integration=MAGNUS_CF4):
    """Defaults and fallbacks of compile_time_evolver, :529-559."""
    if exponentiation is None:
        exponentiation = LIE_TROTTER if spin == ONE else ANALYTIC
    if exponentiation == ANALYTIC and spin != HALF:
        exponentiation = LIE_TROTTER
    if exponentiation == LIE_TROTTER_8 and spin != ONE:
        raise ValueError("LIE_TROTTER_8 is only defined for spin one")  # :1429-1432
    dim = 2 if spin == HALF else 3
    lie = 3 if spin == HALF else (8 if exponentiation == LIE_TROTTER_8 else 4)
    return dim, lie, exponentiation, integration


def _exponential(g, dim, exponentiation, number_of_squares):
    if exponentiation == ANALYTIC:
        return matrix_exponential_analytic(g)
    if dim == 2:
        return matrix_exponential_lie_trotter_half(g, number_of_squares)
    if exponentiation == LIE_TROTTER_8:
        return matrix_exponential_lie_trotter_eight(g, number_of_squares)
    return matrix_exponential_lie_trotter_one(g, number_of_squares)


def _sample(get_field, t, sweep, lie, n):
    field = np.zeros((lie, n), np.float64)
    get_field(t, sweep, field)
    return field


def get_time_evolution(get_field, spin, time_start, time_end, time_step_integration,
                       time_step_output, sweep_parameters=(0.0,), exponentiation=None,
                       integration=MAGNUS_CF4, use_rotating_frame=True, number_of_squares=24,
                       coarse_index=None, record=None):
    """Return (time (N,), U (N, d, d) complex128, dt, n_fine).

    ``get_field(t, sweep, field)`` is evaluated with ``t`` a float64 array of
    lanes and ``field`` a (lie_dimension, lanes) array, so a field written with
    NumPy ufuncs follows the reference's per-sample callback (:620) lane by
    lane.  ``coarse_index`` restricts the run to a subset of coarse steps (the
    result rows then follow that subset).  ``record``, if a list, receives one
    dict per fine step (CF4 only): the lab sample times and their offsets from
    the coarse time, windings, the frame-transformed field samples, and the
    generators g of its two exponentials, lanes along the last axis.
    """
    dim, lie, exponentiation, integration = resolve_methods(spin, exponentiation, integration)
    dt, n_coarse, n_fine, warned = timing(time_start, time_end, time_step_integration,
                                          time_step_output)
    if warned:
        warnings.warn("time_step_output not an integer multiple of time_step_integration")
    dt_out = float(time_step_output)
    sweep = np.asarray(sweep_parameters, np.float64)
    idx = np.arange(n_coarse) if coarse_index is None else np.asarray(coarse_index)
    n = idx.size
    t0 = np.float64(time_start)

    time_coarse = t0 + dt_out * idx.astype(np.float64)            # :818
    time_fine = time_coarse.copy()                                 # :819
    u = identity(dim, n)                                           # :822
    if use_rotating_frame:                                         # :823-829
        rot_field = np.zeros((lie, n), np.float64)
        get_field(time_coarse + dt_out / 2, sweep, rot_field)
        rotating_wave = rot_field[2].copy()
    else:
        rotating_wave = np.zeros(n, np.float64)
    if dim == 2:
        rotating_wave = rotating_wave / 2

    w0 = (1.5 + SQRT3) / 6                                         # :659
    w1 = (1.5 - SQRT3) / 6                                         # :660

    def append(g):
        e = _exponential(g, dim, exponentiation, number_of_squares)
        return matrix_multiply(e, u)                               # :579-580

    for _ in range(n_fine):                                        # :832
        if integration == MAGNUS_CF4:                              # :623-686
            t_rel = (time_fine + 0.5 * dt * (1 - 1 / SQRT3)) - time_coarse
            wa = np.exp(1j * rotating_wave * t_rel)
            fa = _sample(get_field, t_rel + time_coarse, sweep, lie, n)
            t_rel = (time_fine + 0.5 * dt * (1 + 1 / SQRT3)) - time_coarse
            wb = np.exp(1j * rotating_wave * t_rel)
            fb = _sample(get_field, t_rel + time_coarse, sweep, lie, n)
            transform_frame(fa, rotating_wave, wa, dim, use_rotating_frame)
            transform_frame(fb, rotating_wave, wb, dim, use_rotating_frame)
            g = [dt * (w0 * fa[k] + w1 * fb[k]) for k in range(lie)]
            u = append(g)
            g2 = [dt * (w1 * fa[k] + w0 * fb[k]) for k in range(lie)]
            u = append(g2)
            if record is not None:
                ra = (time_fine + 0.5 * dt * (1 - 1 / SQRT3)) - time_coarse
                record.append({"ta": ra + time_coarse, "tb": t_rel + time_coarse, "wa": wa,
                               "wb": wb, "ra": ra, "rb": t_rel, "fa": fa.copy(),
                               "fb": fb.copy(), "tc": time_coarse,
                               "g1": np.array(g), "g2": np.array(g2), "dt": dt,
                               "rot": rotating_wave})
        elif integration == HEUN:                                  # :692-752
            t_rel = time_fine - time_coarse
            wa = np.exp(1j * rotating_wave * t_rel)
            fa = _sample(get_field, t_rel + time_coarse, sweep, lie, n)
            t_rel = time_fine + dt - time_coarse
            wb = np.exp(1j * rotating_wave * t_rel)
            fb = _sample(get_field, t_rel + time_coarse, sweep, lie, n)
            transform_frame(fa, rotating_wave, wa, dim, use_rotating_frame)
            transform_frame(fb, rotating_wave, wb, dim, use_rotating_frame)
            u = append([dt * fa[k] / 2 for k in range(lie)])
            u = append([dt * fb[k] / 2 for k in range(lie)])
        elif integration == EULER:                                 # :758-791
            t_rel = time_fine + 0.5 * dt - time_coarse
            wa = np.exp(1j * rotating_wave * t_rel)
            fa = _sample(get_field, t_rel + time_coarse, sweep, lie, n)
            transform_frame(fa, rotating_wave, wa, dim, use_rotating_frame)
            u = append([dt * fa[k] for k in range(lie)])
        else:
            raise ValueError(integration)
        time_fine = time_fine + dt                                 # :836

    if use_rotating_frame:                                         # :839-852
        w = np.exp(1j * rotating_wave * dt_out)
        last = dim - 1
        for c in range(dim):
            u[0][c] = u[0][c] / w
            u[last][c] = u[last][c] * w

    out = np.empty((n, dim, dim), np.complex128)
    for r in range(dim):
        for c in range(dim):
            out[:, r, c] = u[r][c]
    return time_coarse, out, dt, n_fine


# --------------------------------------------------------------------------
# State product and Bloch vector
# --------------------------------------------------------------------------
def get_state(state_init, time_evolution):
    """state[0] = psi0; state[i] = U[i-1] state[i-1]; U[N-1] unused, :1087-1111."""
    u = np.asarray(time_evolution, np.complex128)
    n, dim = u.shape[0], u.shape[1]
    state = np.empty((n, dim), np.complex128)
    if n == 0:
        return state
    prev = [complex(v) for v in np.asarray(state_init, np.complex128)]
    state[0] = prev
    ul = u.tolist()
    for i in range(1, n):
        m = ul[i - 1]
        cur = []
        for x in range(dim):
            acc = 0j
            for z in range(dim):
                acc += m[x][z] * prev[z]
            cur.append(acc)
        state[i] = cur
        prev = cur
    return state


def get_spin(state):
    """Bloch vector per row, :936-944."""
    s = np.asarray(state, np.complex128)
    spin = np.empty((s.shape[0], 3), np.float64)
    if s.shape[1] == 2:
        spin[:, 0] = (s[:, 0] * np.conj(s[:, 1])).real
        spin[:, 1] = (1j * s[:, 0] * np.conj(s[:, 1])).real
        spin[:, 2] = 0.5 * (s[:, 0].real ** 2 + s[:, 0].imag ** 2
                            - s[:, 1].real ** 2 - s[:, 1].imag ** 2)
    else:
        spin[:, 0] = (2 * np.conj(s[:, 1]) * (s[:, 0] + s[:, 2]) / SQRT2).real
        spin[:, 1] = (2j * np.conj(s[:, 1]) * (s[:, 0] - s[:, 2]) / SQRT2).real
        spin[:, 2] = (s[:, 0].real ** 2 + s[:, 0].imag ** 2
                      - s[:, 2].real ** 2 - s[:, 2].imag ** 2)
    return spin


def evaluate(get_field, spin, time_start, time_end, time_step_integration, time_step_output,
             state_init, sweep_parameters=(0.0,), **kw):
    """The whole evaluate() path, :993-1085: (time, U, state, spin)."""
    time, u, _, _ = get_time_evolution(get_field, spin, time_start, time_end,
                                       time_step_integration, time_step_output,
                                       sweep_parameters, **kw)
    state = get_state(np.asarray(state_init, np.complex128), u)
    return time, u, state, get_spin(state)


# --------------------------------------------------------------------------
# Interpolation sampler, :1941-2142
# --------------------------------------------------------------------------
class InterpolationSampler:
    """Vectorised restatement of generate_interpolation_sampler's sampler."""

    def __init__(self, time, amplitude, order=3, zero_boundary=True):
        self.time = np.asarray(time, np.float64)
        self.amplitude = np.asarray(amplitude, np.float64)
        if self.amplitude.ndim == 2:
            raise NotImplementedError("2-D amplitudes")          # :2062-2063
        if order not in (0, 1, 3):
            raise NotImplementedError(f"Interpolation order {order}")
        self.order = order
        self.zero_boundary = zero_boundary

    def find_time_index(self, t):
        """Uniform-grid guess, then bisection; (-1, 0) out of range, :2026-2061."""
        time = self.time
        size = time.size
        t = np.atleast_1d(np.asarray(t, np.float64))
        index = np.full(t.shape, -1, np.int64)
        param = np.zeros(t.shape, np.float64)
        for e in range(t.size):
            ts = t.flat[e]
            if not (time[0] <= ts <= time[-1]):
                continue
            i = int((ts - time[0]) / (time[1] - time[0]))
            low, high = False, True
            if i + 1 < size:
                low = time[i] <= ts
                high = ts <= time[i + 1]
            else:
                i = size
            failed = False
            if not (low and high):
                imax, imin, budget = size - 1, 0, size
                while not (low and high):
                    if not low:
                        imax = i
                    else:
                        imin = i
                    i = int(float(imax + imin) / 2)
                    low, high = False, True
                    if i + 1 < size:
                        low = time[i] <= ts
                        high = ts <= time[i + 1]
                    budget -= 1
                    if budget == 0:
                        failed = True
                        break
            if failed:
                continue
            index.flat[e] = i
            param.flat[e] = (ts - time[i]) / (time[i + 1] - time[i])
        return index, param

    def __call__(self, t):
        time, amp = self.time, self.amplitude
        size = time.size
        scalar = np.ndim(t) == 0
        t = np.atleast_1d(np.asarray(t, np.float64))
        out = np.zeros(t.shape, np.float64)
        idx, tau = self.find_time_index(t)
        for e in range(t.size):
            ts = t.flat[e]
            if not self.zero_boundary:
                if ts <= time[0]:
                    out.flat[e] = amp[0]
                    continue
                if ts >= time[-1]:
                    out.flat[e] = amp[-1]
                    continue
            i = idx.flat[e]
            p = tau.flat[e]
            if self.order == 0:
                if i >= 0:
                    out.flat[e] = amp[i]
            elif self.order == 1:
                if i >= 0 and i + 1 < size:
                    out.flat[e] = amp[i] * (1 - p) + amp[i + 1] * p
            else:
                if i >= 0 and i + 1 < size:
                    out.flat[e] = self._cubic(i, p)
        return out[0] if scalar else out

    def _cubic(self, i, p):
        """Cubic Hermite with averaged finite-difference tangents, :2108-2141."""
        time, amp = self.time, self.amplitude
        size = time.size
        if time[i] != time[i + 1]:
            mid = (amp[i + 1] - amp[i]) / (time[i + 1] - time[i])
        else:
            mid = 0
        if i == 0:
            left = mid
        elif time[i - 1] != time[i]:
            left = 0.5 * ((amp[i] - amp[i - 1]) / (time[i] - time[i - 1]) + mid)
        else:
            left = 0
        left *= (time[i + 1] - time[i])
        if i + 2 == size:
            right = mid
        elif time[i + 1] != time[i + 2]:
            right = 0.5 * ((amp[i + 2] - amp[i + 1]) / (time[i + 2] - time[i + 1]) + mid)
        else:
            right = 0
        right *= (time[i + 1] - time[i])
        return (amp[i] * (2 * p ** 3 - 3 * p ** 2 + 1)
                + amp[i + 1] * (-2 * p ** 3 + 3 * p ** 2)
                + left * (p ** 3 - 2 * p ** 2 + p)
                + right * (p ** 3 - p ** 2))
