"""Host-side logic and the C ABI, without a GPU."""
from __future__ import annotations

import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

import fields as F
from oracle import spinsim_oracle as O
from oracle.c_oracle import PRELUDE

import spinsim_amd as ss
from spinsim_amd import _native
from spinsim_amd.simulator import FieldSource, derive_steps, resolve_exponentiation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- the C ABI
def _declared_functions():
    with open(os.path.join(ROOT, "include", "spinsim_hip.h")) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ss_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    names = _declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in _native.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.ss_abi_version() == _native.ABI_VERSION == 9
    assert lib.ss_source_hash() != 0


def test_library_rejects_bad_arguments():
    lib = _native.lib()
    assert lib.ss_time_evolution(None, None, 0.0, 1e-7, 1e-6, 10, None, None, None, 0, None) != 0
    assert b"null module" in lib.ss_last_error()
    assert lib.ss_state_scan(None, None, None, None, 5, 4, None, 0, None) != 0
    assert b"dim" in lib.ss_last_error()
    assert lib.ss_spin(None, None, -1, 2, None) != 0
    assert lib.ss_state_scan_workspace(0, 2) == 0
    assert lib.ss_state_scan_workspace(100, 3) > 0
    assert lib.ss_batch_state_scan(None, None, None, None, None, 3, 10, 2, 0, 0, None, 0, None) != 0
    assert b"null buffer" in lib.ss_last_error()
    assert lib.ss_time_evolution(None, None, 0.0, 1e-7, 1e-6, 10, None, None, None, 65, None) != 0
    assert lib.ss_fma_probe(None, 0, 1, None) != 0
    # zero-sized work is a no-op success, even with null buffers
    assert lib.ss_spin(None, None, 0, 2, None) == 0
    assert lib.ss_batch_state_scan(None, None, None, None, None, 0, 0, 2, 0, 0, None, 0, None) == 0
    # the first-column operator layout is for spin half only
    assert lib.ss_batch_state_scan(None, None, None, None, None, 0, 0, 3, 0, 1, None, 0, None) != 0
    assert b"u_layout" in lib.ss_last_error()
    assert lib.ss_batch_state_scan(None, None, None, None, None, 0, 0, 3, 0, 2, None, 0, None) != 0
    assert lib.ss_batch_state_scan(None, None, None, None, None, 0, 0, 2, 0, 3, None, 0, None) != 0
    assert lib.ss_batch_state_scan(None, None, None, None, None, 0, 0, 2, 0, 2, None, 0, None) == 0
    assert lib.ss_batch_state_scan(None, None, None, None, None, 0, 0, 2, 0, 1, None, 0, None) == 0


def test_batch_scan_chunk_sizes():
    """Chunks are whole 1024-row tiles; the workspace holds one record per chunk."""
    lib = _native.lib()
    ws = lib.ss_batch_scan_workspace
    assert ws(5000, 2, 1000) == 0 and ws(5000, 2, -1024) == 0 and ws(5000, 4, 1024) == 0
    assert ws(0, 2, 0) == 0
    # 5000 rows in 1024-row chunks: 5 records; in one 8192-row chunk: 1
    assert ws(5000, 2, 1024) > ws(5000, 2, 8192) > 0
    assert 0 < ws(10 ** 8, 3, 0) <= lib.ss_state_scan_workspace(10 ** 8, 3)
    # a chunk size that is not a whole number of tiles is rejected before launch
    p = ctypes.c_void_p(256)
    for bad in (1000, -1024, 1536):
        assert lib.ss_batch_state_scan(p, p, p, None, p, 1, 5000, 2, bad, 0, p, 1 << 20, None) != 0
        assert b"chunk_rows" in lib.ss_last_error()
        assert lib.ss_batch_costate_scan(p, p, p, p, 1, 5000, 3, bad, p, 1 << 20, None) != 0
        assert b"chunk_rows" in lib.ss_last_error()


@pytest.mark.parametrize("dim,expo", [(2, 0), (2, 1), (3, 1), (3, 2)])
@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("rotating", [True, False])
def test_hiprtc_compiles_every_variant(dim, expo, method, rotating):
    src = (F.TILTED_EIGHT.src if expo == 2 else
           F.RABI_SWEEP_ONE.src if dim == 3 else F.RABI_SWEEP_HALF.src)
    code, _log = _native.compile_code(src, dim, method, expo, rotating, 24, False)
    assert code[:4] == b"\x7fELF"


def test_hiprtc_reports_compile_errors():
    with pytest.raises(_native.NativeError) as err:
        _native.compile_code("__device__ void get_field(double t, const double* s, double* f)"
                             "{ f[0] = undefined_symbol; }", 2, 0, 0, True, 24, False)
    assert "undefined_symbol" in str(err.value)
    with pytest.raises(_native.NativeError):
        _native.compile_code(F.LARMOR.src, 3, 0, 0, True, 24, False)  # analytic is spin-half only
    with pytest.raises(_native.NativeError, match="spin-one only"):
        _native.compile_code(F.LARMOR.src, 2, 0, 2, True, 24, False)
    with pytest.raises(_native.NativeError, match="float32-bulk"):
        _native.compile_code(F.TILTED_EIGHT.src, 3, 0, 2, True, 24, False, 1)


def test_simulator_compile_error_message(capsys):
    with pytest.raises(Exception):
        ss.Simulator("not C at all", ss.SpinQuantumNumber.HALF)
    assert "spinsim error!!!" in capsys.readouterr().out


# ---------------------------------------------------------------- API surface
def test_enums_match_reference():
    assert ss.SpinQuantumNumber.HALF.value == 0.5 and ss.SpinQuantumNumber.HALF.dimension == 2
    assert ss.SpinQuantumNumber.ONE.value == 1 and ss.SpinQuantumNumber.ONE.dimension == 3
    for spin, label in ((ss.SpinQuantumNumber.HALF, "half"), (ss.SpinQuantumNumber.ONE, "one")):
        assert spin.label == label
        for name, vec in O.EIGENSTATES[label].items():
            assert np.array_equal(getattr(spin, name), vec)
    assert ss.IntegrationMethod.EULER is ss.IntegrationMethod.MIDPOINT_SAMPLE
    assert ss.IntegrationMethod.HEUN is ss.IntegrationMethod.HALF_STEP
    assert ss.IntegrationMethod.MAGNUS_CF4.value == "magnus_cf4"
    assert [m.index for m in ss.ExponentiationMethod] == [0, 1, 2]
    assert ss.Device.ROC is ss.Device.ROCM
    assert {d.name for d in ss.Device} >= {"PYTHON", "CPU_SINGLE", "CPU", "CUDA", "ROCM"}


def test_cuda_device_raises():
    with pytest.raises(NotImplementedError):
        ss.Simulator(F.LARMOR.src, ss.SpinQuantumNumber.HALF, device=ss.Device.CUDA)


def test_exponentiation_defaults_and_fallback(capsys):
    E = ss.ExponentiationMethod
    assert resolve_exponentiation(ss.SpinQuantumNumber.HALF, None) == E.ANALYTIC
    assert resolve_exponentiation(ss.SpinQuantumNumber.ONE, None) == E.LIE_TROTTER
    assert resolve_exponentiation(ss.SpinQuantumNumber.ONE, E.ANALYTIC) == E.LIE_TROTTER
    assert "Switching to a Lie Trotter method" in capsys.readouterr().out
    assert resolve_exponentiation(ss.SpinQuantumNumber.ONE, E.LIE_TROTTER_8) == E.LIE_TROTTER_8
    with pytest.raises(ValueError, match="spin-one"):
        resolve_exponentiation(ss.SpinQuantumNumber.HALF, E.LIE_TROTTER_8)


def test_step_arithmetic_matches_oracle(capsys):
    rng = np.random.default_rng(4)
    for _ in range(300):
        dt_out = 10 ** rng.uniform(-7, -3)
        dt = dt_out / rng.choice([1, 2, 3, 7, 10, 50, 2.5, 0.7, 3.3])
        t0 = rng.uniform(-1, 1)
        t1 = t0 + rng.uniform(0, 1e3) * dt_out
        dt_a, n_a, ends = derive_steps(t0, t1, dt, dt_out)
        dt_o, n_o, _, warned = O.timing(t0, t1, dt, dt_out)
        assert dt_a == dt_o and n_a == n_o
        assert ("not an integer multiple" in capsys.readouterr().out) == warned


# ---------------------------------------------------------------- callbacks
def _compile_c_field(src):
    """Compile a C field callback with gcc into a callable (t, sweep) -> field[8]."""
    import hashlib
    import tempfile

    wrapper = ("\nvoid ss_eval(double t, const double* s, double* f)"
               "{ for (int k = 0; k < 8; ++k) f[k] = 0; get_field(t, s, f); }\n")
    stub = "double ss_interp_sample(const ss_ctx *c, int k, double t) { return 0; }\n"
    text = PRELUDE + src + wrapper + stub
    d = tempfile.mkdtemp(prefix="ssfield")
    name = hashlib.sha1(text.encode()).hexdigest()[:12]
    c_path, so_path = os.path.join(d, name + ".c"), os.path.join(d, name + ".so")
    with open(c_path, "w") as fh:
        fh.write(text)
    subprocess.run(["gcc", "-std=c11", "-O1", "-shared", "-fPIC", "-o", so_path, c_path, "-lm"],
                   check=True)
    lib = ctypes.CDLL(so_path)
    lib.ss_eval.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]

    def call(t, sweep):
        s = np.ascontiguousarray(sweep, np.float64)
        f = np.zeros(8)
        lib.ss_eval(t, s.ctypes.data, f.ctypes.data)
        return f
    return call


def _py_eval(fn, t, sweep):
    f = np.zeros(8)
    fn(t, np.asarray(sweep, float), f)
    return f


GLOBAL_SCALE = 2.5


def _helper(x, y):
    if x > y:
        return x - y
    return math.hypot(x, y)


def field_kitchen_sink(time_sample, sweep_parameters, field_sample):
    """Exercises most translatable constructs."""
    a = sweep_parameters[0] * time_sample
    b = 0.0
    for k in range(3):
        b += math.sin(a * (k + 1)) / (k + 1)
    if time_sample < sweep_parameters[1]:
        field_sample[0] = GLOBAL_SCALE * math.cos(a) ** 2
    elif time_sample < 2 * sweep_parameters[1]:
        field_sample[0] = -abs(math.tanh(a)) + min(a, 0.5, b)
    else:
        field_sample[0] = max(b, -1.0) if a > 1 else np.exp(-a)
    field_sample[1] = _helper(a, b) + (a % 0.7) - (a // 0.3) + math.atan2(b, 1.0 + a)
    field_sample[2] = math.tau * 1e3 + math.pi * float(int(a)) + np.sqrt(1 + a * a)
    field_sample[3] = field_sample[0] * 0.5 if not (a > 2 and b < 0) else -1e-3


def test_transpiled_field_evaluates_like_python():
    src, samplers, ctx = ss.transpile(field_kitchen_sink)
    assert not samplers and not ctx
    call = _compile_c_field(src)
    rng = np.random.default_rng(9)
    sweep = [3.7, 0.4]
    for t in rng.uniform(0, 1.5, 200):
        want = _py_eval(field_kitchen_sink, t, sweep)
        got = call(t, sweep)
        assert np.allclose(got, want, rtol=1e-14, atol=1e-14), (t, got, want)


def test_transpiled_field_compiles_for_gfx950():
    src, _, ctx = ss.transpile(field_kitchen_sink)
    code, _ = _native.compile_code(src, 3, 0, 1, True, 24, ctx)
    assert code[:4] == b"\x7fELF"


def test_transpiler_rejects_unsupported_constructs():
    def uses_list(t, s, f):
        f[0] = [1, 2][0]

    def uses_while(t, s, f):
        while t > 0:
            t = t - 1

    def assigns_sweep(t, s, f):
        s[0] = 1.0

    for fn in (uses_list, uses_while, assigns_sweep):
        with pytest.raises(ss.TranspileError):
            ss.transpile(fn)


def test_sampler_reference_makes_context_signature():
    noise = ss.generate_interpolation_sampler([0.0, 1.0, 2.0], [0.0, 1.0, 0.5])
    other = ss.generate_interpolation_sampler([0.0, 1.0, 2.0], [1.0, 1.0, 0.5], order=1)

    def fn(t, s, f):
        f[2] = noise(t) + 2 * other(t) + noise(t + 1)

    src, samplers, ctx = ss.transpile(fn)
    assert ctx and samplers == [noise, other]
    assert "SS_SAMPLE(0," in src and "SS_SAMPLE(1," in src and "SS_CONTEXT" in src
    code, _ = _native.compile_code(src, 2, 0, 0, True, 24, True)
    assert code[:4] == b"\x7fELF"


def test_field_source_named_samplers():
    noise = ss.generate_interpolation_sampler([0.0, 1.0], [0.0, 1.0])
    fs = FieldSource("__device__ void get_field(double t, const double* s, double* f, SS_CONTEXT)"
                     "{ f[2] = noise(t); }", samplers={"noise": noise})
    assert fs.has_context and fs.samplers == [noise]
    assert fs.source.startswith("#define noise(t) SS_SAMPLE(0, (t))")
    code, _ = _native.compile_code(fs.source, 2, 0, 0, True, 24, True)
    assert code[:4] == b"\x7fELF"


def test_sampler_validation():
    with pytest.raises(NotImplementedError):
        ss.generate_interpolation_sampler([0, 1], [0, 1], order=2)
    with pytest.raises(ValueError):
        ss.generate_interpolation_sampler([0, 1, 2], [0, 1])
    s = ss.generate_interpolation_sampler([0, 1, 2], np.ones((4, 3)), per_simulation=True)
    assert s.rows == 4 and s.axes == 1
    m = ss.generate_interpolation_sampler([0, 1, 2], np.ones((4, 3)))
    assert m.rows == 1 and m.axes == 4
    with pytest.raises(ValueError):
        ss.generate_interpolation_sampler([0, 1, 2], np.ones((2, 4, 3)))
    b = ss.generate_interpolation_sampler([0, 1, 2], np.ones((2, 4, 3)), per_simulation=True)
    assert b.rows == 2 and b.axes == 4
    s.update(np.zeros((4, 3)))
    assert s.version == 1
    with pytest.raises(ValueError):
        s.update(np.zeros(3))


def test_simulator_constructs_without_gpu():
    sim = ss.Simulator(F.RABI_ONE.src, ss.SpinQuantumNumber.ONE)
    assert sim.exponentiation_method == ss.ExponentiationMethod.LIE_TROTTER
    assert sim.dimension == 3
    sim = ss.Simulator(field_kitchen_sink, ss.SpinQuantumNumber.HALF,
                       integration_method=ss.IntegrationMethod.HEUN, use_rotating_frame=False)
    assert sim.integration_method.index == 2


def test_autotuner_candidates():
    from spinsim_amd.autotune import lane_candidates, scan_chunk_candidates

    assert lane_candidates(10) == [1, 2, 4, 5, 8, 10]
    assert lane_candidates(100) == [1, 2, 4, 5, 8, 10, 16, 20, 25, 32, 50, 64]
    assert lane_candidates(1) == [1]
    c = scan_chunk_candidates(105_801_688)
    assert all(x % 1024 == 0 for x in c) and len(c) == 5
    assert -(-105_801_688 // max(c)) <= 512 and -(-105_801_688 // min(c)) >= 7900
    assert scan_chunk_candidates(1) == [1024]



def _plan(n_coarse, n_fine, waves):
    lib = _native.lib()
    n = ctypes.c_int()
    lanes = (ctypes.c_int * 8)()
    rows, items = (ctypes.c_int64 * 8)(), (ctypes.c_int64 * 8)()
    _native.check(lib.ss_time_evolution_plan(n_coarse, n_fine, waves, ctypes.byref(n), lanes,
                                             rows, items), "plan")
    return [(lanes[p], rows[p], items[p]) for p in range(n.value)]


def _lane_steps(n_fine, lanes):
    """Fine steps each rank of a group integrates, as the kernel derives them
    (contiguous runs, the first n_fine % lanes ranks one longer)."""
    base, rem = divmod(n_fine, lanes)
    out = []
    for rank in range(lanes):
        j0 = rank * base + min(rank, rem)
        out.append(list(range(j0, j0 + base + (1 if rank < rem else 0))))
    return out


@pytest.mark.parametrize("n_coarse,n_fine,waves", [
    (100000, 10, 3072), (100000, 10, 1024), (100000, 10, 2048), (200000, 5, 1024),
    (100000, 100, 1024), (7, 3, 4096), (1, 1, 1024), (12345, 37, 777), (100000, 0, 1024),
    (50000, 32, 3072), (10 ** 7, 1, 4096)])
def test_phase_plan_covers_every_row_and_step(n_coarse, n_fine, waves):
    plan = _plan(n_coarse, n_fine, waves)
    assert 1 <= len(plan) <= 8
    row = 0
    budget = n_coarse * n_fine / (64 * waves)
    cost = 0
    for p, (lanes, first, items) in enumerate(plan):
        assert first == row and 1 <= lanes <= 64 and items > 0
        last = p == len(plan) - 1
        assert last or items % waves == 0       # whole rounds before the last phase
        row = min(n_coarse, first + items * (64 // lanes)) if last else first + items * (64 // lanes)
        steps = _lane_steps(n_fine, lanes)
        assert [j for r in steps for j in r] == list(range(n_fine))
        cost += -(-items // waves) * max(len(r) for r in steps)
    assert row == n_coarse
    # every wave ends within one finest item of its ideal share (counting the
    # lanes a group of the finest size leaves idle)
    g = plan[-1][0]
    used = (64 // g) * g / 64
    assert cost <= math.ceil(budget / used) + max(len(r) for r in _lane_steps(n_fine, g)) + 1


def test_phase_plan_of_the_spin_one_workload():
    """1e5 rows of 10 steps on three waves per SIMD: one round of two lanes per
    row (5 steps each), the remaining 1696 rows at ten lanes per row."""
    plan = _plan(100000, 10, 3072)
    assert plan == [(2, 0, 3072), (10, 3072 * 32, -(-(100000 - 3072 * 32) // 6))]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """Without the built library every ROCm entry point raises; a Device.ROCM
    simulator never falls back to the CPU devices."""
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "libspinsim_hip.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(_native.NativeError, match="no CPU fallback"):
        _native.lib()
    with pytest.raises(_native.NativeError):
        ss.Simulator(F.LARMOR.src + "\n// variant not in the code cache\n",
                     ss.SpinQuantumNumber.HALF, device=ss.Device.ROCM)


def test_fp32_bulk_override_warns_and_is_recorded():
    """The float32-bulk modes replace the exponentiation method by a residual
    series: an explicit exponentiation_method or number_of_squares warns, and
    Simulator.step_exponential records what computes the step operators."""
    import warnings

    one = ss.SpinQuantumNumber.ONE
    with pytest.warns(UserWarning, match="not used"):
        s = ss.Simulator(F.RABI_ONE.src, one, device=ss.Device.CPU, precision="fp32_bulk",
                         exponentiation_method=ss.ExponentiationMethod.LIE_TROTTER)
    assert s.step_exponential == "residual series, float32"
    with pytest.warns(UserWarning, match="number_of_squares=12"):
        ss.Simulator(F.RABI_ONE.src, one, device=ss.Device.CPU, precision="fp64_series",
                     number_of_squares=12)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        s = ss.Simulator(F.RABI_ONE.src, one, device=ss.Device.CPU, precision="fp32_bulk")
        assert ss.Simulator(F.RABI_ONE.src, one, device=ss.Device.CPU).step_exponential == \
            "lie_trotter"
