"""The state and costate scans through the C ABI, against the sequential
recurrences of get_state (spinsim/__init__.py:1087-1111) on random unitaries:
tile edges, many simulations starting inside one tile (more than the 256
threads of a tile), empty simulations, chunks whose incoming state lies more
than 256 chunks back, and HIP-graph replays."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-9


def _unitaries(rng, n, d):
    """Random unitaries diag(exp(i phi)) (I - 2 v v^+ / |v|^2), built in bulk."""
    v = rng.normal(size=(n, d)) + 1j * rng.normal(size=(n, d))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    h = np.eye(d)[None] - 2 * v[:, :, None] * v.conj()[:, None, :]
    return np.exp(1j * rng.uniform(0, 2 * np.pi, (n, d)))[:, :, None] * h


def _reference(u, offsets, init, reverse):
    out = np.zeros((u.shape[0], u.shape[1]), complex)
    for s in range(len(offsets) - 1):
        a, b = offsets[s], offsets[s + 1]
        if a == b:
            continue
        if not reverse:
            out[a] = init[s]
            for i in range(a + 1, b):
                out[i] = u[i - 1] @ out[i - 1]
        else:
            out[b - 1] = init[s]
            for i in range(b - 2, a - 1, -1):
                out[i] = u[i + 1].conj().T @ out[i + 1]
    return out


def _run(u, offsets, init, reverse, with_spin=False, chunk_rows=0):
    import torch

    from spinsim_amd import _native

    L = _native.lib()
    d = u.shape[1]
    total = int(offsets[-1])
    du = torch.as_tensor(u, device="cuda")
    doff = torch.as_tensor(np.asarray(offsets, np.int64), device="cuda")
    dinit = torch.as_tensor(np.ascontiguousarray(init), device="cuda")
    out = torch.full((total, d), float("nan"), dtype=torch.complex128, device="cuda")
    spin = torch.empty((total, 3), dtype=torch.float64, device="cuda") if with_spin else None
    ws = torch.empty(max(int(L.ss_batch_scan_workspace(total, d, chunk_rows)), 1),
                     dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    b = len(offsets) - 1
    if reverse:
        rc = L.ss_batch_costate_scan(dinit.data_ptr(), du.data_ptr(), out.data_ptr(),
                                     doff.data_ptr(), b, total, d, chunk_rows, ws.data_ptr(),
                                     ws.numel(), stream)
    else:
        rc = L.ss_batch_state_scan(dinit.data_ptr(), du.data_ptr(), out.data_ptr(),
                                   None if spin is None else spin.data_ptr(), doff.data_ptr(), b,
                                   total, d, chunk_rows, 0, ws.data_ptr(), ws.numel(), stream)
    _native.check(rc, "scan")
    return out.cpu().numpy(), (None if spin is None else spin.cpu().numpy())


CASES = {
    # name: simulation lengths
    "one_long": [70001],
    "tile_edges": [255, 256, 257, 1, 1023, 1024, 1025, 2047, 2049],
    "many_tiny": None,      # 3000 simulations of 0-1 rows: ~512 per tile
    "mixed": None,          # long and empty simulations interleaved
}


def _lengths(name, rng):
    if name == "many_tiny":
        return rng.integers(0, 2, 3000)
    if name == "mixed":
        n = rng.integers(0, 3000, 40)
        n[::5] = 0
        n[7] = 66000          # many chunks without a simulation start
        return n
    return np.array(CASES[name])


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("reverse", [False, True], ids=["state", "costate"])
def test_batch_scan_matches_recurrence(dim, name, reverse):
    print(f"case: scan {name} dim={dim} reverse={reverse}", flush=True)
    rng = np.random.default_rng(11)
    n = _lengths(name, rng)
    offsets = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    total = int(offsets[-1])
    u = _unitaries(rng, total, dim)
    init = rng.normal(size=(len(n), dim)) + 1j * rng.normal(size=(len(n), dim))
    got, spin = _run(u, offsets, init, reverse, with_spin=not reverse)
    ref = _reference(u, offsets, init, reverse)
    assert np.abs(got - ref).max(initial=0) <= TOL
    if spin is not None:
        from oracle import spinsim_oracle as O
        assert np.abs(spin - O.get_spin(ref)).max(initial=0) <= TOL


@pytest.mark.parametrize("chunk_rows", [1024, 3072, 16384])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("reverse", [False, True], ids=["state", "costate"])
def test_batch_scan_chunk_sizes(chunk_rows, dim, reverse):
    """Explicit chunk sizes (the autotuner's candidates): one-tile chunks, and
    multi-tile chunks with simulations starting anywhere inside them."""
    print(f"case: scan chunk_rows={chunk_rows} dim={dim} reverse={reverse}", flush=True)
    rng = np.random.default_rng(14)
    n = _lengths("mixed", rng)
    offsets = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    total = int(offsets[-1])
    u = _unitaries(rng, total, dim)
    init = rng.normal(size=(len(n), dim)) + 1j * rng.normal(size=(len(n), dim))
    got, spin = _run(u, offsets, init, reverse, with_spin=not reverse, chunk_rows=chunk_rows)
    ref = _reference(u, offsets, init, reverse)
    assert np.abs(got - ref).max(initial=0) <= TOL
    if spin is not None:
        from oracle import spinsim_oracle as O
        assert np.abs(spin - O.get_spin(ref)).max(initial=0) <= TOL


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("reverse", [False, True], ids=["state", "costate"])
def test_batch_scan_multi_tile_chunks(dim, reverse):
    """Over 2^21 rows, so each workgroup scans a chunk of several tiles;
    checked against the C oracle's sequential product, one simulation at a
    time."""
    import fields as F
    from oracle.c_oracle import COracle

    print(f"case: scan multi-tile chunks dim={dim} reverse={reverse}", flush=True)
    rng = np.random.default_rng(13)
    n = rng.integers(0, 130000, 40)
    n[3] = 0
    n[11] = 1
    n[12] = 257
    offsets = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    total = int(offsets[-1])
    assert total > 1 << 21
    u = _unitaries(rng, total, dim)
    init = rng.normal(size=(len(n), dim)) + 1j * rng.normal(size=(len(n), dim))
    got, spin = _run(u, offsets, init, reverse, with_spin=not reverse)
    co = COracle(F.LARMOR.src if dim == 2 else F.RABI_ONE.src, dim)
    for s in range(len(n)):
        a, b = offsets[s], offsets[s + 1]
        if a == b:
            continue
        if not reverse:
            ref = co.get_state(init[s], u[a:b])
        else:
            # costate[i] = U[i+1]^+ costate[i+1]: the forward product of the
            # reversed, conjugate-transposed operators, reversed
            v = np.ascontiguousarray(u[a:b][::-1].conj().transpose(0, 2, 1))
            ref = co.get_state(init[s], v)[::-1]
        assert np.abs(got[a:b] - ref).max() <= TOL, s
        if spin is not None:
            assert np.abs(spin[a:b] - co.get_spin(ref)).max() <= TOL, s


def test_batch_scan_graph_replays():
    import torch

    from spinsim_amd import _native

    print("case: scan graph replays", flush=True)
    rng = np.random.default_rng(12)
    n = rng.integers(0, 2000, 30)
    offsets = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    total, d = int(offsets[-1]), 2
    u = _unitaries(rng, total, d)
    init = rng.normal(size=(len(n), d)) + 1j * rng.normal(size=(len(n), d))
    ref = _reference(u, offsets, init, False)
    L = _native.lib()
    du = torch.as_tensor(u, device="cuda")
    doff = torch.as_tensor(offsets, device="cuda")
    dinit = torch.as_tensor(init, device="cuda")
    out = torch.empty((total, d), dtype=torch.complex128, device="cuda")
    ws = torch.empty(int(L.ss_batch_scan_workspace(total, d, 0)), dtype=torch.uint8,
                     device="cuda")

    def run():
        s = torch.cuda.current_stream().cuda_stream
        _native.check(L.ss_batch_state_scan(dinit.data_ptr(), du.data_ptr(), out.data_ptr(), None,
                                            doff.data_ptr(), len(n), total, d, 0, 0,
                                            ws.data_ptr(), ws.numel(), s), "scan")

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run()
    for k in range(4):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert np.abs(out.cpu().numpy() - ref).max() <= TOL, k
        x = torch.randn(1000, device="cuda")
        _ = (x * 2).sum().item()
