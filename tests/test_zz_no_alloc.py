"""ss_batch_state_scan_static (the batch state scan on the library's static
scratch, no allocation) against ss_batch_state_scan with a workspace, through
the C ABI: two calls in a row, the second smaller (records of the first call
left in the scratch must not leak into it), spin half and spin one, default
and explicit chunking; and the host-side rejection of a chunking that needs
more records than the scratch holds."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_unitaries(rng, n, d):
    z = rng.normal(size=(n, d, d)) + 1j * rng.normal(size=(n, d, d))
    q, r = np.linalg.qr(z)
    ph = np.diagonal(r, axis1=1, axis2=2)
    return q * (ph / np.abs(ph))[:, None, :]


def _case(rng, d, n_sims, max_rows):
    rows = rng.integers(0, max_rows, n_sims)
    rows[0] = max(rows[0], 1)
    offsets = np.concatenate([[0], np.cumsum(rows)]).astype(np.int64)
    total = int(offsets[-1])
    u = _random_unitaries(rng, total, d)
    psi = rng.normal(size=(n_sims, d)) + 1j * rng.normal(size=(n_sims, d))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    return offsets, u, psi


def _scan_both(torch, L, d, offsets, u, psi, chunk_rows, stream):
    dev = torch.device("cuda", 0)
    total = int(offsets[-1])
    n_sims = len(offsets) - 1
    off = torch.as_tensor(offsets, device=dev)
    u_d = torch.as_tensor(u, device=dev)
    psi_d = torch.as_tensor(psi, device=dev)
    out = {}
    for name in ("workspace", "static"):
        state = torch.full((total, d), float("nan"), dtype=torch.complex128, device=dev)
        spin = torch.full((total, 3), float("nan"), dtype=torch.float64, device=dev)
        if name == "workspace":
            nbytes = L.ss_batch_scan_workspace(total, d, chunk_rows)
            assert nbytes > 0
            ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            rc = L.ss_batch_state_scan(psi_d.data_ptr(), u_d.data_ptr(), state.data_ptr(),
                                       spin.data_ptr(), off.data_ptr(), n_sims, total, d,
                                       chunk_rows, 0, ws.data_ptr(), nbytes, stream)
        else:
            rc = L.ss_batch_state_scan_static(psi_d.data_ptr(), u_d.data_ptr(), state.data_ptr(),
                                              spin.data_ptr(), off.data_ptr(), n_sims, total, d,
                                              chunk_rows, stream)
        assert rc == 0, L.ss_last_error().decode()
        torch.cuda.synchronize()
        out[name] = (state.cpu().numpy(), spin.cpu().numpy())
    return out


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("chunk_rows", [0, 2048])
def test_static_scan_matches_workspace_scan(d, chunk_rows):
    import torch

    from spinsim_amd import _native

    print(f"case: static-scratch batch scan d={d} chunk_rows={chunk_rows}", flush=True)
    L = _native.lib()
    stream = torch.cuda.current_stream(0).cuda_stream
    rng = np.random.default_rng(31 + d)
    # first call: many rows (hundreds of chunks); second: far fewer rows and chunks
    for n_sims, max_rows in ((300, 4000), (7, 900)):
        offsets, u, psi = _case(rng, d, n_sims, max_rows)
        out = _scan_both(torch, L, d, offsets, u, psi, chunk_rows, stream)
        st_w, sp_w = out["workspace"]
        st_s, sp_s = out["static"]
        assert np.isfinite(st_s).all() and np.isfinite(sp_s).all()
        # the same kernels on the same chunking: identical results
        assert np.array_equal(st_s, st_w)
        assert np.array_equal(sp_s, sp_w)
        # and the states are the sequential product (float64 rounding)
        for s in (0, n_sims - 1):
            a, b = int(offsets[s]), int(offsets[s + 1])
            v = psi[s].copy()
            for i in range(a, b):
                assert np.abs(st_s[i] - v).max() <= 1e-12
                v = u[i] @ v


def test_static_scan_rejects_too_many_chunks():
    import torch

    from spinsim_amd import _native

    print("case: static-scratch batch scan, chunking beyond the scratch", flush=True)
    L = _native.lib()
    stream = torch.cuda.current_stream(0).cuda_stream
    dev = torch.device("cuda", 0)
    total = 2049 * 1024  # 2049 chunks of one 1024-row tile
    offsets = torch.tensor([0, total], dtype=torch.int64, device=dev)
    psi = torch.zeros((1, 2), dtype=torch.complex128, device=dev)
    u = torch.zeros((total, 2, 2), dtype=torch.complex128, device=dev)
    state = torch.zeros((total, 2), dtype=torch.complex128, device=dev)
    rc = L.ss_batch_state_scan_static(psi.data_ptr(), u.data_ptr(), state.data_ptr(), None,
                                      offsets.data_ptr(), 1, total, 2, 1024, stream)
    assert rc != 0
    msg = L.ss_last_error().decode()
    assert "2049 chunks" in msg and "at most 2048" in msg
    # the default chunking of the same rows fits
    psi[0, 0] = 1.0
    u[:, 0, 0] = 1.0
    u[:, 1, 1] = 1.0
    rc = L.ss_batch_state_scan_static(psi.data_ptr(), u.data_ptr(), state.data_ptr(), None,
                                      offsets.data_ptr(), 1, total, 2, 0, stream)
    assert rc == 0, L.ss_last_error().decode()
    torch.cuda.synchronize()
    assert bool((state[:, 0] == 1.0).all()) and bool((state[:, 1] == 0.0).all())
