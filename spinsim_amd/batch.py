"""Batched evaluation: many simulations in one persistent launch.

The reference runs one ``evaluate`` per sweep point (spinsim/__init__.py:993;
example.py evaluates twice to sweep a bias).  ``BatchPlan`` packs B
simulations -- each with its own sweep row, start/end time, steps and initial
state -- into one CSR layout of coarse rows (``offsets[s]`` .. ``offsets[s+1]``
belong to simulation s), runs them in one persistent time-evolution launch
that hands out rows from a work counter (uneven durations balance
themselves), then one state scan per simulation, all stream-ordered and
capturable in a HIP graph.
"""
from __future__ import annotations

import math

import numpy as np

from . import _native
from .simulator import Results


def _per_sim(x, b, dtype, name):
    a = np.asarray(x, dtype)
    if a.ndim == 0:
        return np.full(b, a, dtype)
    if a.shape != (b,):
        raise ValueError(f"{name} must be a scalar or have shape ({b},), got {a.shape}")
    return a


def batch_steps(time_start, time_end, time_step_integration, time_step_output, b):
    """Per-simulation (dt, N, n_fine) with evaluate()'s arithmetic (:1025-1034, :832)."""
    t0 = _per_sim(time_start, b, np.float64, "time_start")
    t1 = _per_sim(time_end, b, np.float64, "time_end")
    dti = _per_sim(time_step_integration, b, np.float64, "time_step_integration")
    dto = _per_sim(time_step_output, b, np.float64, "time_step_output")
    dt = np.empty(b)
    n = np.empty(b, np.int64)
    nf = np.empty(b, np.int32)
    warned = 0
    for s in range(b):  # Python float arithmetic, exactly as evaluate() does it
        d_out, d_old = float(dto[s]), float(dti[s])
        d = d_out / round(max(d_out / d_old, 1))
        warned += math.fabs(d_out / d_old - round(d_out / d_old)) > 1e-6
        dt[s] = d
        n[s] = int((t1[s] - t0[s]) / d_out)
        nf[s] = math.floor(d_out / d + 0.5)
    if (n < 0).any():
        raise ValueError("time_end is before time_start for some simulation")
    if warned:
        print(f"\033[33mspinsim warning!!!\n{warned} simulation(s): time_step_output not an "
              f"integer multiple of time_step_integration; time_step_integration reset.\033[0m\n")
    return t0, dt, dto, n, nf


class BatchResults:
    """Results of B simulations in CSR layout (device tensors).

    ``time`` (total,), ``time_evolution`` (total, d, d), ``state`` (total, d),
    lazy ``spin`` (total, 3); rows of simulation s are
    ``offsets[s]:offsets[s+1]``.  ``results[s]`` is that simulation's
    :class:`Results` (numpy).  A compact plan's ``time_evolution`` is expanded
    from the stored first columns when first read.
    """

    def __init__(self, plan):
        self._plan = plan
        self.offsets = plan.offsets_host
        self.time = plan.time
        self.state = plan.state
        self._u = None if plan.compact else plan.u
        self._spin = None

    @property
    def time_evolution(self):
        if self._u is None:
            self._u = su2_from_columns(self._plan.u)
        return self._u

    @property
    def spin(self):
        if self._spin is None:
            out = getattr(self._plan, "spin_out", None)
            self._spin = out if out is not None else self._plan.sim._spin_calculator(self.state)
        return self._spin

    def __len__(self):
        return len(self.offsets) - 1

    def __getitem__(self, s):
        a, b = int(self.offsets[s]), int(self.offsets[s + 1])
        st = self.state[a:b]
        return Results(self.time[a:b].cpu().numpy(), self.time_evolution[a:b].cpu().numpy(),
                       st.cpu().numpy(), self._plan.sim.spin_calculator, _device_state=st)


# Rows of one scan tile (256 threads x 4 rows); chunk sizes are multiples.
SCAN_TILE_ROWS = 1024


def su2_from_columns(col):
    """(n, 2) first columns (a, b) -> (n, 2, 2) U = [[a, -conj b], [b, conj a]]."""
    import torch

    a, b = col[:, 0], col[:, 1]
    return torch.stack([a, -b.conj(), b, a.conj()], 1).resolve_conj().resolve_neg().view(-1, 2, 2)


class BatchPlan:
    """Device buffers and launch parameters for one batched workload.

    ``run()`` enqueues time evolution + state scan (+ Bloch vectors into
    ``spin_out`` when ``with_spin``) on the current stream;
    ``capture()`` records them into a HIP graph and ``replay()`` re-runs it.
    ``compact=True`` (spin half) stores each operator as its first column,
    ``u`` (total, 2): half the operator bytes written and read back by the
    scan, for workloads that want states and Bloch vectors.
    """

    def __init__(self, sim, time_start, time_end, time_step_integration, time_step_output,
                 state_init, sweep_parameters, *, lanes_per_step=None, items_per_grab=None,
                 grid_blocks=0, with_spin=False, compact=False):
        torch = sim._torch()
        self.sim = sim
        d = sim.dimension
        if compact and d != 2:
            raise ValueError("compact operator storage is for spin half (dimension 2)")
        self.compact = bool(compact)
        self.u_layout = _native.U_SU2_COLUMN if compact else _native.U_FULL
        # the state scan composes spin-half operators (special unitary) as
        # first columns, from either store
        self.scan_layout = (_native.U_SU2_COLUMN if compact else
                            _native.U_SU2_FULL if d == 2 else _native.U_FULL)
        sweep = np.asarray(sweep_parameters, np.float64)
        if sweep.ndim == 1:
            sweep = sweep[None, :]
        if sweep.ndim != 2:
            raise ValueError("sweep_parameters must be (P,) or (B, P)")
        b = sweep.shape[0]
        psi = np.asarray(state_init, np.complex128)
        if psi.ndim == 1:
            psi = np.broadcast_to(psi, (b, d))
        if psi.shape != (b, d):
            raise ValueError(f"state_init must be ({d},) or ({b}, {d})")
        for s in sim.field.samplers:
            if s.rows not in (1, b):
                raise ValueError(f"a sampler has {s.rows} amplitude rows for {b} simulations")
        t0, dt, dto, n, nf = batch_steps(time_start, time_end, time_step_integration,
                                         time_step_output, b)
        offsets = np.zeros(b + 1, np.int64)
        np.cumsum(n, out=offsets[1:])
        total = int(offsets[-1])
        dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.n_sims = b
        self.total = total
        self.offsets_host = offsets
        self.fine_steps = int((n * nf.astype(np.int64)).sum())
        f64 = torch.float64
        self.sweep = torch.as_tensor(np.ascontiguousarray(sweep), device=dev)
        self.offsets = torch.as_tensor(offsets, device=dev)
        self.t0 = torch.as_tensor(t0, dtype=f64, device=dev)
        self.dt = torch.as_tensor(dt, dtype=f64, device=dev)
        self.dt_out = torch.as_tensor(dto, dtype=f64, device=dev)
        self.n_fine = torch.as_tensor(nf, dtype=torch.int32, device=dev)
        self.psi0 = torch.as_tensor(np.ascontiguousarray(psi), device=dev)
        self.counter = torch.zeros(1, dtype=torch.int64, device=dev)
        self.time = torch.empty(total, dtype=f64, device=dev)
        self.u = torch.empty((total, 2) if compact else (total, d, d), dtype=torch.complex128,
                             device=dev)
        self.state = torch.empty((total, d), dtype=torch.complex128, device=dev)
        self.lanes = int(lanes_per_step or sim.lanes_per_step or 0)
        if self.lanes <= 0:
            self.lanes = default_batch_lanes(total, int(nf.max(initial=1)))
        self.items_per_grab = int(items_per_grab or 4)
        self.grid_blocks = int(grid_blocks)
        self.graph = None
        self.spin_out = (torch.empty((total, 3), dtype=f64, device=dev) if with_spin else None)
        self.scan_chunk_rows = 0
        self.scan_ws = torch.empty(max(self.scan_workspace_bytes(0), 1), dtype=torch.uint8,
                                   device=dev)

    def scan_workspace_bytes(self, chunk_rows):
        """Workspace of the state and costate scans at this chunk size (0:
        default chunking; else a positive multiple of SCAN_TILE_ROWS)."""
        if not self.total:
            return 0
        return int(_native.lib().ss_batch_scan_workspace(self.total, self.sim.dimension,
                                                         int(chunk_rows)))

    def set_scan_chunk_rows(self, chunk_rows):
        """Rows per workgroup chunk of the scans; grows the workspace if needed
        (not while a captured graph uses it)."""
        chunk_rows = int(chunk_rows)
        if chunk_rows < 0 or chunk_rows % SCAN_TILE_ROWS:
            raise ValueError(f"chunk_rows must be 0 or a positive multiple of {SCAN_TILE_ROWS}")
        need = self.scan_workspace_bytes(chunk_rows)
        if need > self.scan_ws.numel():
            if self.graph is not None:
                raise RuntimeError("the captured graph uses the scan workspace; set the chunk "
                                   "size before capture()")
            torch = self.sim._torch()
            self.scan_ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        self.scan_chunk_rows = chunk_rows

    def run_time_evolution(self, stream=None):
        torch = self.sim._torch()
        if stream is None:
            stream = torch.cuda.current_stream(self.device).cuda_stream
        mod = self.sim._module(self.device)
        _native.check(_native.lib().ss_batch_time_evolution(
            mod.handle, self.sweep.data_ptr(), self.sweep.shape[1], self.n_sims,
            self.offsets.data_ptr(), self.t0.data_ptr(), self.dt.data_ptr(),
            self.dt_out.data_ptr(), self.n_fine.data_ptr(), self.total, self.time.data_ptr(),
            self.u.data_ptr(), self.sim._table_ptr(self.device), self.counter.data_ptr(),
            self.lanes, self.items_per_grab, self.grid_blocks, self.u_layout, stream),
            "ss_batch_time_evolution")

    def run_state_scan(self, stream=None):
        torch = self.sim._torch()
        if stream is None:
            stream = torch.cuda.current_stream(self.device).cuda_stream
        spin = getattr(self, "spin_out", None)
        _native.check(_native.lib().ss_batch_state_scan(
            self.psi0.data_ptr(), self.u.data_ptr(), self.state.data_ptr(),
            None if spin is None else spin.data_ptr(), self.offsets.data_ptr(), self.n_sims,
            self.total, self.sim.dimension, self.scan_chunk_rows, self.scan_layout,
            self.scan_ws.data_ptr(), self.scan_ws.numel(), stream),
            "ss_batch_state_scan")

    def run(self):
        self.run_time_evolution()
        self.run_state_scan()
        return self

    def capture(self):
        """Record run() into a HIP graph (the counter is cleared by a kernel)."""
        torch = self.sim._torch()
        self.sim._module(self.device)
        self.sim._table_ptr(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.run()  # warm up outside the graph
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        self.graph = g
        return self

    def replay(self):
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()
        return self

    def results(self):
        return BatchResults(self)


def default_batch_lanes(total_rows, max_fine):
    """Lanes per coarse step for a batch: enough groups to fill the device."""
    target = 256 * 4 * 8 * 64
    g = 1
    while g < 64 and 2 * g <= max_fine and total_rows * g < target:
        g *= 2
    return g


def evaluate_batch(sim, time_start, time_end, time_step_integration, time_step_output, state_init,
                   sweep_parameters, **kw):
    """Evaluate B simulations in one launch; arguments may be per-simulation arrays."""
    return BatchPlan(sim, time_start, time_end, time_step_integration, time_step_output,
                     state_init, sweep_parameters, **kw).run().results()
