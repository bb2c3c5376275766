"""Interpolation samplers: generate_interpolation_sampler (spinsim/__init__.py:1941-2147).

The reference freezes (time, amplitude) into the numba-compiled sampler as
constants, so new samples mean a recompilation.  Here a sampler is a pair of
device arrays described by an ``ss_table`` record; the integrator reads them
at run time (``SS_SAMPLE(k, t)`` in the generated field code), and
``update()`` swaps in new amplitudes without recompiling.  Evaluation
semantics (uniform-grid index guess, bisection, zero-order hold / linear /
cubic Hermite with averaged finite-difference tangents, zero or clamped
boundaries) are those of :2026-2142.

A 2-D amplitude array (A, M) -- which the reference rejects with "Try
generating a single interpolator for each axis instead" (:2062-2063) -- is a
multi-axis sampler: A components on one time grid, all from one index search
per call (``sampler(t)`` returns A values; in field code ``v = sampler(t)``
then ``v[a]``, or ``sampler(t)[a]``).  Component a equals the one-axis
sampler of ``amplitude[a]``.  With ``per_simulation=True`` the leading axis
holds one row per simulation of a batch instead: (B, M), or (B, A, M) for
multi-axis rows; a single simulation reads row 0.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native


class SsTable(ctypes.Structure):
    """Host image of ``ss_table`` (include/spinsim_hip.h)."""

    _fields_ = [("time", ctypes.c_uint64), ("amplitude", ctypes.c_uint64),
                ("size", ctypes.c_int64), ("amplitude_stride", ctypes.c_int64),
                ("order", ctypes.c_int32), ("zero_boundary", ctypes.c_int32),
                ("axes", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("axis_stride", ctypes.c_int64)]


class InterpolationSampler:
    """Field source interpolated from (time, amplitude) samples.

    amplitude (M,): one component; (A, M): A components (a multi-axis
    sampler).  per_simulation=True: (B, M) or (B, A, M), one row per
    simulation of a batch.
    """

    def __init__(self, time, amplitude, order: int = 3, zero_boundary: bool = True,
                 per_simulation: bool = False):
        time = np.array(time, dtype=np.float64, copy=True)
        amplitude = np.array(amplitude, dtype=np.float64, copy=True)
        if time.ndim != 1 or time.size < 2:
            raise ValueError("time must be a 1-D array of at least two samples")
        shapes = "(B, M) or (B, A, M)" if per_simulation else "(M,) or (A, M)"
        if amplitude.ndim not in ((2, 3) if per_simulation else (1, 2)) \
                or amplitude.shape[-1] != time.size or 0 in amplitude.shape:
            raise ValueError(f"amplitude must have shape {shapes} with M = time.size")
        if order not in (0, 1, 3):
            raise NotImplementedError(
                f"Interpolation order {order} is not implemented. Try order 0 (zero order hold), "
                "1 (linear) or 3 (cubic).")
        self.time = time
        self.amplitude = amplitude
        self.per_simulation = bool(per_simulation)
        self.order = int(order)
        self.zero_boundary = bool(zero_boundary)
        self.version = 0
        self._device = {}
        self._bound = None

    @property
    def rows(self) -> int:
        """Simulation rows (1 unless per_simulation)."""
        return self.amplitude.shape[0] if self.per_simulation else 1

    @property
    def axes(self) -> int:
        """Components per sample (1 for a one-axis sampler)."""
        multi = self.amplitude.ndim == (3 if self.per_simulation else 2)
        return self.amplitude.shape[-2] if multi else 1

    def _fill(self, rec, time_ptr, amp_ptr):
        m = self.time.size
        rec.time = time_ptr
        rec.amplitude = amp_ptr
        rec.size = m
        rec.amplitude_stride = self.axes * m if self.per_simulation else 0
        rec.order = self.order
        rec.zero_boundary = int(self.zero_boundary)
        rec.axes = self.axes
        rec.reserved = 0
        rec.axis_stride = m

    def update(self, amplitude) -> None:
        """Replace the amplitudes (same shape); no recompilation needed."""
        amplitude = np.array(amplitude, dtype=np.float64, copy=True)
        if amplitude.shape != self.amplitude.shape:
            raise ValueError(f"amplitude shape {amplitude.shape} != {self.amplitude.shape}")
        self.amplitude = amplitude
        self.version += 1
        self._device.clear()

    def bind(self, amplitude_rows, stride: int):
        """Read the amplitudes from a device tensor view: row s of the table starts
        at ``amplitude_rows.data_ptr() + s*stride`` doubles (no copy).  Used for
        differentiable control amplitudes."""
        self._bound = (amplitude_rows, int(stride))
        self.version += 1

    def unbind(self):
        self._bound = None
        self.version += 1

    def device_arrays(self, device):
        """(time, amplitude) as float64 tensors on ``device`` (cached)."""
        import torch

        key = (str(device), self.version)
        if key not in self._device:
            self._device.clear()
            t = torch.as_tensor(self.time, device=device)
            a = torch.as_tensor(self.amplitude, device=device).contiguous()
            self._device[key] = (t, a)
        return self._device[key]

    def table_record(self, device) -> SsTable:
        t, a = self.device_arrays(device)
        rec = SsTable()
        self._fill(rec, t.data_ptr(), a.data_ptr())
        bound = getattr(self, "_bound", None)
        if bound is not None:
            rec.amplitude = bound[0].data_ptr()
            rec.amplitude_stride = bound[1]
        return rec

    def host_record(self):
        """(ss_table record, arrays it points at) for the CPU devices; the
        caller keeps the arrays alive while the record is in use."""
        if self._bound is not None:
            raise RuntimeError("a sampler bound to device amplitudes has no host table")
        arrays = (np.ascontiguousarray(self.time), np.ascontiguousarray(self.amplitude))
        rec = SsTable()
        self._fill(rec, arrays[0].ctypes.data, arrays[1].ctypes.data)
        return rec, arrays

    def _check_row(self, row):
        row = int(row)
        if not 0 <= row < self.rows:
            raise IndexError(f"row {row} of a sampler with {self.rows} simulation rows")
        return row

    def _shape_out(self, time_sample, flat):
        """flat (n*axes,) values -> the shape of time_sample (+ (axes,) when multi-axis)."""
        shape = np.shape(time_sample) + ((self.axes,) if self.axes > 1 else ())
        return flat.reshape(shape) if shape else flat[0]

    def evaluate_host(self, time_sample, row: int = 0):
        """Evaluate on the host (the host build of the device sampler)."""
        from . import host

        row = self._check_row(row)
        t = np.ascontiguousarray(np.atleast_1d(np.asarray(time_sample, np.float64)).ravel())
        out = np.empty(t.size * self.axes)
        arr, _keep = build_host_table_array([self])
        host._check(host.host_lib().ss_host_interp_eval(ctypes.addressof(arr), 0, row, self.axes,
                                                        t.ctypes.data, out.ctypes.data, t.size),
                    "ss_host_interp_eval")
        return self._shape_out(time_sample, out)

    def __call__(self, time_sample, row: int = 0):
        """Evaluate at the given time(s), like the reference's host-callable
        sampler: on the HIP device when one is visible, else on the host.
        A multi-axis sampler gives its A components along a last axis."""
        import torch

        if not torch.cuda.is_available():
            return self.evaluate_host(time_sample, row)
        row = self._check_row(row)
        device = torch.device("cuda", torch.cuda.current_device())
        t = torch.as_tensor(np.atleast_1d(np.asarray(time_sample, np.float64)).ravel(),
                            device=device)
        out = torch.empty(t.numel() * self.axes, dtype=torch.float64, device=device)
        tables = build_table_array([self], device)
        stream = torch.cuda.current_stream(device).cuda_stream
        _native.check(_native.lib().ss_interp_eval(tables.data_ptr(), 0, row, self.axes,
                                                   t.data_ptr(), out.data_ptr(), t.numel(),
                                                   stream),
                      "ss_interp_eval")
        return self._shape_out(time_sample, out.cpu().numpy())


def build_table_array(samplers, device):
    """A device byte tensor holding ss_table records for `samplers` (in slot order)."""
    import torch

    arr = (SsTable * max(len(samplers), 1))()
    keep = []
    for k, s in enumerate(samplers):
        arr[k] = s.table_record(device)
        keep.append(s.device_arrays(device))
    host = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8)
    dev = host.to(device)
    dev._ss_keep = keep  # the tables point into these tensors
    return dev


def build_host_table_array(samplers):
    """(ctypes ss_table array, the host arrays its records point at) for the
    CPU devices; keep both while the array is in use."""
    arr = (SsTable * max(len(samplers), 1))()
    keep = []
    for k, s in enumerate(samplers):
        arr[k], arrays = s.host_record()
        keep.append(arrays)
    return arr, keep


def generate_interpolation_sampler(time, amplitude, order=3, zero_boundary=True, device=None, *,
                                   per_simulation=False):
    """Build a sampler usable inside a field function (spinsim/__init__.py:1941).

    ``amplitude`` (M,) gives one component; (A, M) a multi-axis sampler of A
    components on one time grid (the reference raises for 2-D amplitudes,
    :2062-2063).  ``per_simulation=True`` reads the leading axis as one row
    per simulation of a batch: (B, M) or (B, A, M).  ``device`` is accepted
    for drop-in compatibility; samplers live on the HIP device (or the host
    for the CPU devices) and are callable from the host.
    """
    return InterpolationSampler(time, amplitude, order=order, zero_boundary=zero_boundary,
                                per_simulation=per_simulation)
