"""Differentiable simulations: gradients of a loss of the final states with
respect to interpolated control amplitudes (no reference counterpart: the
reference computes no gradients).

A :class:`ControlledSimulator` integrates, for B simulations at once,

    field(t) = static_field(t) + sum_k scale_k * interp_k(c[:, k, :])(t) e_{axis_k}

where each control k is an interpolation table (zero-order hold, linear or
cubic Hermite, exactly the sampler of generate_interpolation_sampler) on a
time grid, with amplitudes ``c`` (B, K, M) given as a float64 device tensor.
:meth:`ControlledSimulator.final_state` returns psi(T) (B, d) and is a
``torch.autograd.Function``: the forward pass runs the batched persistent
time-evolution launch and the per-simulation state scan and keeps only the
coarse step operators and states (checkpoints); the backward pass scans the
costates backward over the coarse operators, then recomputes every fine step
and accumulates dL/dc with the Frechet derivative of each step exponential
(``ss_batch_adjoint``).  In the rotating frame a z-axis control also moves
each coarse step's rotating frequency (the z field at the step's midpoint,
spinsim/__init__.py:823-829); that term is included, so the gradient is the
derivative of the computed forward map for every integration method.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .batch import BatchPlan
from .enums import ExponentiationMethod, IntegrationMethod
from .interpolation import InterpolationSampler
from .simulator import FieldSource, Simulator

AXES = {"x": 0, "y": 1, "z": 2, "q": 3}


class Control:
    """One control: field[axis] += scale * interpolation of amplitudes on ``time``."""

    def __init__(self, axis, time, order: int = 3, zero_boundary: bool = True, scale: float = 1.0):
        self.axis = AXES[axis] if isinstance(axis, str) else int(axis)
        if not 0 <= self.axis <= 3:
            raise ValueError("axis must be x, y, z or q (0..3)")
        self.time = np.asarray(time, np.float64)
        self.order = int(order)
        self.zero_boundary = bool(zero_boundary)
        self.scale = float(scale)
        self.sampler = InterpolationSampler(self.time, np.zeros(self.time.size), order,
                                            zero_boundary)


class SsControlDesc(ctypes.Structure):
    _fields_ = [("axis", ctypes.c_int32), ("table", ctypes.c_int32), ("scale", ctypes.c_double),
                ("grad_offset", ctypes.c_int64)]


def _combined_source(static, controls):
    """HIP-C text: the static field (renamed) plus the control terms; it sets
    SS_ADJ_CONTROLS, the adjoint kernel's number of per-lane gradient moment
    sums, to the control count, and SS_ADJ_AXES to the set of their axes."""
    axes = 0
    for c in controls:
        axes |= 1 << c.axis
    parts = [f"#define SS_ADJ_CONTROLS {len(controls)}\n#define SS_ADJ_AXES {axes}u\n"]
    if static is not None:
        parts.append("#define get_field ss_static_field\n" + static.source + "\n#undef get_field\n")
    call = ""
    if static is not None:
        call = ("    ss_static_field(t, s, f, ss_ctx_ref);\n" if static.has_context
                else "    ss_static_field(t, s, f);\n")
    n0 = len(static.samplers) if static is not None else 0
    terms = "".join(f"    f[{c.axis}] += {c.scale!r} * SS_SAMPLE({n0 + k}, t);\n"
                    for k, c in enumerate(controls))
    parts.append("__device__ void get_field(double t, const double* s, double* f, SS_CONTEXT)\n{\n"
                 + call + terms + "}\n")
    return "".join(parts)


class ControlledSimulator:
    """B simulations driven by differentiable interpolated controls."""

    def __init__(self, spin_quantum_number, controls, static_field=None, *, samplers=None,
                 exponentiation_method=None, use_rotating_frame: bool = True,
                 integration_method: IntegrationMethod = IntegrationMethod.MAGNUS_CF4,
                 number_of_squares: int = 24, precision: str = "fp64"):
        self.controls = list(controls)
        if not 1 <= len(self.controls) <= 4:
            raise ValueError("1 to 4 controls")
        if spin_quantum_number.dimension == 2 and any(c.axis == 3 for c in self.controls):
            raise ValueError("the quadratic shift (axis q) exists for spin one only")
        if exponentiation_method == ExponentiationMethod.LIE_TROTTER_8:
            raise ValueError("the adjoint differentiates the x, y, z and q generators; "
                             "ExponentiationMethod.LIE_TROTTER_8 is not supported here")
        static = FieldSource(static_field, samplers) if static_field is not None else None
        src = _combined_source(static, self.controls)
        table_samplers = (list(static.samplers) if static is not None else []) + \
            [c.sampler for c in self.controls]
        self.n_static_tables = len(static.samplers) if static is not None else 0
        field = FieldSource(src, samplers=table_samplers)
        self.sim = Simulator(field, spin_quantum_number, exponentiation_method=exponentiation_method,
                             use_rotating_frame=use_rotating_frame,
                             integration_method=integration_method,
                             number_of_squares=number_of_squares, precision=precision)
        self.dimension = spin_quantum_number.dimension
        offsets, off = [], 0
        for c in self.controls:
            offsets.append(off)
            off += c.time.size
        self.grad_offsets = offsets
        self.grad_stride = off

    def final_state(self, amplitudes, time_start, time_end, time_step_integration,
                    time_step_output, state_init, sweep_parameters=None, *,
                    deterministic=False):
        """psi(T) (B, d) complex128 of every simulation; differentiable in ``amplitudes``.

        ``amplitudes``: float64 device tensor (B, K, M_k) -- one row of control
        samples per simulation (all controls must share M when K > 1, or pass
        a list of (B, M_k) tensors).  ``deterministic=True`` sums each gradient
        entry's terms in a fixed order (bit-identical from run to run; needs a
        workspace of 72 B per coarse row and control, and coarse steps that
        span at most about five knot intervals); by default they are summed
        with float64 atomics, whose order -- and so the last bits -- can vary.
        """
        import torch

        amps = amplitudes
        if isinstance(amps, (list, tuple)):
            amps = torch.cat([a.reshape(a.shape[0], -1) for a in amps], 1)
        else:
            amps = amps.reshape(amps.shape[0], -1)
        if amps.dtype != torch.float64 or not amps.is_cuda:
            raise TypeError("amplitudes must be a float64 device tensor")
        if amps.shape[1] != self.grad_stride:
            raise ValueError(f"expected {self.grad_stride} control samples per simulation, "
                             f"got {amps.shape[1]}")
        return _FinalState.apply(amps, self, (time_start, time_end, time_step_integration,
                                              time_step_output, state_init, sweep_parameters,
                                              bool(deterministic)))

    # ---- used by the autograd function -----------------------------------------
    def _forward(self, amps, args):
        import torch

        t0, t1, dti, dto, psi0, sweep, _ = args
        b = amps.shape[0]
        amps = amps.detach().contiguous()
        for c, off in zip(self.controls, self.grad_offsets):
            c.sampler.bind(amps[:, off:off + c.time.size], self.grad_stride)
        if sweep is None:
            sweep = np.zeros((b, 1))
        sweep = np.asarray(sweep, np.float64)
        if sweep.ndim == 1:
            sweep = np.broadcast_to(sweep, (b, sweep.size))
        plan = BatchPlan(self.sim, t0, t1, dti, dto, psi0, sweep)
        plan.run()
        out = torch.empty((b, self.dimension), dtype=torch.complex128, device=plan.device)
        stream = torch.cuda.current_stream(plan.device).cuda_stream
        _native.check(_native.lib().ss_batch_final_state(
            plan.psi0.data_ptr(), plan.u.data_ptr(), plan.state.data_ptr(),
            plan.offsets.data_ptr(), out.data_ptr(), b, self.dimension, stream),
            "ss_batch_final_state")
        return out, plan, amps

    def _backward(self, plan, amps, grad_psi, deterministic=False):
        import torch

        b, d = plan.n_sims, self.dimension
        dev = plan.device
        stream = torch.cuda.current_stream(dev).cuda_stream
        lam = grad_psi.detach().to(torch.complex128).contiguous()
        costate = torch.empty((plan.total, d), dtype=torch.complex128, device=dev)
        L = _native.lib()
        _native.check(L.ss_batch_costate_scan(lam.data_ptr(), plan.u.data_ptr(),
                                              costate.data_ptr(), plan.offsets.data_ptr(), b,
                                              plan.total, d, plan.scan_chunk_rows,
                                              plan.scan_ws.data_ptr(), plan.scan_ws.numel(),
                                              stream),
                      "ss_batch_costate_scan")
        for c, off in zip(self.controls, self.grad_offsets):
            c.sampler.bind(amps[:, off:off + c.time.size], self.grad_stride)
        descs = (SsControlDesc * len(self.controls))()
        for k, (c, off) in enumerate(zip(self.controls, self.grad_offsets)):
            descs[k].axis = c.axis
            descs[k].table = self.n_static_tables + k
            descs[k].scale = c.scale
            descs[k].grad_offset = off
        grad = torch.empty((b, self.grad_stride), dtype=torch.float64, device=dev)
        status = torch.empty(1, dtype=torch.int32, device=dev)
        ws, ws_bytes = None, 0
        if deterministic:
            ws_bytes = int(L.ss_batch_adjoint_workspace(plan.total, len(self.controls)))
            ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        mod = self.sim._module(dev)
        _native.check(L.ss_batch_adjoint(
            mod.handle, plan.sweep.data_ptr(), plan.sweep.shape[1], b, plan.offsets.data_ptr(),
            plan.t0.data_ptr(), plan.dt.data_ptr(), plan.dt_out.data_ptr(),
            plan.n_fine.data_ptr(), plan.total, plan.u.data_ptr(), plan.state.data_ptr(),
            costate.data_ptr(), self.sim._table_ptr(dev), len(self.sim.field.samplers),
            ctypes.addressof(descs), len(self.controls), grad.data_ptr(), self.grad_stride,
            status.data_ptr(), None if ws is None else ws.data_ptr(), ws_bytes, stream),
            "ss_batch_adjoint")
        flags = int(status.item())
        if flags & 1:
            raise RuntimeError("ss_batch_adjoint: a step's rotation angle exceeds the derivative "
                               "series range; use a smaller time_step_integration")
        if flags & 2:
            raise RuntimeError("ss_batch_adjoint: a coarse step's samples reach beyond its "
                               "8-knot gradient window, which the deterministic mode needs; "
                               "use a shorter time_step_output or deterministic=False")
        if flags & 4:
            raise RuntimeError("ss_batch_adjoint: a control's axis is not among those the "
                               "module was compiled for (SS_ADJ_AXES)")
        return grad


def _final_state_function():
    import torch

    class _Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, amps, owner, args):
            out, plan, amps_c = owner._forward(amps, args)
            ctx.owner, ctx.plan, ctx.amps, ctx.deterministic = owner, plan, amps_c, args[-1]
            return out

        @staticmethod
        def backward(ctx, grad_out):
            grad = ctx.owner._backward(ctx.plan, ctx.amps, grad_out, ctx.deterministic)
            return grad, None, None

    return _Fn


class _Lazy:
    _fn = None

    @classmethod
    def apply(cls, *args):
        if cls._fn is None:
            cls._fn = _final_state_function()
        return cls._fn.apply(*args)


_FinalState = _Lazy
