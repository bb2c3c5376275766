"""spinsim_amd: spinsim's integration path, MI355X-native (gfx950 HIP kernels).

Drop-in for the reference package's public API (spinsim/__init__.py):
``Simulator``, ``Results``, ``SpinQuantumNumber``, ``IntegrationMethod``,
``ExponentiationMethod``, ``Device`` (plus ``Device.ROCM``) and
``generate_interpolation_sampler``.
"""
from .enums import Device, ExponentiationMethod, IntegrationMethod, SpinQuantumNumber
from .adjoint import Control, ControlledSimulator
from .batch import BatchPlan, BatchResults
from .interpolation import InterpolationSampler, generate_interpolation_sampler
from .simulator import FieldSource, Results, Simulator, precision_error
from .transpile import TranspileError, transpile

__all__ = [
    "BatchPlan", "BatchResults", "Control", "ControlledSimulator", "Device", "ExponentiationMethod", "IntegrationMethod", "SpinQuantumNumber",
    "InterpolationSampler", "generate_interpolation_sampler", "FieldSource", "Results",
    "Simulator", "TranspileError", "precision_error", "transpile",
]
