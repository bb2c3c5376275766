"""The reference's public enums (spinsim/__init__.py:15-269), with a ROCm device.

Values and attributes follow the reference so scripts written against it
run unchanged; ``IntegrationMethod`` also carries the names MIDPOINT_SAMPLE
and HALF_STEP for the reference's EULER and HEUN.
"""
from __future__ import annotations

import math
from enum import Enum

import numpy as np


class SpinQuantumNumber(Enum):
    """Spin quantum number of the system (spinsim/__init__.py:15-66)."""

    def __init__(self, value: float, dimension: int, label: str):
        super().__init__()
        self._value_ = value
        self.dimension = dimension
        self.label = label
        if label == "half":
            self.plus_x = np.array([1, 1], np.complex128) / math.sqrt(2)
            self.minus_x = np.array([-1, 1], np.complex128) / math.sqrt(2)
            self.plus_y = np.array([1, 1j], np.complex128) / math.sqrt(2)
            self.minus_y = np.array([1, -1j], np.complex128) / math.sqrt(2)
            self.plus_z = np.array([1, 0], np.complex128)
            self.minus_z = np.array([0, 1], np.complex128)
        else:
            self.plus_x = np.array([1, math.sqrt(2), 1], np.complex128) / 2
            self.zero_x = np.array([-1, 0, 1], np.complex128) / math.sqrt(2)
            self.minus_x = np.array([1, -math.sqrt(2), 1], np.complex128) / 2
            self.plus_y = np.array([-1, -1j * math.sqrt(2), 1], np.complex128) / 2
            self.zero_y = np.array([1, 0, 1], np.complex128) / math.sqrt(2)
            self.minus_y = np.array([1, -1j * math.sqrt(2), 1], np.complex128) / 2
            self.plus_z = np.array([1, 0, 0], np.complex128)
            self.zero_z = np.array([0, 1, 0], np.complex128)
            self.minus_z = np.array([0, 0, 1], np.complex128)

    HALF = (1 / 2, 2, "half")
    ONE = (1, 3, "one")


class IntegrationMethod(Enum):
    """Integration rule per fine step (spinsim/__init__.py:68-93).

    MAGNUS_CF4: commutator-free fourth-order Magnus, two samples per step.
    MIDPOINT_SAMPLE (alias EULER): one exponential sampled at the midpoint.
    HALF_STEP (alias HEUN): two half steps sampled at the step's ends.
    """

    MAGNUS_CF4 = "magnus_cf4"
    MIDPOINT_SAMPLE = "euler"
    EULER = "euler"
    HALF_STEP = "heun"
    HEUN = "heun"

    @property
    def index(self) -> int:
        return {"magnus_cf4": 0, "euler": 1, "heun": 2}[self.value]


class ExponentiationMethod(Enum):
    """Matrix exponential used per step (spinsim/__init__.py:95-129)."""

    def __init__(self, value: str, index: int):
        super().__init__()
        self._value_ = value
        self.index = index

    ANALYTIC = ("analytic", 0)
    LIE_TROTTER = ("lie_trotter", 1)
    LIE_TROTTER_8 = ("lie_trotter_8", 2)


class Device(Enum):
    """Target device (spinsim/__init__.py:131-269).

    ROCM (alias ROC) runs the hand-written gfx950 HIP kernels.  The
    reference's CPU and CUDA targets are numba back ends; this package has
    no numba and no CUDA, so selecting them raises.
    """

    def __init__(self, value: str, index: int):
        super().__init__()
        self._value_ = value
        self.index = index

    PYTHON = ("python", 0)
    CPU_SINGLE = ("cpu_single", 0)
    CPU = ("cpu", 0)
    CUDA = ("cuda", 1)
    ROCM = ("rocm", 2)
    ROC = ("rocm", 2)
