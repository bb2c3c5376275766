"""Field callbacks used by the tests, each as HIP-C text (compiled by the
product and by the C oracle: the same source) and as a NumPy function (for
the vectorised NumPy oracle, lanes along the last axis)."""
from __future__ import annotations

import math

import numpy as np

TAU = 2 * math.pi


class Field:
    def __init__(self, name, src, fn, spin, sweep=(0.0,)):
        self.name, self.src, self.fn, self.spin, self.sweep = name, src, fn, spin, tuple(sweep)


def _larmor(t, s, f):
    f[0] = 0
    f[1] = 0
    f[2] = TAU * 1000


LARMOR = Field("larmor", """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = 0; f[1] = 0; f[2] = SS_TAU * 1000; }
""", _larmor, "half")


def _rabi_one(t, s, f):
    f[0] = TAU * 2e3 * np.cos(TAU * 7e5 * t)
    f[1] = 0
    f[2] = TAU * 7e5
    f[3] = TAU * 1e3


# Spin-one Rabi drive with quadratic shift (the headline workload).
RABI_ONE = Field("rabi_one", """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = SS_TAU * 2e3 * cos(SS_TAU * 7e5 * t); f[1] = 0; f[2] = SS_TAU * 7e5; f[3] = SS_TAU * 1e3; }
""", _rabi_one, "one")


def _rabi_sweep(t, s, f):
    f[0] = 2 * s[1] * np.cos(s[0] * t)
    f[1] = 0
    f[2] = s[0] + s[2]


# Swept Rabi drive (example.py:33-38 convention with a detuning).
RABI_SWEEP_HALF = Field("rabi_sweep_half", """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = 2 * s[1] * cos(s[0] * t); f[1] = 0; f[2] = s[0] + s[2]; }
""", _rabi_sweep, "half", (TAU * 7e5, TAU * 3e3, TAU * 150.0))


def _rabi_sweep_one(t, s, f):
    f[0] = 2 * s[1] * np.cos(s[0] * t)
    f[1] = s[1] * np.sin(s[0] * t)
    f[2] = s[0] + s[2]
    f[3] = s[3]


RABI_SWEEP_ONE = Field("rabi_sweep_one", """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = 2 * s[1] * cos(s[0] * t); f[1] = s[1] * sin(s[0] * t); f[2] = s[0] + s[2]; f[3] = s[3]; }
""", _rabi_sweep_one, "one", (TAU * 7e5, TAU * 3e3, TAU * -80.0, TAU * 2e3))


def _gaussian(t, s, f):
    f[0] = math.pi / math.sqrt(TAU) * np.exp(-0.5 * (t ** 2))
    f[1] = 0.0
    f[2] = 0.0


# gaussian_example.py:11-14 (a pi pulse along x, time in standard deviations).
GAUSSIAN = Field("gaussian", """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = SS_PI / sqrt(SS_TAU) * exp(-0.5 * (t * t)); f[1] = 0.0; f[2] = 0.0; }
""", _gaussian, "half")


def _tilted_one(t, s, f):
    f[0] = TAU * 3e4 * np.cos(TAU * 2e4 * t) + TAU * 5e3
    f[1] = TAU * 1e4 * np.sin(TAU * 3e4 * t)
    f[2] = TAU * 2e4 + TAU * 5e3 * np.sin(TAU * 1e3 * t)
    f[3] = TAU * 4e3 * np.cos(TAU * 7e3 * t)


# A strongly time-dependent field on every axis (no symmetry to hide errors).
TILTED_ONE = Field("tilted_one", """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = SS_TAU * 3e4 * cos(SS_TAU * 2e4 * t) + SS_TAU * 5e3; f[1] = SS_TAU * 1e4 * sin(SS_TAU * 3e4 * t);
  f[2] = SS_TAU * 2e4 + SS_TAU * 5e3 * sin(SS_TAU * 1e3 * t); f[3] = SS_TAU * 4e3 * cos(SS_TAU * 7e3 * t); }
""", _tilted_one, "one")


def _tilted_eight(t, s, f):
    _tilted_one(t, s, f)
    f[4] = TAU * 8e3 * np.cos(TAU * 1.5e4 * t)
    f[5] = TAU * 3e3
    f[6] = TAU * 6e3 * np.sin(TAU * 2.5e4 * t) - TAU * 2e3
    f[7] = TAU * 4e3 * np.cos(TAU * 5e3 * t)


# All eight su(3) directions of the eight-generator Lie-Trotter exponential.
TILTED_EIGHT = Field("tilted_eight", """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = SS_TAU * 3e4 * cos(SS_TAU * 2e4 * t) + SS_TAU * 5e3; f[1] = SS_TAU * 1e4 * sin(SS_TAU * 3e4 * t);
  f[2] = SS_TAU * 2e4 + SS_TAU * 5e3 * sin(SS_TAU * 1e3 * t); f[3] = SS_TAU * 4e3 * cos(SS_TAU * 7e3 * t);
  f[4] = SS_TAU * 8e3 * cos(SS_TAU * 1.5e4 * t); f[5] = SS_TAU * 3e3;
  f[6] = SS_TAU * 6e3 * sin(SS_TAU * 2.5e4 * t) - SS_TAU * 2e3; f[7] = SS_TAU * 4e3 * cos(SS_TAU * 5e3 * t); }
""", _tilted_eight, "one")


def _tilted_half(t, s, f):
    f[0] = TAU * 3e4 * np.cos(TAU * 2e4 * t) + TAU * 5e3
    f[1] = TAU * 1e4 * np.sin(TAU * 3e4 * t)
    f[2] = TAU * 2e4 + TAU * 5e3 * np.sin(TAU * 1e3 * t)


TILTED_HALF = Field("tilted_half", """
__device__ void get_field(double t, const double* s, double* f)
{ f[0] = SS_TAU * 3e4 * cos(SS_TAU * 2e4 * t) + SS_TAU * 5e3; f[1] = SS_TAU * 1e4 * sin(SS_TAU * 3e4 * t);
  f[2] = SS_TAU * 2e4 + SS_TAU * 5e3 * sin(SS_TAU * 1e3 * t); }
""", _tilted_half, "half")
