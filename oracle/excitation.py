"""Excitation waveforms, restated from the reference for the oracle.

TEST INFRASTRUCTURE ONLY (see oracle/oe_oracle.cpp header).

Excitation::evaluate            src/fdtd/excitation.rs:150-200
Excitation::gaussian            src/fdtd/excitation.rs:67-85
ScheduledExcitation::from_excitation  src/fdtd/batch.rs:110-129  (f64 -> f32)
openEMS Gaussian (f0=0, fc)     matlab/examples/other/gauss_excitation_test.m:14-21
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np


def gaussian(t, t0, tau, amplitude=1.0):
    arg = (np.asarray(t, dtype=np.float64) - t0) / tau
    return amplitude * np.exp(-arg * arg)


def sinusoidal(t, f, amplitude=1.0):
    return amplitude * np.sin(2.0 * math.pi * f * np.asarray(t, dtype=np.float64))


def gaussian_modulated(t, f, t0, tau, amplitude=1.0):
    t = np.asarray(t, dtype=np.float64)
    arg = (t - t0) / tau
    return amplitude * np.exp(-arg * arg) * np.sin(2.0 * math.pi * f * t)


def dirac(t, amplitude=1.0):
    return amplitude * (np.abs(np.asarray(t, dtype=np.float64)) < 1e-15).astype(np.float64)


def step(t, amplitude=1.0):
    return amplitude * (np.asarray(t, dtype=np.float64) >= 0.0).astype(np.float64)


def custom(t: float, times: Sequence[float], values: Sequence[float], amplitude=1.0) -> float:
    """Linear interpolation with the reference's bracket search."""
    if t <= times[0]:
        v = values[0]
    elif t >= times[-1]:
        v = values[-1]
    else:
        i = 0
        while i < len(times) - 1 and times[i + 1] < t:
            i += 1
        a = (t - times[i]) / (times[i + 1] - times[i])
        v = values[i] + a * (values[i + 1] - values[i])
    return amplitude * v


def reference_gaussian_params(center_freq: float, bandwidth: float):
    f_max = center_freq * (1.0 + 0.5 * bandwidth)
    tau = 1.0 / (math.pi * f_max)
    return 4.0 * tau, tau


def openems_gaussian_params(fc: float):
    """(t0, tau) of the openEMS f0=0 Gaussian with cutoff fc."""
    return 9.0 / (2.0 * math.pi * fc), 3.0 / (2.0 * math.pi * fc)


def schedule(fn, start_timestep: int, n: int, dt: float) -> np.ndarray:
    """ScheduledExcitation::from_excitation: f32 samples at (start+s)*dt."""
    t = (start_timestep + np.arange(n, dtype=np.float64)) * dt
    return np.asarray(fn(t), dtype=np.float64).astype(np.float32)
