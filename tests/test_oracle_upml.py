"""Pins the UPML restatement (oracle, Upml of src/extensions/pml.rs):
region boxes of build_regions for an asymmetric layer set, coefficients of
init_region_coefficients at chosen cells against closed forms, the
pass-through cells, and the instability the reference documents for it
(src/fdtd/engine_testing.rs:2319-2354: energy growing by orders of magnitude
within tens of steps)."""
import math

import numpy as np
import pytest

from oracle import excitation as X
from oracle import oracle as O

EPS0 = 8.854187817e-12
Z0 = 376.730313668


def test_region_boxes():
    s = O.OracleSim(20, 18, 16, 1e-3)
    s.add_upml((4, 3, 2, 5, 3, 0))
    assert s.upml_regions() == [((0, 0, 0), (5, 18, 16)),      # x min: layers + 1 planes, whole y-z
                                ((16, 0, 0), (4, 18, 16)),     # x max: from nx - layers - 1
                                ((5, 0, 0), (11, 3, 16)),      # y min: x between the x regions
                                ((5, 12, 0), (11, 6, 16)),     # y max
                                ((5, 3, 0), (11, 9, 4))]       # z min: x and y between; no z max
    t = O.OracleSim(10, 10, 10, 1e-3)
    t.add_upml((0, 0, 0, 0, 0, 0))
    assert t.upml_regions() == []


def test_coefficients_closed_form():
    d = 1e-3
    s = O.OracleSim(20, 18, 16, d)
    s.add_upml((4, 3, 2, 5, 3, 0), grading_order=3.0, reflection_coeff=1e-6)
    dt = s.dt
    smax = -(4.0) * math.log(1e-6) / (2 * Z0 * 4 * d)
    f32 = lambda v: float(np.float32(v))
    # x-min region, outermost plane li = 0: depth 4 d -> kappa_v = smax for every component
    k = smax
    vv = s.upml_array(0, "vv", 1)[0, 7, 5]
    assert vv == f32((2 * EPS0 - k * dt) / (2 * EPS0 + k * dt))
    assert s.upml_array(0, "vvfo", 1)[0, 7, 5] == vv            # kappa_v is the same for every component
    assert s.upml_array(0, "vvfn", 1)[0, 7, 5] == 1.0
    # current grading: half a cell inside for the x component, outside (clamped) for y and z
    kin = smax * (3.5 / 4) ** 3
    kout = smax
    den = 2 * EPS0 + kout * dt  # denominator uses component npp = (c + 2) % 3
    assert s.upml_array(0, "ii", 0)[0, 3, 3] == f32((2 * EPS0 - kout * dt) / den)          # npp = z
    assert s.upml_array(0, "iifn", 0)[0, 3, 3] == f32((2 * EPS0 + kin * dt) / den)         # own = x
    assert s.upml_array(0, "iifo", 0)[0, 3, 3] == f32((2 * EPS0 - kin * dt) / den)
    den1 = 2 * EPS0 + kin * dt                                                              # c = 1: npp = x
    assert s.upml_array(0, "ii", 1)[0, 3, 3] == f32((2 * EPS0 - kin * dt) / den1)
    assert s.upml_array(0, "iifn", 1)[0, 3, 3] == f32((2 * EPS0 + kout * dt) / den1)
    # the innermost plane (depth 0) passes through: vv = vvfn = 1, vvfo = 0, same for ii
    for name, want in (("vv", 1.0), ("vvfn", 1.0), ("vvfo", 0.0), ("ii", 1.0), ("iifn", 1.0), ("iifo", 0.0)):
        assert np.all(s.upml_array(0, name, 2)[4] == want)
    # x-max region: depth counts from nx - layers - 1 = 16 outwards
    k3 = smax * (3 * d / (3 * d)) ** 3 * 0 + (-(4.0) * math.log(1e-6) / (2 * Z0 * 3 * d))  # layers 3 there
    assert s.upml_array(1, "vv", 0)[3, 0, 0] == f32((2 * EPS0 - k3 * dt) / (2 * EPS0 + k3 * dt))
    assert np.all(s.upml_array(1, "vv", 0)[0] == 1.0)


def test_reference_instability():
    """PmlAbsorbingScenario (src/fdtd/engine_testing.rs:1705-1734) on a
    smaller grid: the reference notes energy growing about 1e20 per 30
    steps with its UPML (:2324-2327); the restatement grows the same way."""
    n = 36
    s = O.OracleSim(n, n, n, 0.5e-3)
    s.add_upml((8,) * 6)
    t0, tau = X.reference_gaussian_params(3e9, 0.5)
    smp = X.schedule(lambda t: X.gaussian(t, t0, tau), 0, 90, s.dt).astype(np.float64)
    energy = []
    for b in range(3):
        s.run_batch(30, [(n // 2, n // 2, n // 2, 2, True)], smp[None, 30 * b:30 * (b + 1)])
        energy.append(float((s.fields().astype(np.float64) ** 2).sum()))
    assert energy[2] / energy[1] > 1e15
