"""Pins the UPML restatement (oracle, Upml of src/extensions/pml.rs):
region boxes of build_regions for an asymmetric layer set, coefficients of
init_region_coefficients at chosen cells against closed forms, the
pass-through cells, and the instability the reference documents for it
(src/fdtd/engine_testing.rs:2319-2354: energy growing by orders of magnitude
within tens of steps), and the flux update itself: three steps of pre -> Yee
-> post written out by hand in numpy for both halves, fields and flux arrays
compared at every step, with a check that the variant running the Yee update
on the field instead of the flux would be rejected.  Planting either that
variant or a flipped sign of the vvfo term in the oracle fails all three
layer sets."""
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


# ---------------------------------------------------------------- one-step pin
# The update of pre_voltage_update / post_voltage_update and
# pre_current_update / post_current_update (src/extensions/pml.rs:500-659)
# written out by hand in numpy around a numpy Yee half-step, then compared with
# the FP64 oracle after each of three steps: fields and both flux arrays.


def _yee_h(E, H, da, db):
    """update_h_basic (src/fdtd/basic_engine.rs:191-254): forward differences,
    zero past the last index."""
    def fwd(a, ax):
        d = np.zeros_like(a)
        sl = [slice(None)] * 3
        sl2 = [slice(None)] * 3
        sl[ax] = slice(0, -1)
        sl2[ax] = slice(1, None)
        d[tuple(sl)] = a[tuple(sl2)] - a[tuple(sl)]
        return d
    cx = fwd(E[2], 1) - fwd(E[1], 2)
    cy = fwd(E[0], 2) - fwd(E[2], 0)
    cz = fwd(E[1], 0) - fwd(E[0], 1)
    return np.stack([da[0] * H[0] + db[0] * cx, da[1] * H[1] + db[1] * cy, da[2] * H[2] + db[2] * cz])


def _yee_e(E, H, ca, cb):
    """update_e_basic (src/fdtd/basic_engine.rs:262-306): backward differences,
    the index-0 planes are never written."""
    out = E.copy()
    s = (slice(1, None),) * 3
    sx = (slice(0, -1), slice(1, None), slice(1, None))
    sy = (slice(1, None), slice(0, -1), slice(1, None))
    sz = (slice(1, None), slice(1, None), slice(0, -1))
    cx = (H[2][s] - H[2][sy]) - (H[1][s] - H[1][sz])
    cy = (H[0][s] - H[0][sz]) - (H[2][s] - H[2][sx])
    cz = (H[1][s] - H[1][sx]) - (H[0][s] - H[0][sy])
    out[0][s] = ca[0][s] * E[0][s] + cb[0][s] * cx
    out[1][s] = ca[1][s] * E[1][s] + cb[1][s] * cy
    out[2][s] = ca[2][s] * E[2][s] + cb[2][s] * cz
    return out


def _upml_half(F, flux, coef, regions, yee, swap_first=True):
    """One half-step with the UPML hooks around the Yee update `yee(F) -> F`.
    coef[q] = (c, cfo, cfn) per component for region q.  With swap_first=False
    the Yee update acts on the field itself instead of on the flux: the
    plausible mistake the pin must reject."""
    F = F.copy()
    helps = []
    for q, (st, sz) in enumerate(regions):
        box = tuple(slice(a, a + b) for a, b in zip(st, sz))
        c, cfo, _ = coef[q]
        fh = np.stack([c[n] * F[n][box] - cfo[n] * flux[q][n] for n in range(3)])
        helps.append(fh)
        if swap_first:
            for n in range(3):
                F[n][box] = flux[q][n]
    F = yee(F)
    new_flux = []
    for q, (st, sz) in enumerate(regions):
        box = tuple(slice(a, a + b) for a, b in zip(st, sz))
        _, _, cfn = coef[q]
        fn = np.stack([F[n][box] for n in range(3)])
        new_flux.append(fn)
        for n in range(3):
            F[n][box] = helps[q][n] + cfn[n] * fn[n]
    return F, new_flux


@pytest.mark.parametrize("layers", [(1, 0, 0, 0, 0, 0), (1, 0, 0, 1, 0, 2), (0, 2, 1, 0, 1, 0)])
def test_flux_update_hand_expanded(layers):
    n, d = 6, 1e-3
    s = O.OracleSim(n, n, n, d, prec=64)
    s.set_material(3.0, 1.0, 0.02, 0.0, (0, 4, 1, 5, 0, 3))
    s.add_upml(layers)
    rng = np.random.default_rng(7)
    f0 = rng.standard_normal((6, n, n, n))
    s.set_fields(f0)
    regions = s.upml_regions()
    assert regions, "the layer set must make at least one region"
    C = s.coefficients()
    ca, cb, da, db = C[0:3], C[3:6], C[6:9], C[9:12]
    ce = [tuple(np.stack([s.upml_array(q, w, m) for m in range(3)]) for w in ("vv", "vvfo", "vvfn"))
          for q in range(len(regions))]
    ch = [tuple(np.stack([s.upml_array(q, w, m) for m in range(3)]) for w in ("ii", "iifo", "iifn"))
          for q in range(len(regions))]
    # the graded coefficients must be far from pass-through somewhere, or the pin is vacuous
    assert min(float(np.min(c[0])) for c in ce) < 0.5
    E, H = f0[:3].copy(), f0[3:].copy()
    fe = [np.zeros((3,) + tuple(sz)) for _, sz in regions]
    fhx = [np.zeros((3,) + tuple(sz)) for _, sz in regions]
    Ev, Hv = E.copy(), H.copy()
    fev = [a.copy() for a in fe]
    fhv = [a.copy() for a in fhx]
    for step in range(3):
        H, fhx = _upml_half(H, fhx, ch, regions, lambda X: _yee_h(E, X, da, db))
        E, fe = _upml_half(E, fe, ce, regions, lambda X: _yee_e(X, H, ca, cb))
        Hv, fhv = _upml_half(Hv, fhv, ch, regions, lambda X: _yee_h(Ev, X, da, db), swap_first=False)
        Ev, fev = _upml_half(Ev, fev, ce, regions, lambda X: _yee_e(X, Hv, ca, cb), swap_first=False)
        s.run_batch(1)
        got = s.fields()
        want = np.concatenate([E, H])
        scale = float(np.abs(want).max())
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * scale)
        for q in range(len(regions)):
            for m in range(3):
                np.testing.assert_allclose(s.upml_array(q, "volt_flux", m), fe[q][m], rtol=0, atol=1e-12 * scale)
                np.testing.assert_allclose(s.upml_array(q, "curr_flux", m), fhx[q][m], rtol=0, atol=1e-12 * scale)
        # the Yee-on-the-field variant gives different fields: this check rejects it
        wrong = np.concatenate([Ev, Hv])
        assert float(np.abs(wrong - want).max()) > 1e-3 * scale
