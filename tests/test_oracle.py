"""Pins the CPU oracle (oracle/oe_oracle.cpp) before it is trusted as the checker.

No reference output can be produced here (the Rust crate cannot be built and
no golden vectors exist in the reference), so every check is a closed form, a
known answer from the reference's own unit tests, a hand-derived brute-force
value on a tiny grid, or a physical invariant -- each chosen so that a
plausible restatement mistake (sign, index, dropped term) fails.
"""
import math

import numpy as np
import pytest

from oracle import excitation as X
from oracle import oracle as O

C0 = 299792458.0
EPS0 = 8.854187817e-12
MU0 = 1.256637062e-6
Z0 = 376.730313668


def f32(x):
    return np.float32(x)


# ---------------------------------------------------------------- operator
def test_timestep_closed_form():
    # Operator::calculate_timestep (src/fdtd/operator.rs:82-89)
    for n, d in ((100, 1e-3), (256, 0.5e-3)):
        s = O.OracleSim(n, n, n, d)
        assert s.dt == pytest.approx(0.99 * d / (C0 * math.sqrt(3.0)), rel=1e-15)
    s = O.OracleSim(4, 5, 6, (1e-3, 2e-3, 0.5e-3))
    inv = 1 / 1e-6 + 1 / 4e-6 + 1 / 0.25e-6
    assert s.dt == pytest.approx(0.99 / (C0 * math.sqrt(inv)), rel=1e-15)


def test_vacuum_coefficients():
    s = O.OracleSim(5, 5, 5, (1e-3, 2e-3, 0.5e-3))
    c = s.coefficients()
    dt = s.dt
    for d, dl in enumerate((1e-3, 2e-3, 0.5e-3)):
        assert np.all(c[d] == 1.0) and np.all(c[6 + d] == 1.0)
        assert np.all(c[3 + d] == f32(dt / (EPS0 * dl)))
        assert np.all(c[9 + d] == f32(-(dt / (MU0 * dl))))  # negative Db (operator.rs:129-131)


def test_set_material_known_answers():
    s = O.OracleSim(6, 6, 6, 1e-3)
    cb0 = s.coefficients()[3, 0, 0, 0]
    s.set_material(4.0, 1.0, 0.0, 0.0, (0, 6, 0, 6, 0, 6))
    c = s.coefficients()
    # reference unit test: cb(eps_r=4) / cb(vacuum) == 0.25 (src/fdtd/operator.rs:371)
    assert c[3, 2, 2, 2] / cb0 == pytest.approx(0.25, rel=1e-6)
    # lossy: Ca = (1 - s dt/2e)/(1 + s dt/2e)
    s2 = O.OracleSim(6, 6, 6, 1e-3)
    sig = 0.1
    s2.set_material(4.0, 2.0, sig, 5.0, (1, 3, 0, 6, 2, 4))
    c2 = s2.coefficients()
    dt = s2.dt
    le = sig * dt / (2 * 4 * EPS0)
    lm = 5.0 * dt / (2 * 2 * MU0)
    assert c2[0, 1, 0, 2] == f32((1 - le) / (1 + le))
    assert c2[3, 1, 0, 2] == f32((dt / (4 * EPS0)) / (1 + le) / 1e-3)
    assert c2[6, 2, 5, 3] == f32((1 - lm) / (1 + lm))
    assert c2[9, 2, 5, 3] == f32(-(dt / (2 * MU0)) / (1 + lm) / 1e-3)
    # outside the half-open region untouched
    assert c2[0, 3, 0, 2] == 1.0 and c2[0, 1, 0, 4] == 1.0


def test_later_boxes_overwrite_and_pec():
    s = O.OracleSim(6, 6, 6, 1e-3)
    s.set_material(4.0, 1.0, 0.0, 0.0, (0, 6, 0, 6, 0, 6))
    s.set_material(9.0, 1.0, 0.0, 0.0, (2, 4, 2, 4, 2, 4))
    s.set_pec((3, 4, 3, 4, 3, 4))
    c = s.coefficients()
    cb0 = f32(s.dt / EPS0 / 1e-3)
    assert c[3, 0, 0, 0] == f32(s.dt / (4 * EPS0) / 1e-3)
    assert c[3, 2, 2, 2] == f32(s.dt / (9 * EPS0) / 1e-3)
    assert c[0, 3, 3, 3] == 0 and c[3, 3, 3, 3] == 0 and c[5, 3, 3, 3] == 0
    assert c[6, 3, 3, 3] == 1.0  # PEC leaves H coefficients (operator.rs:249-263)
    assert cb0 > 0


# -------------------------------------------------------------- excitation
def test_excitation_known_answers():
    # reference unit tests (src/fdtd/excitation.rs:358-418)
    assert X.custom(0.5, [0, 1, 2, 3], [0, 1, 0.5, 0]) == pytest.approx(0.5, abs=1e-12)
    assert X.custom(1.5, [0, 1, 2, 3], [0, 1, 0.5, 0]) == pytest.approx(0.75, abs=1e-12)
    assert X.custom(-1.0, [0, 1, 2, 3], [0, 1, 0.5, 0]) == 0.0
    assert X.custom(5.0, [0, 1, 2, 3], [0, 1, 0.5, 0]) == 0.0
    assert X.dirac(0.0) == 1.0 and X.dirac(1e-9) == 0.0 and X.dirac(-1e-9) == 0.0
    assert X.step(0.0, 3.0) == 3.0 and X.step(-1e-9, 3.0) == 0.0
    t0, tau = X.reference_gaussian_params(1e9, 0.5)
    assert X.gaussian(t0, t0, tau) == pytest.approx(1.0, abs=1e-12)
    assert X.gaussian(0.0, t0, tau) < 0.01
    assert X.gaussian(t0, t0, tau, 2.5) == pytest.approx(2.5, abs=1e-12)
    assert X.sinusoidal(0.25e-9, 1e9) == pytest.approx(1.0, abs=1e-10)
    assert X.sinusoidal(1e-9, 1e9) == pytest.approx(X.sinusoidal(0.0, 1e9), abs=1e-10)


def test_openems_gaussian_spectrum():
    # openEMS Gaussian (f0=0, fc): |F(f)| = sigma sqrt(2 pi) exp(-(sigma 2 pi f)^2 / 2)
    # with sigma = tau / sqrt(2) (matlab/examples/other/gauss_excitation_test.m:14-21,50)
    fc = 15e9
    t0, tau = X.openems_gaussian_params(fc)
    assert tau == pytest.approx(3 / (2 * math.pi * fc)) and t0 == pytest.approx(3 * tau)
    dt = 1e-13
    t = np.arange(0, 20000) * dt
    v = X.gaussian(t, t0, tau)
    sigma = tau / math.sqrt(2)
    for f in (1e9, 7.5e9, 15e9):
        F = np.abs(np.sum(v * np.exp(-2j * math.pi * f * t)) * dt)
        ref = sigma * math.sqrt(2 * math.pi) * math.exp(-0.5 * (sigma * 2 * math.pi * f) ** 2)
        assert F == pytest.approx(ref, rel=1e-6)


def test_schedule_is_f32_of_f64():
    dt = 1.9e-12
    t0, tau = X.openems_gaussian_params(15e9)
    s = X.schedule(lambda t: X.gaussian(t, t0, tau), 10, 50, dt)
    assert s.dtype == np.float32
    ref = np.array([np.float32(X.gaussian((10 + q) * dt, t0, tau)) for q in range(50)])
    assert np.array_equal(s, ref)


# --------------------------------------------------------------- Yee step
def _single_step(prec, set_field):
    s = O.OracleSim(5, 5, 5, 1e-3, prec)
    f = np.zeros((6, 5, 5, 5))
    set_field(f)
    s.set_fields(f)
    s.run_batch(1)
    return s, s.fields()


def test_one_step_discrete_laplacian():
    """Ez impulse at (2,2,2): after one leapfrog step the curl-curl gives
    Ez = 1 - 4 (c dt / d)^2 and the four surrounding H edges = +-Db."""
    def init(f):
        f[2, 2, 2, 2] = 1.0

    for prec in (32, 64):
        s, F = _single_step(prec, init)
        dt = s.dt
        db = -(dt / (MU0 * 1e-3))
        cb = dt / (EPS0 * 1e-3)
        tol = 1e-6 if prec == 32 else 1e-13
        assert F[3, 2, 1, 2] == pytest.approx(db, rel=tol)   # Hx(j-1): dEz/dy = +1
        assert F[3, 2, 2, 2] == pytest.approx(-db, rel=tol)  # Hx(j):   dEz/dy = -1
        assert F[4, 1, 2, 2] == pytest.approx(-db, rel=tol)  # Hy(i-1): -dEz/dx = -1
        assert F[4, 2, 2, 2] == pytest.approx(db, rel=tol)   # Hy(i):   +1
        assert np.count_nonzero(F[3:]) == 4
        expect = 1.0 + cb * 4.0 * db  # = 1 - 4 (c dt/d)^2
        assert expect == pytest.approx(1 - 4 * (C0 * dt / 1e-3) ** 2, rel=1e-8)  # eps0 mu0 c0^2 = 1 + O(1e-9)
        assert F[2, 2, 2, 2] == pytest.approx(expect, rel=tol, abs=tol)
        # the four neighbouring Ez edges receive +(c dt/d)^2
        for (i, j) in ((1, 2), (3, 2), (2, 1), (2, 3)):
            assert F[2, i, j, 2] == pytest.approx(-cb * db, rel=tol)


def test_boundary_planes_reference_semantics():
    """E on the index-0 planes is never updated; a difference past the last
    index is zero (src/fdtd/basic_engine.rs:194-306)."""
    s = O.OracleSim(6, 6, 6, 1e-3, 64)
    f = np.zeros((6, 6, 6, 6))
    f[0, 0, 3, 3] = 2.0     # Ex on the i = 0 plane
    f[1, 3, 0, 3] = -1.0    # Ey on the j = 0 plane
    f[2, 5, 5, 5] = 1.0     # Ez at the top corner
    s.set_fields(f)
    s.run_batch(3)
    F = s.fields()
    assert F[0, 0, 3, 3] == 2.0
    assert F[1, 3, 0, 3] == -1.0
    assert np.all(F[:3, 0, 1:, 1:][F[:3, 0, 1:, 1:] != 0] == F[:3, 0, 1:, 1:][F[:3, 0, 1:, 1:] != 0])
    # Top Ez: Hy(i=5) sees dEz/dx = 0 (past the last index); Hx(j=5) likewise.
    s2 = O.OracleSim(6, 6, 6, 1e-3, 64)
    s2.set_fields(f * 0 + np.where(np.indices((6, 6, 6, 6))[0] == 2, f, 0))
    s2.run_batch(1)
    F2 = s2.fields()
    assert F2[4, 5, 5, 5] == 0.0 and F2[3, 5, 5, 5] == 0.0
    assert F2[4, 4, 5, 5] != 0.0 and F2[3, 5, 4, 5] != 0.0


def _leapfrog_energy_series(prec, steps=200, seed=0):
    rng = np.random.default_rng(seed)
    n = 12
    s = O.OracleSim(n, n, n, 1e-3, prec)
    # The reference's upper faces (zero difference past the last index) are not
    # skew-adjoint; true PEC on the top planes (E fixed at 0) makes the whole
    # operator so, leaving the interior stencil as the thing under test.
    s.set_pec((n - 1, n, 0, n, 0, n))
    s.set_pec((0, n, n - 1, n, 0, n))
    s.set_pec((0, n, 0, n, n - 1, n))
    f = np.zeros((6, n, n, n))
    f[:3, 1:-1, 1:-1, 1:-1] = rng.standard_normal((3, n - 2, n - 2, n - 2))
    s.set_fields(f)
    W = []
    prevH = None
    prevE = None
    for _ in range(steps):
        F = s.fields().astype(np.float64)
        if prevH is not None:
            # W^n = eps |E^n|^2 + mu H^{n-1/2} . H^{n+1/2}
            W.append(EPS0 * np.sum(prevE ** 2) + MU0 * np.sum(prevH * F[3:]))
        prevE, prevH = F[:3], F[3:]
        s.run_batch(1)
    return np.array(W)


def test_discrete_energy_invariant_fp64():
    """Leapfrog Yee with D- = -(D+)^T conserves eps|E^n|^2 + mu H^{n-1/2}.H^{n+1/2}
    exactly; a wrong neighbour or sign in any curl term breaks it."""
    W = _leapfrog_energy_series(64)
    assert np.max(np.abs(W - W[0])) / W[0] < 1e-12


def test_discrete_energy_invariant_fp32():
    W = _leapfrog_energy_series(32)
    assert np.max(np.abs(W - W[0])) / W[0] < 1e-5


def test_fp32_tracks_fp64():
    out = []
    for prec in (32, 64):
        s = O.OracleSim(20, 20, 20, 0.5e-3, prec)
        s.set_material(4.0, 1.0, 0.0, 0.0, (0, 20, 0, 20, 8, 12))
        t0, tau = X.reference_gaussian_params(3e9, 0.5)
        smp = X.schedule(lambda t: X.gaussian(t, t0, tau), 0, 300, s.dt).astype(np.float64)[None]
        s.run_batch(300, [(10, 10, 5, 2, True)], smp)
        out.append(s.fields().astype(np.float64))
    rel = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[1])
    assert rel < 1e-5


def test_zero_excitation_stays_zero():
    # ZeroExcitationScenario (src/fdtd/engine_testing.rs:1193-1230)
    s = O.OracleSim(20, 20, 20, 0.5e-3)
    s.run_batch(100)
    assert not np.any(s.fields())


def test_pec_cavity_energy_bounded():
    # PecCavityScenario (src/fdtd/engine_testing.rs:782-839): Dirac in a 20^3 cavity
    s = O.OracleSim(20, 20, 20, 0.5e-3, 64)
    smp = X.schedule(lambda t: X.dirac(t), 0, 200, s.dt).astype(np.float64)[None]
    r = s.run_batch(200, [(10, 10, 10, 2, True)], smp, sample_interval=10, track_peak=True)
    tot = r.energy_samples[:, 1] + r.energy_samples[:, 2]
    assert np.all(np.isfinite(tot)) and tot.max() < 10.0


def test_hard_source_sets_value():
    s = O.OracleSim(8, 8, 8, 1e-3)
    smp = np.array([[0.25, 0.5, 0.75]])
    s.run_batch(3, [(4, 4, 4, 0, False)], smp)
    assert s.fields()[0, 4, 4, 4] == np.float32(0.75)
    # steps past the schedule apply nothing (Sampled(...).get(step) == None)
    s.run_batch(2, [(4, 4, 4, 0, False)], smp[:, :0].copy())


def test_energy_sampling_and_decay_termination():
    s = O.OracleSim(10, 10, 10, 1e-3)
    smp = np.zeros((1, 400))
    smp[0, 0] = 1.0
    r = s.run_batch(400, [(5, 5, 5, 2, True)], smp, sample_interval=7, track_peak=True)
    assert list(r.energy_samples[:5, 0]) == [0, 7, 14, 21, 28]
    e, h = s.energy()
    F = s.fields().astype(np.float64)
    assert e == pytest.approx(np.sum(F[:3] ** 2), rel=1e-12)
    assert h == pytest.approx(np.sum(F[3:] ** 2), rel=1e-12)
    # A lossy box decays; the batch must stop at a check step.
    s2 = O.OracleSim(10, 10, 10, 1e-3)
    s2.set_material(1.0, 1.0, 5.0, 0.0, (0, 10, 0, 10, 0, 10))
    r2 = s2.run_batch(100000, [(5, 5, 5, 2, True)], smp[:, :1].copy(), check_interval=50,
                      energy_decay_db=30.0, sample_interval=1, track_peak=True)
    assert r2.termination_reason == 1
    assert r2.final_decay_db < -30.0
    assert (r2.timesteps_executed - 1) % 50 == 0


# ------------------------------------------------------------------- CPML
def test_cpml_coefficient_closed_forms():
    n, L, d = 40, 8, 0.5e-3
    s = O.OracleSim(n, n, n, d)
    s.add_cpml((L,) * 6, 3.0, 1e-6, 0.0, 1.0, O.MODE_COMPAT)
    dt = s.dt
    smax = -(3 + 1) * math.log(1e-6) / (2 * Z0 * L * d)  # src/extensions/cpml.rs:68-72
    be = s.cpml_array(0, "E", "b")
    ae = s.cpml_array(0, "E", "a")
    bh = s.cpml_array(1, "H", "b")
    # x-min, E: depth (L - i) d -> outermost layer i=0 has rho = 1
    assert be[0] == f32(math.exp(-smax / EPS0 * dt))
    assert be[L - 1] == f32(math.exp(-smax * (1 / L) ** 3 / EPS0 * dt))
    # alpha = 0, kappa = 1: a = 1 - b (computed in double)
    assert ae[0] == f32(1 - math.exp(-smax / EPS0 * dt))
    # x-max, H: depth (i + 1) d - d/2
    assert bh[0] == f32(math.exp(-smax * (0.5 / L) ** 3 / EPS0 * dt))
    assert s.cpml_array(0, "E", "pos").size == L * n * n
    # a face is skipped when layers >= grid size (src/extensions/cpml.rs:262)
    s2 = O.OracleSim(6, 6, 6, d)
    s2.add_cpml((8, 0, 2, 2, 2, 2))
    assert s2.cpml_array(0, "E", "pos") is None and s2.cpml_array(1, "E", "pos") is None
    assert s2.cpml_array(2, "E", "pos").size == 2 * 6 * 6


def test_cpml_compat_psi_first_step():
    """First step from psi = 0 and H = 0: psi_H = a_h * dE, psi_E = a_e * dH,
    and the fields are those of the run without CPML (correction disabled,
    src/extensions/cpml.rs:343-421)."""
    n, L = 12, 3
    rng = np.random.default_rng(1)
    f = np.zeros((6, n, n, n))
    f[:3, 1:, 1:, 1:] = rng.standard_normal((3, n - 1, n - 1, n - 1))
    a = O.OracleSim(n, n, n, 1e-3, 64)
    a.add_cpml((L,) * 6, mode=O.MODE_COMPAT)
    b = O.OracleSim(n, n, n, 1e-3, 64)
    a.set_fields(f)
    b.set_fields(f)
    a.run_batch(1)
    b.run_batch(1)
    assert np.array_equal(a.fields(), b.fields())
    E = f[:3]
    ah = a.cpml_array(0, "H", "a")
    ppos = a.cpml_array(0, "H", "pos").reshape(L, n, n)
    pneg = a.cpml_array(0, "H", "neg").reshape(L, n, n)
    for l in range(L):
        dEz = E[2, l + 1] - E[2, l]
        dEy = E[1, l + 1] - E[1, l]
        assert np.allclose(ppos[l], ah[l] * dEz, rtol=1e-13, atol=0)
        assert np.allclose(pneg[l], ah[l] * dEy, rtol=1e-13, atol=0)
    # z-max face of E: psi_pos <- dHy/dz with the H after this step's update
    H = b.fields()[3:]
    aez = a.cpml_array(5, "E", "a")
    pz = a.cpml_array(5, "E", "pos").reshape(n, n, L)
    for l in range(L):
        k = n - L + l
        assert np.allclose(pz[:, :, l], aez[l] * (H[1][:, :, k] - H[1][:, :, k - 1]), rtol=1e-13, atol=0)


def _radiated_energy(n, L, steps, mode, prec=32):
    """Energy history of a zero-DC modulated Gaussian (15 GHz) radiated by a
    soft Ez source at the centre of an n^3 box with an L-cell CPML."""
    f0, tau = 15e9, 6e-11
    s = O.OracleSim(n, n, n, 0.5e-3, prec)
    if L:
        s.add_cpml((L,) * 6, mode=mode)
    c = n // 2
    smp = X.schedule(lambda t: X.gaussian_modulated(t, f0, 3.5 * tau, tau), 0, steps, s.dt)
    r = s.run_batch(steps, [(c, c, c, 2, True)], smp.astype(np.float64)[None], sample_interval=100)
    tot = r.energy_samples[:, 1] + r.energy_samples[:, 2]
    return tot / tot.max()


def test_cpml_physical_absorbs():
    """After the pulse has left the source, a box with the physical CPML keeps
    only the pulse's static residue (< 1e-4 of the peak energy, monotone), while
    the compat mode -- the reference, which never applies its correction --
    keeps the radiated energy trapped (> 1e-3).  A wrong correction sign makes
    the physical run grow instead."""
    phys = _radiated_energy(40, 8, 1500, O.MODE_PHYSICAL)
    compat = _radiated_energy(40, 8, 1500, O.MODE_COMPAT)
    late = slice(6, None)  # steps >= 600: the pulse (t0 ~ 220 steps) has left
    assert np.all(phys[late] < 1e-4), phys
    assert np.max(phys[late]) - phys[late][0] < 1e-8  # no late growth beyond FP32 rounding
    assert np.min(compat[late]) > 1e-3, compat


# -------------------------------------------------------------------- Mur
def test_mur_coefficient_and_pre_post():
    n = 8
    s = O.OracleSim(n, n, n, 1e-3, 64)
    s.add_mur((False, False, False, False, False, True))
    rng = np.random.default_rng(3)
    f = np.zeros((6, n, n, n))
    f[:3, 1:, 1:, 1:] = rng.standard_normal((3, n - 1, n - 1, n - 1))
    s.set_fields(f)
    ref = O.OracleSim(n, n, n, 1e-3, 64)
    ref.set_fields(f)
    ref.run_batch(1)
    s.run_batch(1)
    coeff = (C0 * s.dt - 1e-3) / (C0 * s.dt + 1e-3)
    assert coeff < 0  # src/extensions/mur_abc.rs:536-546
    R = ref.fields()
    F = s.fields()
    # boundary k = n-1: V = V_shift_old - c V_bnd_old + c V_shift_new for Ex, Ey
    for comp in (0, 1):
        expect = f[comp, :, :, n - 2] - coeff * f[comp, :, :, n - 1] + coeff * R[comp, :, :, n - 2]
        assert np.allclose(F[comp, :, :, n - 1], expect, rtol=1e-12, atol=1e-15)
    assert np.array_equal(F[2], R[2]) and np.array_equal(F[:, :, :, : n - 1], R[:, :, :, : n - 1])


# -------------------------------------------------------------------- ADE
def _uniform_medium_response(kind, params, eps_inf, steps, mode, n=6, prec=64):
    """Ex(t) of a medium filled with a uniform Ex = 1 (index-0 planes included)
    and the ADE on every cell: the curl is exactly zero, so every cell follows
    the material ODE  eps0 eps_inf E' = -J_pol."""
    s = O.OracleSim(n, n, n, 1e-3, prec)
    s.set_material(eps_inf, 1.0, 0.0, 0.0, (0, n, 0, n, 0, n))
    pos = np.array([(i, j, k) for i in range(n) for j in range(n) for k in range(n)], np.uint32)
    s.add_ade(kind, params, pos, mode)
    f = np.zeros((6, n, n, n))
    f[0] = 1.0
    s.set_fields(f)
    p = s.add_field_probe((n // 2, n // 2, n // 2))
    s.run_batch(steps)
    F = s.fields()
    assert np.allclose(F[0], F[0, 1, 1, 1], rtol=1e-12, atol=1e-14) and not np.any(F[3:])
    return s.dt, s.probe_data(p)[:, 0]


def _mean_period(x):
    """Mean zero-crossing period (in samples) of a zero-mean oscillation."""
    sgn = np.signbit(x)
    idx = np.nonzero(sgn[1:] != sgn[:-1])[0]
    # linear interpolation of each crossing
    z = idx + x[idx] / (x[idx] - x[idx + 1])
    return 2.0 * np.mean(np.diff(z))


def test_drude_physical_plasma_oscillation():
    # eps0 eps_inf E' = -J ; J' + g J = eps0 wp^2 E  ->  E'' + g E' + (wp^2/eps_inf) E = 0
    eps_inf = 2.0
    dt0 = O.OracleSim(6, 6, 6, 1e-3).dt
    wp, g = 0.1 / dt0, 0.002 / dt0
    dt, e = _uniform_medium_response(O.ADE_DRUDE, (wp, g), eps_inf, 2000, O.MODE_PHYSICAL)
    t = np.arange(1, 2001) * dt
    wd = math.sqrt(wp * wp / eps_inf - g * g / 4)
    ref = np.exp(-g * t / 2) * (np.cos(wd * t) + (g / (2 * wd)) * np.sin(wd * t))
    assert np.max(np.abs(e[:20] - ref[:20])) < 2e-3
    assert _mean_period(e) * dt == pytest.approx(2 * math.pi / wd, rel=2e-3)
    # envelope decays at g/2 (trapezoidal coupling is non-dissipative for g = 0)
    q = 1800 + int(np.argmax(np.abs(e[1800:])))
    assert abs(e[q]) == pytest.approx(math.exp(-g / 2 * (q + 1) * dt), rel=0.02)


def test_drude_physical_stable_at_large_wp_dt():
    # config-3 Drude: wp = 1e15 rad/s, gamma = 1e13 /s, wp*dt ~ 1e3
    _, e = _uniform_medium_response(O.ADE_DRUDE, (1e15, 1e13), 1.0, 2000, O.MODE_PHYSICAL)
    assert np.all(np.isfinite(e)) and np.max(np.abs(e)) <= 1.0 + 1e-9


def test_lorentz_physical_oscillation():
    # eps0 eps_inf E' = -P'; P'' + g P' + w0^2 P = eps0 de w0^2 E
    #  -> oscillation at W = w0 sqrt(1 + de/eps_inf) about the static E0 eps_inf/(eps_inf+de)
    eps_inf, de = 2.0, 3.0
    dt0 = O.OracleSim(6, 6, 6, 1e-3).dt
    w0 = 0.05 / dt0
    dt, e = _uniform_medium_response(O.ADE_LORENTZ, (w0, 0.0, de), eps_inf, 4000, O.MODE_PHYSICAL)
    W = w0 * math.sqrt(1 + de / eps_inf)
    e_static = eps_inf / (eps_inf + de)
    period = _mean_period(e - e_static) * dt
    assert period == pytest.approx(2 * math.pi / W, rel=2e-3)
    nper = int(len(e) // (period / dt))
    span = int(round(nper * period / dt))
    assert np.mean(e[:span]) == pytest.approx(e_static, abs=5e-3)


def test_debye_physical_relaxation():
    # tau P' + P = eps0 de E, eps0 eps_inf E' = -P'  ->  E -> E0 eps_inf/(eps_inf+de)
    # at rate (1 + de/eps_inf)/tau
    eps_inf, de = 2.0, 4.0
    dt0 = O.OracleSim(6, 6, 6, 1e-3).dt
    tau = 50 * dt0
    dt, e = _uniform_medium_response(O.ADE_DEBYE, (tau, de), eps_inf, 400, O.MODE_PHYSICAL)
    e_inf = eps_inf / (eps_inf + de)
    r = (1 + de / eps_inf) / tau
    assert e[-1] == pytest.approx(e_inf, abs=1e-6)
    slope = (math.log(e[100] - e_inf) - math.log(e[20] - e_inf)) / (80 * dt)
    assert -slope == pytest.approx(r, rel=1e-2)


def test_ade_compat_first_step_known_values():
    # One step of the reference bodies from J = 0 (src/extensions/dispersive.rs:171-196,510-535)
    n = 6
    for kind, prm in ((O.ADE_LORENTZ, (2 * math.pi * 10e9, 1e9, 3.0)), (O.ADE_DRUDE, (1e11, 1e9)),
                      (O.ADE_DEBYE, (1e-11, 2.0))):
        s = O.OracleSim(n, n, n, 1e-3, 64)
        s.add_ade(kind, prm, np.array([[3, 3, 3]], np.uint32), O.MODE_COMPAT)
        ref = O.OracleSim(n, n, n, 1e-3, 64)
        f = np.zeros((6, n, n, n))
        f[2, 3, 3, 3] = 1.0
        s.set_fields(f)
        ref.set_fields(f)
        s.run_batch(1)
        ref.run_batch(1)
        e = ref.fields()[2, 3, 3, 3]
        dt = s.dt
        if kind == O.ADE_LORENTZ:
            w0, g, de = prm
            den = 1 + g * dt / 2
            j = EPS0 * de * w0 ** 2 * dt ** 2 / den * e
            enew = e - ((g * dt / 2 - 1) / den) * j
        elif kind == O.ADE_DRUDE:
            wp, g = prm
            den = 1 + g * dt / 2
            j = EPS0 * wp ** 2 * dt / den * e
            enew = e - dt / EPS0 * j
        else:
            tau, de = prm
            j = 2 * EPS0 * de * dt / (2 * tau + dt) * e
            enew = e - dt / EPS0 * (j / dt)
        assert s.ade_state(0)[2, 0] == pytest.approx(j, rel=1e-12)
        assert s.fields()[2, 3, 3, 3] == pytest.approx(enew, rel=1e-12)
