"""GPU NF2FF: the matrix-core transform against the oracle's float64
restatement of matlab/AnalyzeNF2FF.m, and the recording box end to end.

Tolerances (stated here, reported on every run):
  far-field patterns  rel L2 <= 1e-5 against the float64 oracle (the
                      contraction runs on FP32 matrix cores; phases, the
                      second-axis reduction and P_rad are float64);
  P_rad               rel 1e-12 (float64 throughout);
  surface phasors     rel 1e-11 of the maximum against a float64 DFT of the
                      oracle's signal series (FP64 running DFT);
  FP32-Kahan DFT      far field within rel L2 1e-5 of the FP64-DFT one.
"""
import math

import numpy as np
import pytest

from openems_amd import MODE_PHYSICAL, CpmlBoundaries, EngineBatch, HipEngine, Nf2ff, Operator, ScheduledExcitation
from oracle import excitation as X
from oracle import nf2ff as N
from oracle import oracle as O
from oracle import port as P
from tests.parity_util import say

pytestmark = pytest.mark.gpu

PATTERN_TOL = 1e-5


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _gpu_far_field(faces, f, th, ph, r=1.0, center=(0, 0, 0)):
    nf = Nf2ff(f)
    for fc in faces:
        a1, a2 = (fc.axis + 1) % 3, (fc.axis + 2) % 3
        nf.add_surface(fc.axis, fc.sign, fc.c0, fc.c1, fc.c2, fc.E[[a1, a2]], fc.H[[a1, a2]])
    g = nf.calculate(th, ph, r, center)
    # the split output form of the C-ABI holds the same values
    s = nf.calculate(th, ph, r, center, split=True)
    assert np.array_equal(g.e_theta, s.e_theta) and np.array_equal(g.e_phi, s.e_phi)
    assert g.prad == s.prad and g.dmax == s.dmax
    return g


def _check(name, faces, f, th, ph, r=1.0, center=(0, 0, 0)):
    Et, Ep, prad, dmax = N.far_field(faces, f, th, ph, r=r, center=center)
    g = _gpu_far_field(faces, f, th, ph, r, center)
    et, ep = _rel_l2(g.e_theta, Et), _rel_l2(g.e_phi, Ep)
    tot = math.sqrt((np.linalg.norm(g.e_theta - Et) ** 2 + np.linalg.norm(g.e_phi - Ep) ** 2)
                    / (np.linalg.norm(Et) ** 2 + np.linalg.norm(Ep) ** 2))
    say(f"[nf2ff] {name}: rel L2 E_theta {et:.2e} E_phi {ep:.2e} |E| {tot:.2e}; "
        f"P_rad {g.prad:.6e} vs {prad:.6e}; D_max {g.dmax:.6f} vs {dmax:.6f}; "
        f"gemm {g.gemm_ms:.3f} ms of {g.total_ms:.3f} ms")
    assert tot <= PATTERN_TOL
    assert abs(g.prad - prad) <= 1e-12 * abs(prad)
    assert abs(g.dmax - dmax) <= 1e-5 * abs(dmax)
    return g


def test_hertzian_dipole_surfaces():
    f = 3e9
    lam = N.C0 / f
    faces = N.box_faces([-0.5 * lam] * 3, [0.5 * lam] * 3, (41, 37, 45),
                        lambda X_, Y, Z: N.hertzian_dipole_fields(1e-3, f, X_, Y, Z))
    th = np.linspace(0, math.pi, 37)
    ph = np.linspace(0, 2 * math.pi, 73)[:-1]
    g = _check("hertzian dipole 41x37x45", faces, f, th, ph)
    assert abs(g.dmax - 1.5) < 5e-3


def test_random_phasors_nonuniform_offset_centre():
    rng = np.random.default_rng(11)
    lo, hi = [-0.03, -0.02, -0.025], [0.02, 0.035, 0.03]
    steps = (19, 70, 33)
    axes = [np.sort(rng.uniform(lo[a], hi[a], steps[a])) for a in range(3)]
    for a in range(3):
        axes[a][0], axes[a][-1] = lo[a], hi[a]
    faces = []
    for axis in range(3):
        a1, a2 = (axis + 1) % 3, (axis + 2) % 3
        for sign, c0 in ((-1, lo[axis]), (1, hi[axis])):
            shape = (3, steps[a1], steps[a2])
            E = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
            H = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / N.Z0
            faces.append(N.Face(axis, sign, c0, axes[a1], axes[a2], E, H))
    th = np.linspace(0.05, math.pi - 0.05, 23)
    ph = np.linspace(0, 2 * math.pi, 41)
    _check("random phasors 19x70x33", faces, 8e9, th, ph, r=3.0, center=(0.004, -0.002, 0.001))


def test_angle_chunks_and_wide_faces():
    """z faces 6 x 300 nodes and 181 x 180 irregular angles: every direction
    cosine is distinct, so the GEMM rows are split over several launches;
    the oracle checks every 20th theta row.  A second call with other angles
    and a third with the first ones exercise the cached angle plan."""
    f = 10e9
    lam = N.C0 / f
    faces = N.box_faces([-0.1 * lam, -2 * lam, -0.08 * lam], [0.1 * lam, 2 * lam, 0.08 * lam], (6, 300, 5),
                        lambda X_, Y, Z: N.hertzian_dipole_fields(1e-3, f, X_ + 0.001, Y, Z))
    rng = np.random.default_rng(12)
    th = np.sort(rng.uniform(0, math.pi, 181))
    ph = np.sort(rng.uniform(0, 2 * math.pi, 180))
    nf = Nf2ff(f)
    for fc in faces:
        a1, a2 = (fc.axis + 1) % 3, (fc.axis + 2) % 3
        nf.add_surface(fc.axis, fc.sign, fc.c0, fc.c1, fc.c2, fc.E[[a1, a2]], fc.H[[a1, a2]])
    g = nf.calculate(th, ph)
    sub = slice(0, 181, 20)
    Et, Ep, prad, _ = N.far_field(faces, f, th[sub], ph)
    err = math.sqrt((np.linalg.norm(g.e_theta[sub] - Et) ** 2 + np.linalg.norm(g.e_phi[sub] - Ep) ** 2)
                    / (np.linalg.norm(Et) ** 2 + np.linalg.norm(Ep) ** 2))
    say(f"[nf2ff] 181x180 irregular angles, 6x300x5 box: rel L2 {err:.2e} on every 20th theta; "
        f"gemm {g.gemm_ms:.3f} ms, {g.gemm_flop / 1e9:.1f} GFLOP")
    assert err <= PATTERN_TOL
    assert abs(g.prad - prad) <= 1e-12 * abs(prad)
    th2 = np.linspace(0.2, 2.9, 7)
    g2 = nf.calculate(th2, ph[:50])
    E2t, E2p, _, _ = N.far_field(faces, f, th2, ph[:50])
    assert _rel_l2(g2.e_theta, E2t) <= PATTERN_TOL
    g3 = nf.calculate(th, ph)
    assert np.array_equal(g3.e_theta, g.e_theta) and np.array_equal(g3.e_phi, g.e_phi)


def test_mirror_groups_against_per_angle_reduction(monkeypatch):
    """181 x 360 regular angles with the poles: the second contraction runs
    over mirror groups (theta, pi - theta) x (phi, -phi, pi - phi, pi + phi)
    sharing |kappa_p| and |kappa_q|, and at the poles (720 angles with
    kappa_q = 0) over groups of 16.  It must agree with the per-angle reduction
    (OE_NF2FF_MIRROR=0, which rounds GC -/+ j GS to float32 before the
    float64 sum) to 1e-6 and with the float64 oracle to the pattern
    tolerance."""
    f = 9e9
    lam = N.C0 / f
    faces = N.box_faces([-0.6 * lam, -0.5 * lam, -0.7 * lam], [0.6 * lam, 0.5 * lam, 0.7 * lam], (33, 29, 41),
                        lambda X_, Y, Z: N.hertzian_dipole_fields(1e-3, f, X_ + 0.001, Y - 0.002, Z))
    th = np.linspace(0, math.pi, 181)
    ph = np.linspace(0, 2 * math.pi, 361)[:-1]
    out = {}
    for mirror in ("1", "0"):
        monkeypatch.setenv("OE_NF2FF_MIRROR", mirror)
        out[mirror] = _gpu_far_field(faces, f, th, ph)
    g, g0 = out["1"], out["0"]
    d = math.sqrt((np.linalg.norm(g.e_theta - g0.e_theta) ** 2 + np.linalg.norm(g.e_phi - g0.e_phi) ** 2)
                  / (np.linalg.norm(g0.e_theta) ** 2 + np.linalg.norm(g0.e_phi) ** 2))
    sub = slice(0, 181, 15)
    Et, Ep, prad, _ = N.far_field(faces, f, th[sub], ph)
    err = math.sqrt((np.linalg.norm(g.e_theta[sub] - Et) ** 2 + np.linalg.norm(g.e_phi[sub] - Ep) ** 2)
                    / (np.linalg.norm(Et) ** 2 + np.linalg.norm(Ep) ** 2))
    say(f"[nf2ff] mirror groups, 181x360 angles, 33x29x41 box: rel L2 {d:.2e} vs per-angle, {err:.2e} vs oracle "
        f"(every 15th theta); device {g.total_ms:.3f} ms vs {g0.total_ms:.3f} ms")
    assert d <= 1e-6
    assert err <= PATTERN_TOL
    assert abs(g.prad - prad) <= 1e-12 * abs(prad)
    assert g.prad == g0.prad


def test_side_stream_faces_equal_one_stream(monkeypatch):
    """The x and y faces on a side stream with their own buffers and N/L
    accumulator (the default) give the one-stream transform
    (OE_NF2FF_STREAMS=1) up to the order of the final N/L sum; a second call
    reuses the side stream and buffers and repeats the first bit for bit."""
    f = 6e9
    lam = N.C0 / f
    faces = N.box_faces([-0.7 * lam, -0.6 * lam, -0.8 * lam], [0.7 * lam, 0.6 * lam, 0.8 * lam], (37, 31, 45),
                        lambda X_, Y, Z: N.hertzian_dipole_fields(1e-3, f, X_ - 0.002, Y + 0.001, Z))
    th = np.linspace(0, math.pi, 61)
    ph = np.linspace(0, 2 * math.pi, 121)[:-1]
    out = {}
    for streams in ("1", "2"):
        monkeypatch.setenv("OE_NF2FF_STREAMS", streams)
        out[streams] = _gpu_far_field(faces, f, th, ph)
    g1, g2 = out["1"], out["2"]
    d = math.sqrt((np.linalg.norm(g2.e_theta - g1.e_theta) ** 2 + np.linalg.norm(g2.e_phi - g1.e_phi) ** 2)
                  / (np.linalg.norm(g1.e_theta) ** 2 + np.linalg.norm(g1.e_phi) ** 2))
    Et, Ep, prad, _ = N.far_field(faces, f, th[::6], ph)
    err = math.sqrt((np.linalg.norm(g2.e_theta[::6] - Et) ** 2 + np.linalg.norm(g2.e_phi[::6] - Ep) ** 2)
                    / (np.linalg.norm(Et) ** 2 + np.linalg.norm(Ep) ** 2))
    say(f"[nf2ff] side-stream faces: rel L2 {d:.2e} vs one stream, {err:.2e} vs oracle; "
        f"device {g2.total_ms:.3f} ms vs {g1.total_ms:.3f} ms")
    assert d <= 1e-12
    assert err <= PATTERN_TOL
    assert g2.prad == g1.prad
    again = _gpu_far_field(faces, f, th, ph)
    assert np.array_equal(again.e_theta, g2.e_theta) and np.array_equal(again.e_phi, g2.e_phi)


def _fdtd_dipole(precision, n=48, steps=900, stride=3):
    d = 1e-3
    op = Operator.uniform(n, n, n, d)
    eng = HipEngine(op)
    cp = eng.cpml(CpmlBoundaries(layers=(8,) * 6, mode=MODE_PHYSICAL))
    box = (12, 36, 12, 36, 12, 36)
    freqs = [8e9, 10e9, 12e9]
    nf = eng.nf2ff_box(box, freqs, precision=precision, time_stride=stride)
    t0, tau = X.openems_gaussian_params(5e9)
    smp = X.schedule(lambda t: X.gaussian_modulated(t, 10e9, t0, tau), 0, steps, op.dt)
    src = (24, 24, 24)
    eng.run_batch(EngineBatch(num_steps=steps, excitations=[ScheduledExcitation(src, 2, smp)], extensions=[cp, nf]))
    return eng, nf, freqs, op.dt, smp, box


def test_fdtd_dipole_box_against_oracle():
    n, steps, stride = 48, 900, 3
    say(f"[nf2ff] fdtd dipole {n}^3, {steps} steps, stride {stride}")
    eng, nf, freqs, dt, smp, box = _fdtd_dipole(0, n, steps, stride)
    terms, faces = N.surface_terms((n, n, n), (1e-3,) * 3, box)
    # surface phasors against the oracle's signals
    sim = O.OracleSim(n, n, n, 1e-3)
    sim.add_cpml((8,) * 6, mode=O.MODE_PHYSICAL)
    sim.add_signals(terms)
    sim.run_batch(steps, [(24, 24, 24, 2, True)], smp.astype(np.float64)[None])
    E, H = eng.read_fields()
    assert np.array_equal(np.concatenate([E, H]), sim.fields())
    ser = sim.signal_series(0, len(terms))[::stride]
    st = np.arange(0, steps, stride)
    is_h = np.array([sig[0][0] >= 3 for sig in terms])
    ref = np.empty((len(freqs), len(terms)), np.complex128)
    ref[:, ~is_h] = P.dft(ser[:, ~is_h], (st + 1) * dt, freqs)  # sample spacing stride * dt
    ref[:, is_h] = P.dft(ser[:, is_h], (st + 0.5) * dt, freqs)
    got = nf.dft_values()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say(f"[nf2ff] surface phasors: max |gpu - oracle| / max = {err:.2e} over {len(terms)} signals")
    assert err < 1e-11
    # far field of the GPU's phasors: transform parity and the dipole pattern
    th = np.linspace(0, math.pi, 37)
    ph = np.linspace(0, 2 * math.pi, 37)[:-1]
    for q, f in enumerate(freqs):
        g = nf.far_field(q, th, ph)
        Et, Ep, prad, dmax = N.far_field(N.faces_from_phasors(got[q], (1e-3,) * 3, faces), f, th, ph)
        e = math.sqrt((np.linalg.norm(g.e_theta - Et) ** 2 + np.linalg.norm(g.e_phi - Ep) ** 2)
                      / (np.linalg.norm(Et) ** 2 + np.linalg.norm(Ep) ** 2))
        say(f"[nf2ff] {f / 1e9:.0f} GHz: rel L2 {e:.2e}; P_rad {g.prad:.4e}; D_max {g.dmax:.4f} (oracle {dmax:.4f}, "
            f"Hertzian 1.5)")
        assert e <= PATTERN_TOL
        assert abs(g.prad - prad) <= 1e-12 * abs(prad)
        assert g.prad > 0
        # physics: a single-edge source is a Hertzian dipole.  The pattern is
        # sin(theta) and its sphere-integrated directivity 1.5; D_max divides
        # by the surface Poynting flux, which the absorber's residual
        # reflections in this small box move by a few percent (the oracle's
        # own phasors give the same D_max).
        pat = np.abs(g.e_theta[:, 0])
        assert np.max(np.abs(pat / pat.max() - np.sin(th))) < 0.03
        U = np.abs(g.e_theta) ** 2 + np.abs(g.e_phi) ** 2
        p_int = np.trapezoid(U.mean(axis=1) * 2 * math.pi * np.sin(th), th)
        d_int = 4 * math.pi * U.max() / p_int
        say(f"[nf2ff] {f / 1e9:.0f} GHz: pattern-integrated directivity {d_int:.4f}")
        assert abs(d_int - 1.5) < 0.01
        assert abs(g.dmax - 1.5) < 0.2


def test_fdtd_dipole_fp32_kahan_far_field():
    _, a, freqs, _, _, _ = _fdtd_dipole(0)
    _, b, _, _, _, _ = _fdtd_dipole(1)
    th = np.linspace(0, math.pi, 19)
    ph = np.linspace(0, 2 * math.pi, 19)[:-1]
    for q in range(len(freqs)):
        ga = a.far_field(q, th, ph)
        gb = b.far_field(q, th, ph)
        e = math.sqrt((np.linalg.norm(gb.e_theta - ga.e_theta) ** 2 + np.linalg.norm(gb.e_phi - ga.e_phi) ** 2)
                      / (np.linalg.norm(ga.e_theta) ** 2 + np.linalg.norm(ga.e_phi) ** 2))
        say(f"[nf2ff] fp32-kahan vs fp64 DFT, {freqs[q] / 1e9:.0f} GHz: rel L2 {e:.2e}, "
            f"P_rad rel {abs(gb.prad / ga.prad - 1):.2e}")
        assert e <= PATTERN_TOL
        assert abs(gb.prad / ga.prad - 1) < 1e-5


def test_split_fp16_contraction_against_oracle_and_fp32(monkeypatch):
    """z faces of 100 x 300 nodes and 181 x 180 irregular angles: the large
    contractions run on the FP16 matrix cores with split operands over
    several K steps (K = 128 or 320).  The far field must match the float64
    oracle to the pattern tolerance and the FP32 contraction closely."""
    f = 10e9
    lam = N.C0 / f
    faces = N.box_faces([-1.2 * lam, -2 * lam, -0.1 * lam], [1.2 * lam, 2 * lam, 0.1 * lam], (100, 300, 6),
                        lambda X_, Y, Z: N.hertzian_dipole_fields(1e-3, f, X_ + 0.002, Y - 0.001, Z))
    rng = np.random.default_rng(13)
    th = np.sort(rng.uniform(0, math.pi, 181))
    ph = np.sort(rng.uniform(0, 2 * math.pi, 180))
    out = {}
    for split in ("1", "0"):
        monkeypatch.setenv("OE_NF2FF_SPLIT16", split)
        nf = Nf2ff(f)
        for fc in faces:
            a1, a2 = (fc.axis + 1) % 3, (fc.axis + 2) % 3
            nf.add_surface(fc.axis, fc.sign, fc.c0, fc.c1, fc.c2, fc.E[[a1, a2]], fc.H[[a1, a2]])
        out[split] = nf.calculate(th, ph)
    g, g32 = out["1"], out["0"]
    assert g.gemm_split_ms > 0 and g32.gemm_split_ms == 0
    sub = slice(0, 181, 30)
    Et, Ep, prad, _ = N.far_field(faces, f, th[sub], ph)
    err = math.sqrt((np.linalg.norm(g.e_theta[sub] - Et) ** 2 + np.linalg.norm(g.e_phi[sub] - Ep) ** 2)
                    / (np.linalg.norm(Et) ** 2 + np.linalg.norm(Ep) ** 2))
    d32 = math.sqrt((np.linalg.norm(g.e_theta - g32.e_theta) ** 2 + np.linalg.norm(g.e_phi - g32.e_phi) ** 2)
                    / (np.linalg.norm(g32.e_theta) ** 2 + np.linalg.norm(g32.e_phi) ** 2))
    say(f"[nf2ff] split fp16, 100x300x6 box: rel L2 {err:.2e} vs oracle, {d32:.2e} vs fp32; split gemm "
        f"{g.gemm_split_ms:.3f} ms of {g.gemm_ms:.3f} ms ({g.gemm_split_flop / 1e9:.1f} GFLOP)")
    assert err <= PATTERN_TOL
    assert d32 <= 1e-6
    assert abs(g.prad - prad) <= 1e-12 * abs(prad)


def test_fp32_contraction_after_non_finite_phasors(monkeypatch):
    """A recording box whose phasors were non-finite (a diverged run) and
    then finite again (engine and extension reset, run again) gives the far
    field of a fresh box, bit for bit: the transform's reused buffers keep
    nothing from the earlier call (the FP32 contraction's padding rows are
    cleared per face)."""
    monkeypatch.setenv("OE_NF2FF_SPLIT16", "0")
    shape, steps, freqs = (40, 36, 32), 160, (20e9,)
    box = (10, 30, 9, 27, 8, 24)
    th = np.linspace(0, math.pi, 19)
    ph = np.linspace(0, 2 * math.pi, 25)[:-1]
    t0, tau = X.openems_gaussian_params(30e9)

    def build():
        op = Operator.uniform(*shape, 0.5e-3)
        eng = HipEngine(op)
        cp = eng.cpml(CpmlBoundaries(layers=(6,) * 6, mode=MODE_PHYSICAL))
        nf = eng.nf2ff_box(box, freqs)
        return eng, cp, nf, X.schedule(lambda t: X.gaussian(t, t0, tau), 0, steps, op.dt)

    def run(eng, cp, nf, smp, extra=()):
        eng.run_batch(EngineBatch(num_steps=steps, excitations=[ScheduledExcitation((20, 18, 16), 2, smp, True)],
                                  extensions=[cp, nf, *extra]))

    eng, cp, nf, smp = build()

    def poison(which, step, f):
        if step == 40:
            for c in range(6):
                f[c][...] = np.nan

    run(eng, cp, nf, smp, [eng.host_callback(poison, hooks=("post_update_e",))])
    bad = nf.far_field(0, th, ph)
    assert not np.all(np.isfinite(bad.e_theta))
    eng.reset()
    cp.reset()
    nf.reset()
    run(eng, cp, nf, smp)
    again = nf.far_field(0, th, ph)
    e2, cp2, nf2, _ = build()
    run(e2, cp2, nf2, smp)
    fresh = nf2.far_field(0, th, ph)
    say(f"[nf2ff] after non-finite phasors: finite {bool(np.all(np.isfinite(again.e_theta)))}, "
        f"max |E_theta| {np.abs(fresh.e_theta).max():.3e}")
    assert np.array_equal(nf.dft_values(), nf2.dft_values())
    assert np.array_equal(again.e_theta, fresh.e_theta) and np.array_equal(again.e_phi, fresh.e_phi)
    assert again.prad == fresh.prad and np.abs(fresh.e_theta).max() > 0


def _box_phasors(shape, box, directions, precision, stride, steps=240):
    d = 1e-3
    op = Operator.uniform(*shape, d)
    eng = HipEngine(op)
    cp = eng.cpml(CpmlBoundaries(layers=(6,) * 6, mode=MODE_PHYSICAL))
    nf = eng.nf2ff_box(box, [7e9, 11e9, 13e9], directions=directions, precision=precision, time_stride=stride)
    t0, tau = X.openems_gaussian_params(5e9)
    smp = X.schedule(lambda t: X.gaussian_modulated(t, 10e9, t0, tau), 0, steps, op.dt)
    src = tuple(n // 2 for n in shape)
    eng.run_batch(EngineBatch(num_steps=steps, excitations=[ScheduledExcitation(src, 0, smp),
                                                            ScheduledExcitation(src, 2, smp)],
                              extensions=[cp, nf]))
    return nf.dft_values(), np.concatenate(eng.read_fields())


@pytest.mark.parametrize("box", [(9, 31, 8, 27, 10, 34), (1, 39, 5, 30, 1, 43)], ids=["inner", "to-the-edges"])
@pytest.mark.parametrize("directions", [0, 0b100110])
@pytest.mark.parametrize("precision,stride", [(0, 1), (1, 1), (0, 3)])
def test_surface_dft_equals_term_tables(monkeypatch, directions, precision, stride, box):
    """The recording box's DFT formed from the faces' own indices equals the
    table-driven one bit for bit (same terms, same order, same float64
    sums), on a box of unequal sides in a grid of unequal sides (also one
    reaching the first and last nodes the box may hold, 1 and n - 1 on x
    and z), with every face or a subset (x-hi, y-lo, z-hi), in FP64 and
    FP32-Kahan, sampled every step (graph replay) or every 3rd."""
    shape = (40, 36, 44)
    say(f"[nf2ff] surface DFT vs tables: directions {directions:#b}, precision {precision}, stride {stride}")
    monkeypatch.setenv("OE_SURFACE_DFT", "0")
    ref, f_ref = _box_phasors(shape, box, directions, precision, stride)
    monkeypatch.setenv("OE_SURFACE_DFT", "1")
    got, f_got = _box_phasors(shape, box, directions, precision, stride)
    assert np.array_equal(f_got, f_ref)
    assert got.shape == ref.shape and np.abs(ref).max() > 0
    assert np.array_equal(got, ref)
