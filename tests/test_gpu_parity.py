"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Scenarios are the reference's cross-engine list (src/fdtd/engine_testing.rs:
782-2130) plus the plain-cavity and CPML workloads.  The bar is relative L2
<= 1e-5 over all six field arrays; because the kernels evaluate the same
uncontracted FP32 expressions as the oracle, results are also required to be
bit-identical.
"""
import math

import numpy as np
import pytest

from openems_amd import (ADE_DEBYE, ADE_DRUDE, ADE_LORENTZ, MODE_COMPAT, MODE_PHYSICAL, EngineBatch, HipEngine,
                         Operator, ScheduledExcitation)
from openems_amd import workloads as W
from oracle import excitation as X
from oracle import oracle as O
from tests.parity_util import Scenario, assert_parity, gauss_ref, rel_l2, run_pair, say

pytestmark = pytest.mark.gpu


def dirac(t):
    return X.dirac(t)


REFERENCE_SCENARIOS = [
    Scenario("pec_cavity_energy_conservation", (20, 20, 20), 0.5e-3, 200, excitations=[((10, 10, 10), 2, True, dirac)]),
    Scenario("free_propagation_no_reflections", (40, 40, 60), 0.5e-3, 150,
             excitations=[((20, 20, 10), 2, True, gauss_ref(5e9))]),
    Scenario("dielectric_slab_interface", (30, 30, 60), 0.3e-3, 200,
             materials=[(4.0, 1.0, 0.0, 0.0, (0, 30, 0, 30, 25, 35))],
             excitations=[((15, 15, 10), 2, True, gauss_ref(3e9))]),
    Scenario("resonant_cavity_mode", (20, 20, 40), 0.5e-3, 300,
             excitations=[((10, 10, 20), 2, True, lambda t: X.sinusoidal(t, 3e9))]),
    Scenario("small_grid_stability", (10, 10, 10), 1e-3, 100, excitations=[((5, 5, 5), 0, True, gauss_ref(2e9))]),
    Scenario("multi_direction_excitation", (25, 25, 25), 0.5e-3, 150,
             excitations=[((12, 12, 12), 0, True, gauss_ref(3e9)), ((13, 13, 13), 1, True, gauss_ref(3e9)),
                          ((14, 14, 14), 2, True, gauss_ref(3e9))]),
    Scenario("hard_soft_source_behavior", (30, 30, 30), 0.5e-3, 100,
             excitations=[((15, 15, 15), 2, True, gauss_ref(3e9)), ((20, 20, 20), 0, False, gauss_ref(3e9))]),
    Scenario("long_simulation_stability", (15, 15, 15), 1e-3, 1000, excitations=[((7, 7, 7), 1, True, dirac)]),
    Scenario("zero_excitation_stays_zero", (20, 20, 20), 0.5e-3, 100),
    Scenario("high_permittivity_material", (30, 30, 60), 0.3e-3, 300,
             materials=[(100.0, 1.0, 0.0, 0.0, (0, 30, 0, 30, 25, 35))],
             excitations=[((15, 15, 10), 2, True, gauss_ref(1e9))]),
    Scenario("extreme_permittivity_material", (25, 25, 50), 0.4e-3, 400,
             materials=[(1000.0, 1.0, 0.0, 0.0, (0, 25, 0, 25, 20, 30))],
             excitations=[((12, 12, 8), 2, True, gauss_ref(0.5e9))]),
    Scenario("lossy_dielectric_material", (30, 30, 60), 0.3e-3, 300,
             materials=[(4.0, 1.0, 0.1, 0.0, (0, 30, 0, 30, 20, 40))],
             excitations=[((15, 15, 10), 2, True, gauss_ref(3e9))]),
    Scenario("magnetic_material", (25, 25, 50), 0.4e-3, 300,
             materials=[(1.0, 10.0, 0.0, 0.0, (0, 25, 0, 25, 18, 32))],
             excitations=[((12, 12, 8), 2, True, gauss_ref(2e9))]),
    Scenario("very_long_simulation_stability", (12, 12, 12), 1e-3, 10000, excitations=[((6, 6, 6), 1, True, dirac)]),
    Scenario("multi_layer_dielectric_stack", (25, 25, 80), 0.25e-3, 400,
             materials=[(2.0, 1.0, 0.0, 0.0, (0, 25, 0, 25, 20, 30)), (4.0, 1.0, 0.0, 0.0, (0, 25, 0, 25, 30, 40)),
                        (9.0, 1.0, 0.0, 0.0, (0, 25, 0, 25, 40, 50)), (16.0, 1.0, 0.0, 0.0, (0, 25, 0, 25, 50, 60))],
             excitations=[((12, 12, 8), 2, True, gauss_ref(3e9))]),
    Scenario("mixed_lossy_lossless_materials", (30, 30, 60), 0.3e-3, 300,
             materials=[(4.0, 1.0, 0.0, 0.0, (0, 30, 0, 15, 20, 40)), (4.0, 1.0, 0.05, 0.0, (0, 30, 15, 30, 20, 40))],
             excitations=[((15, 15, 10), 2, True, gauss_ref(3e9))]),
]


@pytest.mark.parametrize("sc", REFERENCE_SCENARIOS, ids=lambda s: s.name)
def test_reference_scenario_parity(sc):
    res = run_pair(sc)
    assert_parity(res, sc.name)
    if sc.name == "zero_excitation_stays_zero":
        assert not np.any(res["gpu"])


def test_plain_cavity_config_parity_and_fp64_error():
    """100^3 vacuum, openEMS Gaussian (fc = 15 GHz) soft Ez at the centre, 1000
    steps, probe 20 cells off-centre: bit-identical to the FP32 oracle; the
    error against the FP64 oracle is reported."""
    wl = W.plain_cavity()
    t0, tau = wl.gaussian_t0_tau()
    sc = Scenario(wl.name, wl.shape, wl.delta, wl.steps,
                  excitations=[(wl.source, wl.source_dir, True, lambda t: X.gaussian(t, t0, tau))], probes=[wl.probe])
    res = run_pair(sc)
    assert_parity(res, sc.name)
    res64 = run_pair(sc, prec=64)
    err64 = rel_l2(res["gpu"], res64["oracle"])
    perr = rel_l2(res["gpu_probes"][0], res64["oracle_probes"][0])
    say(f"[parity] plain_cavity: rel_l2 vs FP64 oracle fields={err64:.3e} probe={perr:.3e}")
    assert err64 < 1e-3
    probe = res["gpu_probes"][0]
    assert probe.shape == (1000, 6) and np.max(np.abs(probe[:, 2])) > 0


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_PHYSICAL], ids=["compat", "physical"])
def test_cpml_parity_fields_and_psi(mode):
    sc = Scenario(f"cpml_{mode}", (40, 36, 44), 0.5e-3, 300, cpml=dict(layers=(8, 6, 5, 8, 8, 7), mode=mode),
                  materials=[(3.0, 1.0, 0.01, 0.0, (10, 30, 12, 20, 15, 30))],
                  excitations=[((20, 18, 22), 2, True, gauss_ref(10e9)), ((22, 17, 20), 0, True, gauss_ref(8e9))],
                  probes=[(30, 18, 22)])
    res = run_pair(sc)
    assert_parity(res, sc.name)
    cp = res["exts"][0]
    sim = res["sim"]
    for face in range(6):
        for field in ("E", "H"):
            for which in ("pos", "neg"):
                g = cp.psi(face, field, which)
                o = sim.cpml_array(face, field, which)
                assert g is not None and o is not None
                assert np.array_equal(g, o.astype(np.float32)), (face, field, which)


def test_cpml_kappa_alpha_parity():
    sc = Scenario("cpml_kappa_alpha", (32, 32, 32), 0.5e-3, 200,
                  cpml=dict(layers=(8,) * 6, kappa_max=5.0, alpha_max=0.05, mode=MODE_PHYSICAL),
                  excitations=[((16, 16, 16), 2, True, gauss_ref(10e9))])
    assert_parity(run_pair(sc), sc.name)


def test_cpml_boxes_reduced_parity():
    wl = W.cpml_boxes(n=72, steps=300, count=20)
    t0, tau = wl.gaussian_t0_tau()
    for mode in (MODE_PHYSICAL, MODE_COMPAT):
        sc = Scenario(f"cpml_boxes72_{mode}", wl.shape, wl.delta, wl.steps, materials=wl.boxes,
                      cpml=dict(layers=(8,) * 6, mode=mode),
                      excitations=[(wl.source, 2, True, lambda t: X.gaussian(t, t0, tau))], probes=[wl.probe])
        assert_parity(run_pair(sc), sc.name)


def test_cpml_boxes_full_size_short_run():
    """The 256^3 CPML workload at full size for 60 steps (the oracle finishes
    in seconds)."""
    wl = W.cpml_boxes(n=256, steps=60)
    t0, tau = wl.gaussian_t0_tau()
    sc = Scenario("cpml_boxes256", wl.shape, wl.delta, wl.steps, materials=wl.boxes,
                  cpml=dict(layers=(8,) * 6, mode=MODE_PHYSICAL),
                  excitations=[(wl.source, 2, True, lambda t: X.gaussian(t, t0, tau))], probes=[wl.probe])
    assert_parity(run_pair(sc), sc.name)


def test_mur_parity():
    sc = Scenario("mur_zmax", (24, 28, 32), 0.5e-3, 250, mur=(False, False, False, False, False, True),
                  excitations=[((12, 14, 16), 2, True, gauss_ref(10e9)), ((12, 14, 16), 0, True, gauss_ref(12e9))])
    assert_parity(run_pair(sc), sc.name)
    sc2 = Scenario("mur_all_faces", (20, 22, 24), 0.5e-3, 150, mur=(True,) * 6,
                   excitations=[((10, 11, 12), 1, True, gauss_ref(10e9))])
    assert_parity(run_pair(sc2), sc2.name)


def _sphere(center, r, shape):
    c = np.array(center)
    g = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1)
    m = np.sum((g - c) ** 2, -1) <= r * r
    return np.argwhere(m).astype(np.uint32)


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_PHYSICAL], ids=["compat", "physical"])
def test_ade_parity(mode):
    shape = (32, 32, 32)
    ades = [
        (ADE_DRUDE, (1e15, 1e13, 0.0), _sphere((10, 10, 16), 4, shape), mode),
        (ADE_LORENTZ, (2 * math.pi * 10e9, 1e9, 3.0), _sphere((22, 20, 16), 5, shape), mode),
        (ADE_DEBYE, (1e-11, 4.0, 0.0), _sphere((16, 24, 10), 3, shape), mode),
    ]
    if mode == MODE_COMPAT:  # the reference's explicit Drude diverges at wp*dt ~ 1e3
        ades[0] = (ADE_DRUDE, (1e11, 1e9, 0.0), ades[0][2], mode)
    sc = Scenario(f"ade_{mode}", shape, 0.5e-3, 200, ades=ades,
                  materials=[(2.0, 1.0, 0.0, 0.0, (17, 28, 15, 26, 11, 22))],
                  excitations=[((16, 16, 16), 2, True, gauss_ref(10e9))], probes=[(10, 10, 16)])
    res = run_pair(sc)
    assert_parity(res, sc.name)
    for g, o in zip(res["gpu_ade"], res["oracle_ade"]):
        assert np.array_equal(g, o.astype(np.float32))


@pytest.mark.parametrize("tiles", [dict(bx=16, ci=3), dict(bx=64, ci=40), dict(vec=1, bx=32, ci=5),
                                   dict(vec=1, bx=256, ci=7)], ids=str)
def test_tile_shapes_give_identical_results(tiles):
    sc = Scenario("tiles", (44, 36, 40), 0.5e-3, 120, cpml=dict(layers=(6,) * 6, mode=MODE_PHYSICAL),
                  materials=[(5.0, 2.0, 0.02, 0.0, (5, 30, 3, 20, 10, 35))],
                  excitations=[((22, 18, 20), 1, True, gauss_ref(10e9))], tiles=tiles)
    assert_parity(run_pair(sc), f"tiles {tiles}")


def test_batch_consistency():
    # BatchConsistencyScenario (src/fdtd/engine_testing.rs:1233-1276): 2 x 100 == 1 x 200
    sc = Scenario("batch_consistency", (15, 15, 15), 0.8e-3, 200, excitations=[((7, 7, 7), 2, True, gauss_ref(4e9))],
                  cpml=dict(layers=(3,) * 6, mode=MODE_PHYSICAL))
    one = run_pair(sc)
    two = run_pair(sc, batches=[100, 100])
    assert np.array_equal(one["gpu"], two["gpu"])
    assert_parity(two, "batch_consistency")


def test_energy_samples_and_decay_termination():
    n = 16
    op = Operator.uniform(n, n, n, 1e-3)
    op.set_material(1.0, 1.0, 5.0, 0.0, (0, n, 0, n, 0, n))
    sim = O.OracleSim(n, n, n, 1e-3)
    sim.set_material(1.0, 1.0, 5.0, 0.0, (0, n, 0, n, 0, n))
    eng = HipEngine(op)
    smp = np.zeros(1, np.float32)
    smp[0] = 1.0
    res = eng.run_batch(EngineBatch(num_steps=100000, excitations=[ScheduledExcitation((8, 8, 8), 2, smp)],
                                    check_interval=50, energy_decay_db=30.0, sample_interval=1, track_peak=True))
    r2 = sim.run_batch(100000, [(8, 8, 8, 2, True)], smp.astype(np.float64)[None], check_interval=50,
                       energy_decay_db=30.0, sample_interval=1, track_peak=True)
    assert res.termination_reason == 1 == r2.termination_reason
    assert res.timesteps_executed == r2.timesteps_executed
    assert res.final_decay_db == pytest.approx(r2.final_decay_db, rel=1e-12)
    assert np.array_equal(res.energy_samples[:, 0], r2.energy_samples[:, 0])
    assert np.allclose(res.energy_samples[:, 1:], r2.energy_samples[:, 1:], rtol=1e-12, atol=0)


def test_read_write_reset_roundtrip():
    op = Operator.uniform(9, 10, 11, 1e-3)
    eng = HipEngine(op)
    rng = np.random.default_rng(5)
    E = rng.standard_normal((3, 9, 10, 11)).astype(np.float32)
    H = rng.standard_normal((3, 9, 10, 11)).astype(np.float32)
    eng.write_fields(E, H)
    e2, h2 = eng.read_fields()
    assert np.array_equal(e2, E) and np.array_equal(h2, H)
    ee, he = eng.energy()
    assert ee == pytest.approx(np.sum(E.astype(np.float64) ** 2), rel=1e-12)
    assert he == pytest.approx(np.sum(H.astype(np.float64) ** 2), rel=1e-12)
    eng.run_batch(EngineBatch(num_steps=3))
    assert eng.timestep == 3
    eng.reset()
    e3, h3 = eng.read_fields()
    assert eng.timestep == 0 and not np.any(e3) and not np.any(h3)
    # coefficients read back equal the oracle's
    sim = O.OracleSim(9, 10, 11, 1e-3)
    assert np.array_equal(eng.coefficients(), sim.coefficients())


def test_coefficients_with_materials_match_oracle():
    wl = W.cpml_boxes(n=64, steps=1, count=30, seed=7)
    op = Operator.uniform(*wl.shape, wl.delta)
    sim = O.OracleSim(*wl.shape, wl.delta)
    W.apply_materials(op, wl)
    W.apply_materials(sim, wl)
    op.set_pec((30, 34, 30, 34, 0, 64))
    sim.set_pec((30, 34, 30, 34, 0, 64))
    eng = HipEngine(op)
    assert np.array_equal(eng.coefficients(), sim.coefficients())


def test_extension_outliving_engine_is_detached():
    """Regression: destroying the engine before its extensions used to leave
    them pointing at freed memory; they are now detached and report an error."""
    from openems_amd import CpmlBoundaries, OEError

    op = Operator.uniform(16, 16, 16, 1e-3)
    eng = HipEngine(op)
    cp = eng.cpml(CpmlBoundaries(layers=(4,) * 6))
    pr = eng.field_probe((8, 8, 8))
    eng.run_batch(EngineBatch(num_steps=5, extensions=[cp, pr]))
    assert pr.data().shape == (5, 6)
    eng.close()
    with pytest.raises(OEError):
        pr.data()
    with pytest.raises(OEError):
        cp.reset()
    del cp, pr  # oe_extension_destroy on detached handles


@pytest.mark.parametrize("overlap", [False, True], ids=["disjoint", "overlapping"])
def test_ade_merged_launches(overlap, monkeypatch):
    """Two dispersive extensions share one pre and one post launch when their
    cells are disjoint; with overlapping cells the engine keeps one launch per
    extension in list order.  Both match the oracle, merged or not."""
    shape = (30, 30, 30)
    c2 = (17, 15, 15) if overlap else (21, 21, 15)
    ades = [(ADE_DRUDE, (1e11, 1e9, 0.0), _sphere((12, 12, 15), 4, shape), MODE_PHYSICAL),
            (ADE_LORENTZ, (2 * math.pi * 10e9, 1e9, 3.0), _sphere(c2, 5, shape), MODE_PHYSICAL)]
    out = []
    for merge in ("1", "0"):
        monkeypatch.setenv("OE_MERGE_ADE", merge)
        sc = Scenario(f"ade_merge{merge}_{'ov' if overlap else 'dj'}", shape, 0.5e-3, 160, ades=ades,
                      excitations=[((15, 15, 15), 2, True, gauss_ref(10e9))], probes=[(12, 12, 15)])
        res = run_pair(sc)
        assert_parity(res, sc.name)
        for g, o in zip(res["gpu_ade"], res["oracle_ade"]):
            assert np.array_equal(g, o.astype(np.float32))
        out.append(np.concatenate(res["eng"].read_fields()))
    assert np.array_equal(out[0], out[1])
