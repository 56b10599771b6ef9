"""The fused timestep (one pass computing H and E, ping-pong field buffers)
against the FP32 oracle, bit for bit.

Cases: no CPML (the fused pass covers the grid; halo rows and planes are
recomputed), CPML on every face in both modes (the x/y slabs run around the
fused interior, whose halos are then read from the new buffer; every psi
array is compared), CPML on the z faces only, the material table off, k-rows
shorter than a wave, batches that alternate fused and two-half-step schemes,
and probes and a port DFT sampled after fused steps.
"""
import numpy as np
import pytest

from openems_amd import MODE_COMPAT, MODE_PHYSICAL
from tests.parity_util import Scenario, assert_parity, gauss_ref, run_pair, say

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["1", "2"], ids=["lds", "rows"])
def fused_kind(request, monkeypatch):
    """Both fused kernels (OE_FUSED_KIND): rows exchanged through LDS with a
    barrier per plane (1), and each wave recomputing the H of the row below
    (2, the default)."""
    monkeypatch.setenv("OE_FUSED_KIND", request.param)
    return request.param

BOXES = [(3.0, 1.0, 0.01, 0.0, (10, 30, 12, 20, 15, 60)), (6.0, 2.0, 0.0, 0.0, (5, 15, 3, 30, 40, 100))]
PEC = [(20, 22, 8, 28, 70, 72)]


def _check_psi(res):
    cp = res["exts"][0]
    sim = res["sim"]
    for face in range(6):
        for field in ("E", "H"):
            for which in ("pos", "neg"):
                g = cp.psi(face, field, which)
                o = sim.cpml_array(face, field, which)
                if o is None:
                    assert g is None
                    continue
                assert g is not None and np.array_equal(g, o.astype(np.float32)), (face, field, which)


def _assert_fused(eng):
    """The fused kernel ran (the engine's launch counts, timing on)."""
    n = sum(eng.kernel_time(k)[0] for k in ("fused:interior", "fused:grid"))
    say(f"[fused] fused launches: {n}")
    assert n > 0


def _exc():
    return [((20, 18, 64), 2, True, gauss_ref(10e9)), ((22, 17, 60), 0, True, gauss_ref(8e9))]


@pytest.mark.parametrize("table", [1, 0])
def test_fused_no_cpml(table):
    sc = Scenario(f"fused_plain_t{table}", (40, 36, 128), 0.5e-3, 150, materials=BOXES, pecs=PEC,
                  excitations=_exc(), probes=[(30, 18, 64)], fused=1, table=table)
    res = run_pair(sc)
    assert res["eng"].fused_steps() == 1
    _assert_fused(res["eng"])
    assert_parity(res, sc.name)


@pytest.mark.parametrize("mode", [MODE_PHYSICAL, MODE_COMPAT])
def test_fused_cpml_all_faces(mode):
    sc = Scenario(f"fused_cpml_{mode}", (44, 40, 128), 0.5e-3, 200,
                  cpml=dict(layers=(8, 6, 5, 8, 8, 12), mode=mode), materials=BOXES, pecs=PEC,
                  excitations=_exc(), probes=[(30, 18, 64)], fused=1)
    res = run_pair(sc)
    _assert_fused(res["eng"])
    assert_parity(res, sc.name)
    _check_psi(res)


def _many_boxes(n, shape, seed=7):
    """n small dielectric boxes of distinct permittivity: a material table
    of n + 1 E rows (more than a byte indexes once n > 255)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        x0 = int(rng.integers(0, shape[0] - 4))
        y0 = int(rng.integers(0, shape[1] - 4))
        z0 = int(rng.integers(0, shape[2] - 8))
        out.append((1.5 + 0.01 * q, 1.0, 0.0, 0.0, (x0, x0 + 4, y0, y0 + 4, z0, z0 + 8)))
    return out


def test_fused_many_materials_two_byte_index():
    """More than 256 E table rows: the fused pass reads the 16-bit row index
    (the byte-wide copy exists only up to 256 rows)."""
    shape = (44, 40, 128)
    sc = Scenario("fused_many_rows", shape, 0.5e-3, 120, cpml=dict(layers=(8,) * 6, mode=MODE_PHYSICAL),
                  materials=_many_boxes(300, shape), excitations=_exc(), probes=[(30, 18, 64)], fused=1)
    res = run_pair(sc)
    rows = res["eng"].material_table(-1)
    say(f"[fused] table rows (E, H): {rows}")
    assert rows[0] > 256 and rows[1] == 1
    _assert_fused(res["eng"])
    assert_parity(res, sc.name)
    _check_psi(res)


def test_fused_cpml_z_faces_only_and_kappa():
    sc = Scenario("fused_cpml_z", (36, 30, 136), 0.5e-3, 180,
                  cpml=dict(layers=(0, 0, 0, 0, 8, 8), kappa_max=4.0, alpha_max=0.03, mode=MODE_PHYSICAL),
                  materials=BOXES[:1], excitations=_exc(), fused=1)
    res = run_pair(sc)
    _assert_fused(res["eng"])
    assert_parity(res, sc.name)
    _check_psi(res)


def test_fused_short_rows_forced():
    # nz = 44: rows fill 11 of 64 lanes; mode 2 fuses them anyway (z-face
    # layers a multiple of 4, which float4 lanes and so the fused pass need)
    sc = Scenario("fused_short_rows", (30, 26, 44), 0.5e-3, 160,
                  cpml=dict(layers=(6, 6, 6, 6, 8, 8), mode=MODE_PHYSICAL),
                  materials=[(4.0, 1.0, 0.02, 0.0, (8, 20, 6, 18, 10, 30))],
                  excitations=[((15, 13, 22), 2, True, gauss_ref(10e9))], fused=2)
    res = run_pair(sc)
    _assert_fused(res["eng"])
    assert_parity(res, sc.name)
    _check_psi(res)


def test_fused_batches_alternate_with_split():
    """Fused and two-half-step batches in turn: the current buffer and the
    z-face psi buffers carry over correctly."""
    from openems_amd import EngineBatch, ScheduledExcitation
    from oracle import excitation as X
    sc = Scenario("fused_alternate", (40, 36, 128), 0.5e-3, 1, cpml=dict(layers=(8,) * 6, mode=MODE_PHYSICAL),
                  materials=BOXES, excitations=[((20, 18, 64), 2, True, gauss_ref(10e9))])
    res = run_pair(sc)  # one step: engine, oracle and CPML built identically
    eng, sim, exts = res["eng"], res["sim"], res["exts"]
    fn = sc.excitations[0][3]
    start = 1
    eng.set_profiling(1)
    for nb, mode in ((37, 1), (20, 0), (41, 1), (1, 1), (15, 0)):
        eng.fused_steps(mode)
        smp = X.schedule(fn, start, nb, eng.dt)
        eng.run_batch(EngineBatch(num_steps=nb, excitations=[ScheduledExcitation((20, 18, 64), 2, smp, True)],
                                  extensions=exts))
        sim.run_batch(nb, [(20, 18, 64, 2, True)], smp[None].astype(np.float64))
        start += nb
        E, H = eng.read_fields()
        g = np.concatenate([E, H])
        o = sim.fields()
        say(f"[fused] after {start} steps (last batch mode {mode}): equal={np.array_equal(g, o)}")
        assert np.array_equal(g, o)
    _assert_fused(eng)
    _check_psi(res)


def test_fused_port_dft_matches_split():
    """A lumped port (U/I probes + running DFT) sampled after fused steps gives
    the same series and spectra as with the two half-steps."""
    from openems_amd import workloads as W
    out = []
    for mode in (2, 0):  # microstrip rows are short (nz <= 48): force the fused pass
        ms = W.microstrip(np.random.default_rng(11))
        op, eng, cp, port, exc = W.build_microstrip_hip(ms, 600, record_steps=600)
        eng.fused_steps(mode)
        if mode:
            eng.set_profiling(1)
        from openems_amd import EngineBatch
        eng.run_batch(EngineBatch(num_steps=600, excitations=exc, extensions=[cp, port]))
        if mode:
            _assert_fused(eng)
        out.append((port.series(), port.dft_values(), np.concatenate(eng.read_fields())))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
