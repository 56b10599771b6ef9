"""The material table (per-cell row index + coefficient rows in LDS) against
the coefficient arrays and the oracle.

The table is built from the arrays themselves, so every run must be
bit-identical with the table on, with it off, and to the FP32 oracle:
  * material boxes + physical CPML + a soft source (the bench workload's
    ingredients), table on and off;
  * the drop-in path (coefficient arrays handed over whole), including
    grids with more distinct coefficient sets than the table holds, where
    one or both half-steps fall back to the arrays;
  * the row index read as bytes (tables of 2 to 256 rows, the default) and
    as 16-bit words (OE_BYTE_INDEX=0, and every table of 257 rows or more).
"""
import numpy as np
import pytest

from openems_amd import CpmlBoundaries, EngineBatch, HipEngine, MODE_PHYSICAL, Operator, ScheduledExcitation
from oracle import excitation as X
from oracle import oracle as O
from tests.parity_util import rel_l2, say

pytestmark = pytest.mark.gpu

SHAPE = (48, 40, 36)
DELTA = 5e-4


def _boxes(rng, n):
    nx, ny, nz = SHAPE
    out = []
    for q in range(n):
        s = rng.integers(3, 12, size=3)
        o = [rng.integers(0, dim - sz) for dim, sz in zip(SHAPE, s)]
        mu = 2.0 if q % 5 == 0 else 1.0        # every fifth box magnetic: distinct H sets
        out.append((float(rng.uniform(1, 10)), mu, float(rng.uniform(0, 0.05)), 0.0,
                    (o[0], o[0] + s[0], o[1], o[1] + s[1], o[2], o[2] + s[2])))
    return out


def _source(dt, steps):
    t0, tau = X.reference_gaussian_params(20e9, 0.5)
    return X.schedule(lambda t: X.gaussian(t, t0, tau), 0, steps, dt)


def _run(eng, exts, smp, steps, pos=(24, 20, 18)):
    eng.run_batch(EngineBatch(num_steps=steps, excitations=[ScheduledExcitation(pos, 2, smp, True)], extensions=exts))
    E, H = eng.read_fields()
    return np.concatenate([E, H])


def _oracle(sim, smp, steps, pos=(24, 20, 18)):
    sim.run_batch(steps, [(pos[0], pos[1], pos[2], 2, True)], smp[None].astype(np.float64))
    return sim.fields()


def test_boxes_cpml_table_on_off_and_oracle():
    rng = np.random.default_rng(7)
    mats = _boxes(rng, 30)
    steps = 80
    outs = []
    for enable in (1, 0):
        op = Operator.uniform(*SHAPE, DELTA)
        for m in mats:
            op.set_material(*m)
        op.set_pec((10, 14, 10, 12, 8, 9))
        eng = HipEngine(op)
        rows = eng.material_table(enable)
        say(f"[table] enable={enable} rows (E, H) = {rows}")
        if enable:
            # E: vacuum + 30 boxes + PEC at most; H: vacuum + the 6 magnetic
            # boxes at most (a non-magnetic box keeps the vacuum Da, Db)
            assert 2 <= rows[0] <= 32 and 2 <= rows[1] <= 7
        else:
            assert rows == (0, 0)
        cp = eng.cpml(CpmlBoundaries(layers=(8,) * 6, mode=MODE_PHYSICAL))
        outs.append(_run(eng, [cp], _source(op.dt, steps), steps))
    sim = O.OracleSim(*SHAPE, DELTA, 32)
    for m in mats:
        sim.set_material(*m)
    sim.set_pec((10, 14, 10, 12, 8, 9))
    sim.add_cpml((8,) * 6, 3.0, 1e-6, 0.0, 1.0, MODE_PHYSICAL)
    o = _oracle(sim, _source(sim.dt, steps), steps)
    say(f"[table] rel L2 table vs oracle {rel_l2(outs[0], o):.2e}")
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], o)


def _coeff_sets(rng, n_e, n_h, dt):
    """12 coefficient arrays with n_e distinct E sets and n_h distinct H sets
    (per-cell choice among random sets of plausible values)."""
    N = int(np.prod(SHAPE))
    cb0 = dt / (8.854187817e-12 * DELTA)
    db0 = -dt / (1.256637062e-6 * DELTA)
    c = np.empty((12, N), np.float32)
    se = rng.integers(0, n_e, size=N)
    sh = rng.integers(0, n_h, size=N)
    ca_sets = rng.uniform(0.9, 1.0, size=(n_e, 3))
    cb_sets = cb0 * rng.uniform(0.1, 1.0, size=(n_e, 3))
    da_sets = rng.uniform(0.9, 1.0, size=(n_h, 3))
    db_sets = db0 * rng.uniform(0.5, 1.0, size=(n_h, 3))
    for d in range(3):
        c[d] = ca_sets[se, d]
        c[3 + d] = cb_sets[se, d]
        c[6 + d] = da_sets[sh, d]
        c[9 + d] = db_sets[sh, d]
    return c.reshape((12,) + SHAPE)


@pytest.mark.parametrize("n_e,n_h", [(7, 3), (300, 600), (5000, 4), (5000, 5000)])
def test_drop_in_arrays(n_e, n_h):
    rng = np.random.default_rng(n_e + n_h)
    sim = O.OracleSim(*SHAPE, DELTA, 32)
    coeff = _coeff_sets(rng, n_e, n_h, sim.dt)
    sim.set_coefficients(coeff)
    eng = HipEngine.from_arrays(SHAPE, DELTA, sim.dt, coeff)
    rows = eng.material_table()
    say(f"[table] arrays with {n_e} E / {n_h} H sets: rows {rows}")
    assert rows == (n_e if n_e <= 1024 else 0, n_h if n_h <= 1024 else 0)
    assert np.array_equal(eng.coefficients(), coeff)
    steps = 40
    g = _run(eng, [], _source(sim.dt, steps), steps)
    o = _oracle(sim, _source(sim.dt, steps), steps)
    assert np.all(np.isfinite(g))
    assert np.array_equal(g, o)


def test_table_toggle_keeps_state():
    op = Operator.uniform(*SHAPE, DELTA)
    op.set_material(4.0, 2.0, 0.01, 0.0, (5, 20, 5, 20, 5, 20))
    eng = HipEngine(op)
    assert eng.material_table(-1) == (2, 2)
    smp = _source(op.dt, 60)
    eng.run_batch(EngineBatch(num_steps=30, excitations=[ScheduledExcitation((24, 20, 18), 2, smp[:30], True)]))
    assert eng.material_table(0) == (0, 0)
    eng.run_batch(EngineBatch(num_steps=30, excitations=[ScheduledExcitation((24, 20, 18), 2, smp[30:], True)]))
    assert eng.material_table(1) == (2, 2)
    a = np.concatenate(eng.read_fields())
    ref = HipEngine(op)
    b = _run(ref, [], smp, 60)
    assert np.array_equal(a, b)


def test_byte_index_equals_16bit(monkeypatch):
    rng = np.random.default_rng(11)
    mats = _boxes(rng, 30)
    steps = 60
    outs = []
    for byte_index in ("1", "0"):
        monkeypatch.setenv("OE_BYTE_INDEX", byte_index)
        op = Operator.uniform(*SHAPE, DELTA)
        for m in mats:
            op.set_material(*m)
        eng = HipEngine(op)
        rows = eng.material_table(-1)
        say(f"[table] byte index {byte_index}: rows (E, H) = {rows}")
        assert 2 <= rows[0] <= 256 and 2 <= rows[1] <= 256
        cp = eng.cpml(CpmlBoundaries(layers=(8,) * 6, mode=MODE_PHYSICAL))
        outs.append(_run(eng, [cp], _source(op.dt, steps), steps))
    assert np.all(np.isfinite(outs[0]))
    assert np.array_equal(outs[0], outs[1])
