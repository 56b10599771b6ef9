"""The multi-structure sweep against single-engine runs of the same
structures (and one structure against the oracle).

A sweep must give every structure exactly what its own engine gives with
the matching termination configuration (decay test every check_interval
steps, energy sampled at the same steps, whole-run peak): the same number of
steps, the same reason, bit-identical fields and port signals, identical
port spectra.
"""
import numpy as np
import pytest

from openems_amd import EngineBatch, Sweep
from openems_amd import workloads as W
from oracle import oracle as O
from oracle import port as P
from tests.parity_util import say

pytestmark = pytest.mark.gpu

MAX_STEPS = 700
CHECK = 50
DECAY = 20.0


def _structures():
    rng = np.random.default_rng(21)
    out = []
    for shape in ((44, 40, 32), (60, 48, 36), (52, 44, 28), (40, 52, 40)):
        ms = W.microstrip(rng)
        # shrink to test size, keeping the generator's geometry rules
        w = min(ms.width, 4)
        x0, y0 = 16, 16
        segs = [(x0, shape[0] - 16, y0, y0 + w), (shape[0] - 16 - w, shape[0] - 16, y0, shape[1] - 16)]
        out.append(W.Microstrip(shape, ms.eps_r, min(ms.h, 5), w, segs, (x0 + w // 2, y0 + w // 2)))
    # one grid whose rows do not hold whole float4 lanes (scalar lanes)
    out.append(W.Microstrip((42, 38, 30), 4.4, 3, 3, [(14, 28, 14, 17)], (15, 15)))
    return out


def _single(ms):
    op, eng, cp, port, exc = W.build_microstrip_hip(ms, MAX_STEPS, record_steps=MAX_STEPS)
    res = eng.run_batch(EngineBatch(max_timesteps=MAX_STEPS, excitations=exc, extensions=[cp, port],
                                    energy_decay_db=DECAY, check_interval=CHECK, sample_interval=CHECK,
                                    track_peak=True, persistent_peak=True))
    return eng, port, res


def _sweep(structs, use_graph, autotune):
    sw = Sweep()
    keep = []
    for ms in structs:
        op, eng, cp, port, exc = W.build_microstrip_hip(ms, MAX_STEPS, record_steps=MAX_STEPS)
        sw.add(eng, cp, port, exc)
        keep.append((eng, port))
    res = sw.run(MAX_STEPS, energy_decay_db=DECAY, check_interval=CHECK, use_graph=use_graph, autotune=autotune)
    return sw, keep, res


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("overlap", ["0", "1", "2"])
def test_sweep_matches_single_engines(use_graph, overlap, monkeypatch):
    """overlap 1: the CPML slab launches of each half-step run on a second
    stream beside the interior launch; 2: on three (OE_SWEEP_OVERLAP)."""
    monkeypatch.setenv("OE_SWEEP_OVERLAP", overlap)
    structs = _structures()
    sw, keep, res = _sweep(structs, use_graph, autotune=use_graph)
    say(f"[sweep] graph={use_graph} overlap={overlap} tuning {sw.tuning()}")
    for q, ms in enumerate(structs):
        eng1, port1, r1 = _single(ms)
        eng2, port2 = keep[q]
        say(f"[sweep] {ms.shape}: sweep {res[q].timesteps_executed} steps (reason {res[q].termination_reason}), "
            f"single {r1.timesteps_executed} (reason {r1.termination_reason})")
        assert res[q].timesteps_executed == r1.timesteps_executed
        assert res[q].termination_reason == r1.termination_reason
        assert eng2.timestep == eng1.timestep
        E1, H1 = eng1.read_fields()
        E2, H2 = eng2.read_fields()
        assert np.array_equal(E1, E2) and np.array_equal(H1, H2)
        assert np.array_equal(port1.series(), port2.series())
        assert np.array_equal(port1.dft_values(), port2.dft_values())
        assert np.array_equal(port1.s11(50.0), port2.s11(50.0))
    assert any(r.termination_reason == 1 for r in res), "no structure reached the decay criterion"


def test_sweep_structure_against_oracle():
    ms = _structures()[1]
    steps = 300
    sw = Sweep()
    op, eng, cp, port, exc = W.build_microstrip_hip(ms, steps, record_steps=steps)
    sw.add(eng, cp, port, exc)
    sw.run(steps, check_interval=CHECK)
    sim = O.OracleSim(*ms.shape, ms.delta)
    ms.apply(sim)
    sim.add_cpml((ms.layers,) * 6, mode=O.MODE_PHYSICAL)
    sim.add_signals(P.lumped_port_terms(ms.shape, (ms.delta,) * 3, 2, ms.port_region))
    sim.run_batch(steps, [(e.position[0], e.position[1], e.position[2], e.direction, True) for e in exc],
                  np.array([e.samples.astype(np.float64) for e in exc]))
    E, H = eng.read_fields()
    assert np.array_equal(np.concatenate([E, H]), sim.fields())
    assert np.array_equal(port.series(), sim.signal_series(0, 2))


def _shifted(exc, k):
    from openems_amd import ScheduledExcitation
    return [ScheduledExcitation(e.position, e.direction, e.samples[k:], e.soft_source) for e in exc]


@pytest.mark.parametrize("first", [173, 200])
def test_sweep_continued_equals_one_run(first):
    """A run of `first` steps, then a second sweep over the same engines
    (excitations shifted by the steps already run) to the end: decay checks
    fall on the same timesteps and the peak carries over, so steps, reason,
    fields, port series and spectra equal those of one run."""
    structs = _structures()
    _, keep1, res1 = _sweep(structs, True, autotune=False)
    sw = Sweep()
    built = []
    for ms in structs:
        op, eng, cp, port, exc = W.build_microstrip_hip(ms, MAX_STEPS, record_steps=MAX_STEPS)
        sw.add(eng, cp, port, exc)
        built.append((eng, cp, port, exc))
    ra = sw.run(first, energy_decay_db=DECAY, check_interval=CHECK, autotune=False)
    cont = [q for q in range(len(structs)) if ra[q].termination_reason == 0]
    sw2 = Sweep()
    for q in cont:
        eng, cp, port, exc = built[q]
        sw2.add(eng, cp, port, _shifted(exc, first))
    rb = sw2.run(MAX_STEPS - first, energy_decay_db=DECAY, check_interval=CHECK, autotune=False)
    total = [r.timesteps_executed for r in ra]
    reason = [r.termination_reason for r in ra]
    for k, q in enumerate(cont):
        total[q] += rb[k].timesteps_executed
        reason[q] = rb[k].termination_reason
    for q in range(len(structs)):
        eng1, port1 = keep1[q]
        eng2, _, port2, _ = built[q]
        say(f"[sweep] continued at {first}: {structs[q].shape} {total[q]} steps (reason {reason[q]}), "
            f"one run {res1[q].timesteps_executed} (reason {res1[q].termination_reason})")
        assert total[q] == res1[q].timesteps_executed and reason[q] == res1[q].termination_reason
        assert eng2.timestep == eng1.timestep
        assert np.array_equal(np.concatenate(eng1.read_fields()), np.concatenate(eng2.read_fields()))
        assert np.array_equal(port1.series(), port2.series())
        assert np.array_equal(port1.dft_values(), port2.dft_values())
    assert cont, "every structure stopped in the first run"


def test_sweep_rejects_mixed_timesteps():
    structs = _structures()[:2]
    built = [W.build_microstrip_hip(ms, 50) for ms in structs]
    sw = Sweep()
    _, eng, cp, port, exc = built[0]
    sw.add(eng, cp, port, exc)
    sw.run(3, check_interval=CHECK, autotune=False)
    sw2 = Sweep()
    for _, eng, cp, port, exc in built:
        sw2.add(eng, cp, port, exc)
    with pytest.raises(Exception, match="one timestep"):
        sw2.run(5, check_interval=CHECK, autotune=False)


def test_sweep_sources_on_repeated_cells_keep_list_order():
    """Sources of one structure that share a cell (a soft and then a hard
    source on one port edge) are applied in list order by the sweep's point
    kernel, as a single engine applies them; the other structure's distinct
    cells take the parallel path."""
    from openems_amd import ScheduledExcitation
    structs = _structures()[:2]
    steps = 300

    def sources(exc):
        e0, e1 = exc[0], exc[1]
        return exc + [ScheduledExcitation(e0.position, e0.direction, 0.5 * e0.samples, True),
                      ScheduledExcitation(e1.position, e1.direction, 0.25 * e1.samples, False)]

    sw = Sweep()
    keep = []
    for q, ms in enumerate(structs):
        op, eng, cp, port, exc = W.build_microstrip_hip(ms, steps, record_steps=steps)
        sw.add(eng, cp, port, sources(exc) if q == 0 else exc)
        keep.append((eng, port))
    sw.run(steps, check_interval=CHECK, autotune=False)
    for q, ms in enumerate(structs):
        op, eng1, cp1, port1, exc1 = W.build_microstrip_hip(ms, steps, record_steps=steps)
        eng1.run_batch(EngineBatch(num_steps=steps, excitations=sources(exc1) if q == 0 else exc1,
                                   extensions=[cp1, port1]))
        eng2, port2 = keep[q]
        assert np.array_equal(np.concatenate(eng1.read_fields()), np.concatenate(eng2.read_fields()))
        assert np.array_equal(port1.series(), port2.series())
