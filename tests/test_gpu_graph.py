"""Steps replayed from captured HIP graphs against steps launched one by one,
and both against the oracle.

The kernels of a step read the step index from a device counter; runs of
steps without host work in between are replayed from one graph per batch.
Every case must give bit-identical fields, probe records, DFT spectra and
series, energy samples and termination steps with graphs on (several run
lengths) and off, and match the oracle.
"""
import numpy as np
import pytest

from openems_amd import ADE_DRUDE, MODE_PHYSICAL, EngineBatch, ScheduledExcitation
from oracle import excitation as X
from tests.parity_util import Scenario, assert_parity, gauss_ref, run_pair, say

pytestmark = pytest.mark.gpu


def _mixed(graph, steps=230):
    pos = np.array([[12, 12, 14], [13, 12, 14], [12, 13, 15]], np.uint32)
    return Scenario(f"graph_{graph}", (28, 26, 32), 0.5e-3, steps, cpml=dict(layers=(6,) * 6, mode=MODE_PHYSICAL),
                    materials=[(3.0, 1.0, 0.01, 0.0, (6, 20, 6, 18, 8, 24))],
                    excitations=[((14, 13, 16), 2, True, gauss_ref(12e9)), ((10, 9, 12), 0, True, gauss_ref(9e9))],
                    mur=[False] * 5 + [True], ades=[(ADE_DRUDE, (1e11, 1e9, 0.0), pos, MODE_PHYSICAL)],
                    probes=[(20, 13, 16)], graph=graph)


@pytest.mark.parametrize("graph", [0, 16, 7])
def test_hooks_with_and_without_graphs(graph):
    sc = _mixed(graph)
    res = run_pair(sc, batches=[37, 193])
    assert res["eng"].graph_steps() == graph
    assert_parity(res, sc.name)


def _port_run(graph, steps=700):
    from openems_amd import workloads as W
    ms = W.microstrip(np.random.default_rng(5))
    op, eng, cp, port, exc = W.build_microstrip_hip(ms, steps, record_steps=steps)
    eng.graph_steps(graph)
    res = eng.run_batch(EngineBatch(num_steps=steps, excitations=exc, extensions=[cp, port], sample_interval=50,
                                    track_peak=True, energy_capacity=100))
    return res, port.series(), port.dft_values(), np.concatenate(eng.read_fields())


def test_port_dft_energy_graph_equals_direct():
    ref = _port_run(0)
    for g in (16, 5):
        got = _port_run(g)
        say(f"[graph] {g} steps per graph: steps {got[0].timesteps_executed}, samples {len(got[0].energy_samples)}")
        assert got[0].timesteps_executed == ref[0].timesteps_executed
        assert np.array_equal(np.asarray(got[0].energy_samples), np.asarray(ref[0].energy_samples))
        for a, b in zip(got[1:], ref[1:]):
            assert np.array_equal(a, b)


def test_decay_termination_graph_equals_direct():
    """Energy-decay termination (checks every 40 steps) stops on the same step."""
    out = []
    for g in (0, 16):
        sc = Scenario(f"graph_decay_{g}", (30, 28, 32), 0.5e-3, 0, cpml=dict(layers=(8,) * 6, mode=MODE_PHYSICAL),
                      excitations=[((15, 14, 16), 2, True, gauss_ref(15e9))], graph=g)
        res = run_pair(sc, batches=[1])
        eng, exts = res["eng"], res["exts"]
        smp = X.schedule(sc.excitations[0][3], 1, 4000, eng.dt)
        r = eng.run_batch(EngineBatch(num_steps=4000, excitations=[ScheduledExcitation((15, 14, 16), 2, smp, True)],
                                      extensions=exts, energy_decay_db=30.0, check_interval=40, track_peak=True,
                                      sample_interval=40))
        out.append((r.timesteps_executed, r.termination_reason, np.concatenate(eng.read_fields())))
        say(f"[graph] {g}: stopped after {r.timesteps_executed} steps, reason {r.termination_reason}")
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("graph", [0, 16])
def test_slab_overlap_equals_serial(graph, monkeypatch):
    """The CPML slab launches on the side stream (forked and joined around
    each interior launch) give the same fields as all launches on one
    stream, eager and replayed from graphs, and both match the oracle."""
    out = []
    for ovl in ("1", "0"):
        monkeypatch.setenv("OE_OVERLAP_SLABS", ovl)
        sc = Scenario(f"overlap_{ovl}_g{graph}", (40, 36, 48), 0.5e-3, 150, cpml=dict(layers=(6,) * 6, mode=MODE_PHYSICAL),
                      materials=[(3.0, 1.0, 0.01, 0.0, (8, 30, 8, 28, 10, 38))],
                      excitations=[((20, 18, 24), 2, True, gauss_ref(12e9)), ((3, 2, 20), 1, True, gauss_ref(9e9))],
                      probes=[(25, 18, 24)], graph=graph)
        res = run_pair(sc, batches=[31, 119])
        assert_parity(res, sc.name)
        out.append(np.concatenate(res["eng"].read_fields()))
    assert np.array_equal(out[0], out[1])
