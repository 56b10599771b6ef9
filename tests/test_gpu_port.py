"""GPU parity of the lumped port, V/I signals and the running DFT.

The port's U/I series are float64 sums of the same float32 field values in
the same term order as the oracle, so they must be bit-identical; the FP64
running DFT is compared with a direct float64 DFT of the oracle's series, and
the FP32-Kahan running DFT with the FP64 one.
"""
import math

import numpy as np
import pytest

from openems_amd import MODE_PHYSICAL, CpmlBoundaries, EngineBatch, HipEngine, Operator
from openems_amd.engine import port_excitations
from oracle import excitation as X
from oracle import oracle as O
from oracle import port as P
from tests.parity_util import say

pytestmark = pytest.mark.gpu


def _port_case(precision, steps=600, n=40, R=50.0):
    d = 0.5e-3
    reg = (20, 21, 20, 21, 18, 22)
    op = Operator.uniform(n, n, n, d)
    sim = O.OracleSim(n, n, n, d)
    for s in (op, sim):
        s.set_material(3.0, 1.0, 0.0, 0.0, (0, n, 0, n, 0, 18))
    op.add_lumped_element(2, reg, R)
    sim.add_lumped(2, reg, R)
    eng = HipEngine(op)
    cp = eng.cpml(CpmlBoundaries(layers=(8,) * 6, mode=MODE_PHYSICAL))
    sim.add_cpml((8,) * 6, mode=O.MODE_PHYSICAL)
    freqs = np.linspace(1e9, 20e9, 64)
    port = eng.lumped_port(2, reg, freqs, precision=precision, record_steps=steps)
    sim.add_signals(P.lumped_port_terms((n, n, n), sim.delta, 2, reg))
    t0, tau = X.openems_gaussian_params(15e9)
    smp = X.schedule(lambda t: X.gaussian(t, t0, tau), 0, steps, op.dt)
    gexc = port_excitations(2, reg, smp)
    eng.run_batch(EngineBatch(num_steps=steps, excitations=gexc, extensions=[cp, port]))
    oexc = [(e.position[0], e.position[1], e.position[2], e.direction, True) for e in gexc]
    sim.run_batch(steps, oexc, np.array([e.samples.astype(np.float64) for e in gexc]))
    return eng, port, sim, freqs, op.dt


def test_port_series_dft_and_s11_fp64():
    say("[port] fp64 dft, 40^3, 600 steps")
    eng, port, sim, freqs, dt = _port_case(0)
    g = port.series()
    o = sim.signal_series(0, 2)
    assert g.shape == o.shape == (600, 2)
    assert np.array_equal(g, o)
    E, H = eng.read_fields()
    assert np.array_equal(np.concatenate([E, H]), sim.fields())
    steps = np.arange(600)
    tE = (steps + 1) * dt
    tH = (steps + 0.5) * dt
    U = P.dft(o[:, :1], tE, freqs)[:, 0]
    I = P.dft(o[:, 1:], tH, freqs)[:, 0]
    got = port.dft_values()
    assert np.allclose(got[:, 0], U, rtol=1e-11, atol=1e-11 * np.abs(U).max())
    assert np.allclose(got[:, 1], I, rtol=1e-11, atol=1e-11 * np.abs(I).max())
    s11 = port.s11(50.0)
    ref = P.calc_port(U, I, 50.0)["s11"]
    assert np.allclose(s11, ref, rtol=1e-9, atol=1e-9)
    assert np.all(np.abs(s11) < 1.5)


def test_port_dft_fp32_kahan_against_fp64():
    say("[port] fp32-kahan vs fp64")
    _, p64, _, freqs, _ = _port_case(0, steps=2000)
    _, p32, _, _, _ = _port_case(1, steps=2000)
    a = p64.dft_values()
    b = p32.dft_values()
    for m in range(2):
        err = np.abs(b[:, m] - a[:, m]) / np.abs(a[:, m]).max()
        say(f"[port] signal {m}: max |fp32k - fp64| / max|fp64| = {err.max():.2e}")
        assert err.max() < 1e-6
    s64 = p64.s11(50.0)
    s32 = p32.s11(50.0)
    assert np.max(np.abs(s32 - s64)) < 1e-4


def test_generic_dft_surface_signals():
    """Single-entry E and H signals (the NF2FF surface form) with FP64 sums."""
    n, steps = 24, 300
    op = Operator.uniform(n, n, n, 1e-3)
    sim = O.OracleSim(n, n, n, 1e-3)
    eng = HipEngine(op)
    terms = [[(c, 6, j, k, 1.0)] for c in (1, 2) for j in range(8, 16, 3) for k in range(8, 16, 3)]
    terms += [[(c, 6, j, k, 0.5), (c, 5, j, k, 0.5)] for c in (4, 5) for j in range(8, 16, 3) for k in range(8, 16, 3)]
    freqs = [3e9, 7e9, 11e9]
    x = eng.dft(terms, freqs, record_steps=steps)
    sim.add_signals(terms)
    t0, tau = X.openems_gaussian_params(15e9)
    smp = X.schedule(lambda t: X.gaussian(t, t0, tau), 0, steps, op.dt)
    from openems_amd import ScheduledExcitation
    eng.run_batch(EngineBatch(num_steps=steps, excitations=[ScheduledExcitation((12, 12, 12), 2, smp)], extensions=[x]))
    sim.run_batch(steps, [(12, 12, 12, 2, True)], smp.astype(np.float64)[None])
    g = x.series()
    o = sim.signal_series(0, len(terms))
    assert np.array_equal(g, o)
    half = len(terms) // 2
    st = np.arange(steps)
    ref = np.concatenate([P.dft(o[:, :half], (st + 1) * op.dt, freqs), P.dft(o[:, half:], (st + 0.5) * op.dt, freqs)], 1)
    got = x.dft_values()
    assert np.allclose(got, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
