"""The host-callback extension (an Extension without a device form, run on the
host: src/extensions/extension_trait.rs:32-166) against the oracle running the
same hooks on its own fields.

Every hook point changes the fields differently, next to device CPML, Mur and
ADE extensions and a source, so a hook run at the wrong point, on the wrong
half, before instead of after the device extensions of its point, or with the
wrong timestep changes the result.  The bar is bit-identical fields and the
same call sequence; check_termination and error propagation follow
src/fdtd/basic_engine.rs:115-147 and the reference's `?`.
"""
import numpy as np
import pytest

from openems_amd import (ADE_DRUDE, MODE_PHYSICAL, CpmlBoundaries, EngineBatch, HipEngine, Operator,
                         ScheduledExcitation)
from openems_amd._ffi import OEError
from oracle import excitation as X
from oracle import oracle as O
from tests.parity_util import say

pytestmark = pytest.mark.gpu

SHAPE = (20, 22, 24)
D = 0.5e-3
SRC = (10, 11, 12)
F32 = np.float32


def _sphere(center, r):
    g = np.stack(np.meshgrid(*[np.arange(s) for s in SHAPE], indexing="ij"), -1)
    return np.argwhere(np.sum((g - np.array(center)) ** 2, -1) <= r * r).astype(np.uint32)


def _pair(mur=True, ade=True):
    """GPU engine and oracle with the same materials and device extensions."""
    op = Operator.uniform(*SHAPE, D)
    sim = O.OracleSim(*SHAPE, D)
    for s in (op, sim):
        s.set_material(3.0, 1.0, 0.02, 0.0, (4, 15, 5, 17, 6, 18))
    eng = HipEngine(op)
    layers = (3, 3, 3, 3, 3, 0 if mur else 3)
    exts = [eng.cpml(CpmlBoundaries(layers=layers, mode=MODE_PHYSICAL))]
    sim.add_cpml(layers, mode=O.MODE_PHYSICAL)
    if mur:
        en = (False,) * 5 + (True,)
        exts.append(eng.mur(en))
        sim.add_mur(en)
    if ade:
        cells = _sphere((12, 8, 9), 3)
        exts.append(eng.ade(ADE_DRUDE, (2e11, 1e9, 0.0), cells, MODE_PHYSICAL))
        sim.add_ade(ADE_DRUDE, (2e11, 1e9, 0.0), cells, O.MODE_PHYSICAL)
    return eng, sim, exts


def _samples(n, dt):
    t0, tau = X.openems_gaussian_params(40e9)
    return X.schedule(lambda t: X.gaussian(t, t0, tau), 0, n, dt)


class Hooks:
    """The same hook body for both sides; records (which, step)."""

    def __init__(self, smp):
        self.smp = smp
        self.calls = []

    def __call__(self, which, step, f):
        self.calls.append((which, step))
        if which == 0:    # pre_update_h: scale one H value
            f[3][(9, 9, 9)] = F32(f[3][(9, 9, 9)] * F32(0.5))
        elif which == 1:  # post_update_h: add to one H value
            f[5][(8, 12, 10)] = F32(f[5][(8, 12, 10)] + F32(1e-6 * step))
        elif which == 2:  # pre_update_e: scale one E value on the Mur line plane
            f[0][(10, 11, SHAPE[2] - 2)] = F32(f[0][(10, 11, SHAPE[2] - 2)] * F32(0.75))
        elif which == 3:  # post_update_e: a soft source written as a hook
            f[2][SRC] = F32(f[2][SRC] + F32(self.smp[step]))
        elif which == 4:  # pre_batch
            f[1][(5, 5, 5)] = F32(f[1][(5, 5, 5)] + F32(1e-3))
        elif which == 5:  # post_batch
            f[4][(6, 7, 8)] = F32(0.0)
        return 0


def test_every_hook_point_matches_oracle():
    steps = 150
    eng, sim, exts = _pair()
    smp = _samples(steps, eng.dt)
    hg, ho = Hooks(smp), Hooks(smp)
    src = [ScheduledExcitation((14, 9, 7), 1, _samples(steps, eng.dt), True)]
    hx = eng.host_callback(hg, hooks=range(6))
    eng.run_batch(EngineBatch(num_steps=steps, excitations=src, extensions=exts + [hx]))
    sim.set_hook(ho, hooks=range(6))
    sim.run_batch(steps, [(14, 9, 7, 1, True)], _samples(steps, eng.dt).astype(np.float64)[None])
    E, H = eng.read_fields()
    g = np.concatenate([E, H])
    o = sim.fields()
    diff = int(np.count_nonzero(g != o))
    say(f"[host] all hooks, CPML+Mur+Drude, {steps} steps: differing {diff}, max |g| {np.abs(g).max():.3e}")
    assert hg.calls == ho.calls
    assert hg.calls[0] == (4, 0) and hg.calls[-1] == (5, steps)
    assert len(hg.calls) == 4 * steps + 2
    assert diff == 0
    assert np.abs(g).max() > 0


def test_post_update_e_hook_equals_device_source():
    """A post_update_e hook adding the samples equals the same soft source
    given as a device excitation (sources act after the post-E hooks)."""
    steps = 120
    a, _, ea = _pair(mur=False)
    b, _, eb = _pair(mur=False)
    smp = _samples(steps, a.dt)
    a.run_batch(EngineBatch(num_steps=steps, excitations=[ScheduledExcitation(SRC, 2, smp, True)], extensions=ea))

    def hook(which, step, f):
        f[2][SRC] = F32(f[2][SRC] + F32(smp[step]))

    b.run_batch(EngineBatch(num_steps=steps, extensions=eb + [b.host_callback(hook, hooks=("post_update_e",))]))
    fa = np.concatenate(a.read_fields())
    fb = np.concatenate(b.read_fields())
    assert np.array_equal(fa, fb) and np.abs(fa).max() > 0


def test_check_termination_and_reason():
    eng, sim, exts = _pair(mur=False, ade=False)
    asked = []

    def stop():
        asked.append(eng.timestep)
        return "three checks" if len(asked) >= 3 else None

    hx = eng.host_callback(check_termination=stop)
    r = eng.run_batch(EngineBatch(num_steps=100, extensions=exts + [hx], check_interval=5))
    # asked at timesteps 5, 10 and 15 (after each of those steps): 16 steps
    assert r.termination_reason == 2 and r.timesteps_executed == 16
    assert hx.stop_reason() == "three checks"
    n = []
    sim.set_hook(check_termination=lambda: n.append(1) or len(n) >= 3)
    ro = sim.run_batch(100, check_interval=5)
    assert (ro.termination_reason, ro.timesteps_executed) == (2, 16)
    assert np.array_equal(np.concatenate(eng.read_fields()), sim.fields())
    # the decay test wins when both end the batch at the same check
    for term, want in ((None, 1), (lambda: "always", 1)):
        e2, _, x2 = _pair(mur=False, ade=False)
        smp = _samples(400, e2.dt)
        extra = [e2.host_callback(check_termination=term)] if term else []
        r2 = e2.run_batch(EngineBatch(num_steps=400, excitations=[ScheduledExcitation(SRC, 2, smp, True)],
                                      extensions=x2 + extra, check_interval=100, energy_decay_db=1e-9,
                                      sample_interval=1, track_peak=True))
        assert (r2.termination_reason, r2.timesteps_executed) == (want, 101)
    # without a decay test the request ends the batch at the first check
    e3, _, x3 = _pair(mur=False, ade=False)
    r3 = e3.run_batch(EngineBatch(num_steps=400, extensions=x3 + [e3.host_callback(check_termination=lambda: "now")],
                                  check_interval=100))
    assert (r3.termination_reason, r3.timesteps_executed) == (2, 101)


def test_hook_errors_propagate():
    eng, _, exts = _pair(mur=False, ade=False)
    seen = []

    def hook(which, step, f):
        seen.append((which, step))
        return 4 if step == 3 else 0  # OE_ERR_MATERIAL

    with pytest.raises(OEError) as ei:
        eng.run_batch(EngineBatch(num_steps=10, extensions=exts + [eng.host_callback(hook, hooks=(1, 5))]))
    assert ei.value.code == 4
    assert seen[-1] == (1, 3) and (5, 4) not in seen  # no post_batch after an error

    def boom(which, step, f):
        raise ValueError("hook failed")

    eng2, _, exts2 = _pair(mur=False, ade=False)
    with pytest.raises(ValueError, match="hook failed"):
        eng2.run_batch(EngineBatch(num_steps=4, extensions=exts2 + [eng2.host_callback(boom, hooks=(2,))]))
