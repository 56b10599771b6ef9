"""Pins the oracle's extension hooks (the checker of the engine's
host-callback extension) to the reference's per-step order
(src/extensions/extension_trait.rs:8-17; src/fdtd/basic_engine.rs:56-163):

- post_update_e runs after the E update and before the excitations, so adding
  v to one E cell there equals a soft source of v at that cell;
- pre_update_h of step n runs after step n-1 ended, so changing H there
  equals changing it between two one-step batches;
- pre_batch / post_batch run once around the batch with the timestep at its
  ends; check_termination ends the batch at the first check step it asks for
  (timestep > 0, multiple of check_interval), and the energy-decay test wins
  when both fire at the same check.
"""
import numpy as np
import pytest

from oracle import excitation as X
from oracle import oracle as O

N = (14, 12, 10)
CELL = (7, 6, 5)


def _sim(prec=32):
    s = O.OracleSim(*N, 1e-3, prec=prec)
    s.set_material(2.0, 1.0, 0.01, 0.0, (2, 9, 3, 10, 1, 8))
    s.add_cpml((3,) * 6, mode=O.MODE_PHYSICAL)
    return s


def _samples(n, dt):
    t0, tau = X.openems_gaussian_params(40e9)
    return X.schedule(lambda t: X.gaussian(t, t0, tau), 0, n, dt).astype(np.float64)


def test_post_update_e_equals_soft_source():
    steps = 30
    a, b = _sim(), _sim()
    smp = _samples(steps, a.dt)
    a.run_batch(steps, [(*CELL, 2, True)], smp[None])

    def hook(which, step, f):
        f[2][CELL] += np.float32(smp[step])

    b.set_hook(hook, hooks=(3,))
    b.run_batch(steps)
    assert np.array_equal(a.fields(), b.fields())
    assert np.abs(a.fields()).max() > 0


def test_pre_update_h_equals_change_between_batches():
    steps = 12
    a, b = _sim(), _sim()
    smp = _samples(steps, a.dt)
    for st in range(steps):
        f = a.fields()
        f[3][CELL] = np.float32(np.float32(f[3][CELL] * np.float32(0.5)) + np.float32(1e-3 * st))
        a.set_fields(f)
        a.run_batch(1, [(3, 3, 3, 1, True)], smp[None, st:st + 1])
    calls = []

    def hook(which, step, f):
        calls.append((which, step))
        f[3][CELL] = np.float32(np.float32(f[3][CELL] * np.float32(0.5)) + np.float32(1e-3 * step))

    b.set_hook(hook, hooks=(0,))
    b.run_batch(steps, [(3, 3, 3, 1, True)], smp[None])
    assert calls == [(0, t) for t in range(steps)]
    assert np.array_equal(a.fields(), b.fields())


def test_batch_hooks_and_termination():
    s = _sim()
    seen = []
    asked = []
    s.set_hook(lambda w, t, f: seen.append((w, t)), hooks=(4, 5),
               check_termination=lambda: asked.append(1) or len(asked) >= 3)
    r = s.run_batch(100, check_interval=5)
    # asked at timesteps 5, 10, 15: stops after the step of timestep 15 (16 steps)
    assert r.termination_reason == 2 and r.timesteps_executed == 16
    assert seen == [(4, 0), (5, 16)]
    # the decay test takes precedence when both end the batch at one check
    smp = _samples(400, s.dt)
    runs = {}
    for name, term in (("plain", None), ("stop", lambda: True)):
        t = _sim()
        if term:
            t.set_hook(check_termination=term)
        runs[name] = t.run_batch(400, [(*CELL, 2, True)], smp[None], check_interval=100, energy_decay_db=1e-9,
                                 sample_interval=1, track_peak=True)
    assert runs["plain"].termination_reason == 1 and runs["plain"].timesteps_executed == 101
    assert runs["stop"].termination_reason == 1 and runs["stop"].timesteps_executed == 101
    # without a decay test the request ends the batch at the first check
    t = _sim()
    t.set_hook(check_termination=lambda: True)
    r = t.run_batch(400, [(*CELL, 2, True)], smp[None], check_interval=100)
    assert r.termination_reason == 2 and r.timesteps_executed == 101


def test_hook_error_propagates():
    s = _sim()
    s.set_hook(lambda w, t, f: 5 if t == 3 else 0, hooks=(1,))
    with pytest.raises(RuntimeError):
        s.run_batch(10)
