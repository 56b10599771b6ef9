"""The learned-stop sweep end to end on small microstrip structures.

With gates that never open, every structure runs checkpoint to checkpoint
in new sweeps over the same engines (excitations shifted by the steps
already run) and must end exactly as the plain sweep: same steps, same
S11, bit for bit.  With gates that always open, every structure stops at
the first checkpoint with S11 from the simulated DFT plus the modes' tail.
"""
import numpy as np
import pytest

from openems_amd import learned_stop as LS
from openems_amd import workloads as W
from openems_amd.extrapolator import Extrapolator, ExtrapolatorDims
from tests.parity_util import say

pytestmark = pytest.mark.gpu

WINDOW, DEC, CHECK, MAX_STEPS = 512, 8, 128, 3000


def _structs():
    rng = np.random.default_rng(31)
    out = []
    for shape in ((60, 60, 32), (72, 64, 36), (64, 60, 40)):
        ms = W.microstrip(rng)
        w = min(ms.width, 4)
        segs = [(16, shape[0] - 16, 16, 16 + w)]
        out.append(W.Microstrip(shape, ms.eps_r, min(ms.h, 5), w, segs, (16 + w // 2, 16 + w // 2)))
    return out


def _model():
    return Extrapolator(ExtrapolatorDims(L0=WINDOW // DEC, c1=8, c2=16, c3=16, hidden=32, modes=4, batch=4,
                                         tail=64), seed=3)


def test_rejected_structures_continue_to_the_plain_result():
    structs = _structs()
    plain = LS.run_sweep(structs, MAX_STEPS, 50.0, record_steps=0, autotune=False)
    res = LS.learned_sweep(_model(), structs, MAX_STEPS, WINDOW, DEC, CHECK, max_residual=-1.0, stride=448)
    say(f"[learned] checkpoints {res.checkpoints}; steps {res.steps_learned.tolist()} vs plain {plain.steps.tolist()}")
    assert not res.accepted.any()
    assert len(res.checkpoints) >= 2
    assert np.array_equal(res.steps_learned, plain.steps)
    assert np.array_equal(res.s11, plain.s11)


def test_open_gates_stop_at_the_first_checkpoint():
    structs = _structs()
    kw = dict(max_residual=np.inf, max_tail_energy=np.inf, max_horizon_change=np.inf, min_decay_db=0.0, stride=448)
    res = LS.learned_sweep(_model(), structs, MAX_STEPS, WINDOW, DEC, CHECK, overlap=False, **kw)
    say(f"[learned] open gates: steps {res.steps_learned.tolist()}, checkpoints {res.checkpoints}")
    early = WINDOW + CHECK
    # a structure stops at the first checkpoint unless it decayed before it
    for q in range(len(structs)):
        assert res.accepted[q] == (res.steps_learned[q] == early)
        assert res.steps_learned[q] <= early
    assert res.accepted.any()
    assert np.isfinite(res.s11).all()
    # assessed while the next segment runs: the same structures accepted with
    # the same S11 (formed from the state at the checkpoint); each accepted
    # one ran that segment too
    ov = LS.learned_sweep(_model(), structs, MAX_STEPS, WINDOW, DEC, CHECK, overlap=True, **kw)
    say(f"[learned] open gates, overlapped: steps {ov.steps_learned.tolist()}")
    assert np.array_equal(ov.accepted, res.accepted)
    assert np.array_equal(ov.s11, res.s11)
    acc = res.accepted
    assert np.all(ov.steps_learned[acc] > early)
    assert np.array_equal(ov.steps_learned[~acc], res.steps_learned[~acc])
