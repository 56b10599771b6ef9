"""Host logic of the learned-stop sweep (no GPU): the closed-form DFT of a
mode sum against a direct sum, the training-pair layout, and CalcPort."""
import math

import numpy as np

from openems_amd.extrapolator import training_pair
from openems_amd.learned_stop import calc_s11, mode_params, tail_dft


def test_tail_dft_closed_form_matches_direct_sum():
    rng = np.random.default_rng(0)
    s = -rng.uniform(0.001, 0.05, 3) + 2j * math.pi * rng.uniform(0.01, 0.2, 3)
    a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    dt, D, start, k0, k1 = 1e-12, 8, 100, 140, 900
    freqs = np.array([1e9, 7e9, 19e9])
    k = np.arange(k0, k1)
    x = np.real((a[:, None] * np.exp(s[:, None] * (k - start) / D)).sum(0))
    for off in (1.0, 0.5):
        ref = np.array([2 * dt * (x * np.exp(-2j * math.pi * f * (k + off) * dt)).sum() for f in freqs])
        got = tail_dft(s, a, start, k0, k1, D, dt, off, freqs)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_mode_params_match_network_definition():
    o = np.array([[0.3, -1.0, 1.0, 2.0, -0.5, 0.25]], np.float32)
    s, aV, aI = mode_params(o, 1)
    alpha = math.log1p(math.exp(0.3)) * 0.05
    f = 0.5 / (1 + math.exp(1.0))
    assert abs(s[0, 0] - (-alpha + 2j * math.pi * f)) < 1e-7
    assert aV[0, 0] == 1 + 2j and aI[0, 0] == -0.5 + 0.25j


def test_training_pair_layout():
    v = np.arange(40, dtype=float)
    i = -np.arange(40, dtype=float)
    x, t, (sv, si) = training_pair(v, i, window=16, decimate=4, tail=5)
    assert x.shape == (2, 4) and t.shape == (2, 5)
    assert sv == 15 and si == 15
    assert np.allclose(x[0], np.array([0, 4, 8, 12]) / 15)
    assert np.allclose(t[0], np.array([16, 20, 24, 28, 32]) / 15)
    assert np.allclose(t[1], -np.array([16, 20, 24, 28, 32]) / 15)


def test_calc_s11_limits():
    I = np.array([1 + 1j, 2.0])
    assert np.allclose(calc_s11(50 * I, I), 0)
    assert np.allclose(calc_s11(0 * I, I), -1)


def test_tail_schedule():
    from openems_amd.learned_stop import tail_schedule
    s = tail_schedule(10, 512)
    # 16, 32, ..., 256 over the first 7 epochs, then the full tail
    assert s[:1] == [16] and s[7:] == [512, 512, 512]
    assert all(a <= b for a, b in zip(s, s[1:]))
    assert set(s[:7]) <= {16, 32, 64, 128, 256}
    assert tail_schedule(4, 10) == [10] * 4          # shorter than the first stage
    assert tail_schedule(3, 64, first=16, share=0.0) == [64] * 3


def test_refit_residues_recovers_modes_and_ignores_huge_damping():
    from openems_amd.learned_stop import modes_eval, refit_residues
    rng = np.random.default_rng(1)
    s = np.array([-0.01 + 2j * math.pi * 0.05, -0.003 + 2j * math.pi * 0.11, -0.02 + 2j * math.pi * 0.3])
    a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    k = np.arange(3000)
    x = modes_eval(s, a, k, 2000, 8)
    # an extra pole decaying by e^-50 per sample overflows a naive design matrix
    got, r = refit_residues(x, np.append(s, -50 + 1j), 0, 2000, 2000, 8)
    assert np.abs(got[:3] - a).max() < 1e-10 and abs(got[3]) < 1e-10 and r < 1e-20
    # held-out continuation
    assert np.abs(modes_eval(s, got[:3], k[2000:], 2000, 8) - x[2000:]).max() < 1e-10


def test_refit_of_stacked_signals_equals_one_at_a_time():
    """V and I share the poles, so one factorisation serves both: the stacked
    refit gives each signal's residues and residual as its own refit does."""
    from openems_amd.learned_stop import modes_eval, refit_residues
    rng = np.random.default_rng(2)
    s = np.array([-0.01 + 2j * math.pi * 0.05, -0.003 + 2j * math.pi * 0.11, -0.02 + 2j * math.pi * 0.3])
    k = np.arange(4000)
    sig = [modes_eval(s, rng.standard_normal(3) + 1j * rng.standard_normal(3), k, 3000, 8)
           + 1e-3 * rng.standard_normal(len(k)) for _ in range(2)]
    a2, r2 = refit_residues(np.stack(sig), s, 1000, 3000, 3000, 8, 4)
    assert a2.shape == (2, 3) and r2.shape == (2,)
    for c in range(2):
        a1, r1 = refit_residues(sig[c], s, 1000, 3000, 3000, 8, 4)
        assert isinstance(r1, float)
        assert np.abs(a2[c] - a1).max() <= 1e-12 * np.abs(a1).max()
        assert abs(r2[c] - r1) <= 1e-12 * r1
        assert 0 < r1 < 1e-4  # the noise floor, not a failed fit


def test_predict_stop_on_exponential_decay():
    from openems_amd.learned_stop import predict_stop
    # energy exp(-t / 400) from a peak of 1: -50 dB at t = 400 * 5 ln 10 = 4605.2,
    # the first check below is 4700, the run then holds 4701 steps
    ts = np.arange(0, 3000, 100)
    en = np.exp(-ts / 400.0)
    assert predict_stop(ts, en, 3000, 20000) == 4701
    assert predict_stop(ts, en, 3000, 4000) == 4000           # capped
    assert predict_stop(ts, np.ones_like(en), 3000, 20000) == 20000  # no decay
    # only checks before pos count
    assert predict_stop(ts, en, 1000, 20000) == 4701


def test_pencil_poles_recover_damped_modes():
    from openems_amd.learned_stop import modes_eval, pencil_poles
    s = np.array([-0.02 + 2j * math.pi * 0.04, -0.005 + 2j * math.pi * 0.13])
    k = np.arange(4096)
    u = modes_eval(s, np.array([1.0 + 0.5j, -0.3 + 0.2j]), k, 0, 8)
    i = modes_eval(s, np.array([0.2 - 0.1j, 0.7 + 0.0j]), k, 0, 8)
    got = pencil_poles(u, i, 0, 4096, 8, 4)
    for sm in s:
        assert np.min(np.abs(got - sm)) < 1e-8


def test_learned_s11_equals_plain_for_exact_modes():
    """For port signals that are exact damped modes, the learned S11 (the
    simulated DFT to the checkpoint, modes refitted from the window, their
    closed-form DFT from the checkpoint to the plain run's stop) equals the
    plain run's own S11 over [0, S): the horizon rule leaves nothing but the
    extrapolation itself."""
    from openems_amd.learned_stop import calc_s11, modes_eval, refit_residues, tail_dft
    dt, D, P, W, S = 9.5e-13, 8, 2304, 2048, 5001
    s = np.array([-0.004 + 2j * math.pi * 0.03, -0.01 + 2j * math.pi * 0.09, -0.002 + 2j * math.pi * 0.15])
    k = np.arange(S)
    u = modes_eval(s, np.array([1.0 - 0.4j, 0.3 + 0.2j, -0.2 + 0.1j]), k, 0, D)
    i = modes_eval(s, np.array([0.02 + 0.01j, -0.004j, 0.003 + 0.0j]), k, 0, D)
    fr = np.linspace(1e9, 20e9, 64)

    def dft(x, k0, k1, off):
        kk = np.arange(k0, k1)
        return 2 * dt * np.exp(-2j * math.pi * np.outer(fr, (kk + off) * dt)) @ x[k0:k1]

    plain = calc_s11(dft(u, 0, S, 1.0), dft(i, 0, S, 0.5))
    aU, _ = refit_residues(u, s, P - W, P, P, D, 4)
    aI, _ = refit_residues(i, s, P - W, P, P, D, 4)
    U = dft(u, 0, P, 1.0) + tail_dft(s, aU, P, P, S, D, dt, 1.0, fr)
    I = dft(i, 0, P, 0.5) + tail_dft(s, aI, P, P, S, D, dt, 0.5, fr)
    assert np.abs(calc_s11(U, I) - plain).max() < 1e-9
    # stopping the tail at the cap instead of the plain run's stop does not
    # reproduce the plain run (the bias the horizon rule removes)
    U2 = dft(u, 0, P, 1.0) + tail_dft(s, aU, P, P, 20000, D, dt, 1.0, fr)
    I2 = dft(i, 0, P, 0.5) + tail_dft(s, aI, P, P, 20000, D, dt, 0.5, fr)
    assert np.abs(calc_s11(U2, I2) - plain).max() > 1e-4
