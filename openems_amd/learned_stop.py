"""Learned-stop sweeps: stop each structure early and continue its port
signals with the extrapolator's modes instead of running to -50 dB.

  plain run    every structure runs until its energy is 50 dB below its
               whole-run peak (or max_steps); S11 from the port's running DFT
  learned run  every structure runs `window + check` steps; the network maps
               the decimated first `window` steps of V and I to damped modes;
               the modes' prediction over the next `check` steps is compared
               with the simulated samples there: if the squared error exceeds
               `max_residual` times the energy of the window, the structure
               falls back to the plain run.  Otherwise its S11 combines the
               simulated DFT over [0, window + check) with the closed-form DFT
               of the predicted modes from window + check to max_steps.

The DFT of a mode sum is a geometric series per mode (U sampled at (k + 1)
dt, I at (k + 1/2) dt, as the engine samples E and H), so no tail samples
are synthesised for S11.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import workloads as W
from .engine import Sweep
from .extrapolator import Extrapolator, ExtrapolatorDims, training_pair

A_SCALE = 0.05
F_MAX = 0.5


@dataclass
class SweepRun:
    steps: np.ndarray               # per structure
    u: List[np.ndarray]             # recorded port U series (float64)
    i: List[np.ndarray]
    s11: np.ndarray                 # [n_struct, n_freq]
    dft: np.ndarray                 # [n_struct, n_freq, 2] complex (U, I), 2 dt X convention
    wall: float
    dt: float
    freqs: np.ndarray


def run_sweep(structs: List[W.Microstrip], max_steps: int, decay_db: Optional[float], record_steps: int,
              freqs=None, check: int = 100, autotune: bool = True) -> SweepRun:
    fr = np.linspace(1e9, 20e9, 64) if freqs is None else np.asarray(freqs)
    sw = Sweep()
    keep = []
    dt = None
    for ms in structs:
        op, eng, cp, port, exc = W.build_microstrip_hip(ms, max_steps, record_steps=record_steps, freqs=fr)
        sw.add(eng, cp, port, exc)
        keep.append((eng, port))
        dt = op.dt
    t0 = time.perf_counter()
    res = sw.run(max_steps, energy_decay_db=decay_db, check_interval=check, autotune=autotune)
    wall = time.perf_counter() - t0
    u, i, s11, dft = [], [], [], []
    for eng, port in keep:
        ser = port.series()
        u.append(ser[:, 0].copy())
        i.append(ser[:, 1].copy())
        s11.append(port.s11(50.0))
        dft.append(port.dft_values())
    return SweepRun(np.array([r.timesteps_executed for r in res]), u, i, np.array(s11), np.array(dft), wall, dt, fr)


def mode_params(o: np.ndarray, n_modes: int):
    """Network outputs [B, 6 M] -> (s [B, M] complex per decimated sample, aV, aI)."""
    o = np.asarray(o, np.float64).reshape(o.shape[0], n_modes, 6)
    alpha = np.logaddexp(0.0, o[..., 0]) * A_SCALE
    f = F_MAX / (1.0 + np.exp(-o[..., 1]))
    return -alpha + 2j * math.pi * f, o[..., 2] + 1j * o[..., 3], o[..., 4] + 1j * o[..., 5]


def tail_dft(s, a, start: int, k0: int, k1: int, decimate: int, dt: float, offset: float, freqs) -> np.ndarray:
    """2 dt sum_{k=k0}^{k1-1} x(k) exp(-j 2 pi f (k + offset) dt) for
    x(k) = sum_m Re(a_m exp(s_m (k - start) / decimate)); s, a: [M]."""
    out = np.zeros(len(freqs), np.complex128)
    n = k1 - k0
    if n <= 0:
        return out
    for q, f in enumerate(freqs):
        wdt = 2 * math.pi * f * dt
        tot = 0.0 + 0.0j
        for sm, am in ((s, a), (np.conj(s), np.conj(a))):
            lq = sm / decimate - 1j * wdt                  # log of the per-step ratio
            first = np.exp((k0 - start) * sm / decimate - 1j * wdt * (k0 + offset))
            ratio = np.where(np.abs(lq) > 1e-12, (1 - np.exp(n * lq)) / (1 - np.exp(lq)), n)
            tot += 0.5 * (am * first * ratio).sum()
        out[q] = 2 * dt * tot
    return out


def calc_s11(U, I, z_ref=50.0):
    inc = 0.5 * (U + I * z_ref)
    return (U - inc) / inc


@dataclass
class LearnedResult:
    s11: np.ndarray
    accepted: np.ndarray            # bool per structure
    residual: np.ndarray            # validation error / window energy
    wall_learned: float             # early sweep + inference + fallback runs
    wall_fallback: float
    steps_learned: np.ndarray


def learned_sweep(model: Extrapolator, structs, max_steps: int, window: int, decimate: int, check: int,
                  max_residual: float = 1e-4, freqs=None, decay_db: float = 50.0) -> LearnedResult:
    d = model.dims
    early = window + check
    t0 = time.perf_counter()
    run = run_sweep(structs, early, None, record_steps=early, freqs=freqs)
    X = np.zeros((len(structs), 2, d.L0), np.float32)
    scales = []
    for q in range(len(structs)):
        x, _, sc = training_pair(run.u[q], run.i[q], window, decimate, 1)
        X[q] = x
        scales.append(sc)
    modes = np.zeros((len(structs), 6 * d.modes), np.float32)
    for b0 in range(0, len(structs), d.batch):
        m, _ = model.forward(X[b0:b0 + d.batch], tails=False)
        modes[b0:b0 + d.batch] = m
    s, aV, aI = mode_params(modes, d.modes)
    n_chk = check // decimate
    n = np.arange(n_chk)
    s11 = run.s11.copy()
    accepted = np.zeros(len(structs), bool)
    resid = np.zeros(len(structs))
    for q in range(len(structs)):
        sv, si = scales[q]
        z = np.exp(s[q][:, None] * n)
        pv = np.real((aV[q][:, None] * z).sum(0))
        pi = np.real((aI[q][:, None] * z).sum(0))
        av = run.u[q][window:early:decimate][:n_chk] / sv
        ai = run.i[q][window:early:decimate][:n_chk] / si
        e_win = (run.u[q][:window:decimate] / sv) ** 2 + (run.i[q][:window:decimate] / si) ** 2
        resid[q] = float(((pv - av) ** 2 + (pi - ai) ** 2).sum() / max(e_win.sum(), 1e-30))
        if resid[q] <= max_residual:
            accepted[q] = True
            U = run.dft[q, :, 0] + sv * tail_dft(s[q], aV[q], window, early, max_steps, decimate, run.dt, 1.0, run.freqs)
            I = run.dft[q, :, 1] + si * tail_dft(s[q], aI[q], window, early, max_steps, decimate, run.dt, 0.5, run.freqs)
            s11[q] = calc_s11(U, I)
    wall_early = time.perf_counter() - t0
    t1 = time.perf_counter()
    steps = np.full(len(structs), early)
    rej = [q for q in range(len(structs)) if not accepted[q]]
    if rej:
        full = run_sweep([structs[q] for q in rej], max_steps, decay_db, record_steps=0, freqs=freqs)
        for k, q in enumerate(rej):
            s11[q] = full.s11[k]
            steps[q] = full.steps[k]
    wall_fb = time.perf_counter() - t1
    return LearnedResult(s11, accepted, resid, wall_early + wall_fb, wall_fb, steps)


def dataset(run: SweepRun, window: int, decimate: int, tail: int):
    X, T = [], []
    for q in range(len(run.u)):
        x, t, _ = training_pair(run.u[q], run.i[q], window, decimate, tail)
        X.append(x)
        T.append(t)
    return np.array(X), np.array(T)


def train(model: Extrapolator, X: np.ndarray, T: np.ndarray, epochs: int, lr: float, seed: int = 0,
          log=None) -> List[float]:
    rng = np.random.default_rng(seed)
    B = model.dims.batch
    losses = []
    steps = epochs * max(1, len(X) // B)
    k = 0
    for ep in range(epochs):
        perm = rng.permutation(len(X))
        tot = 0.0
        nb = 0
        for b0 in range(0, len(X) - B + 1 if len(X) >= B else 1, B):
            idx = perm[b0:b0 + B]
            lr_k = lr * 0.5 * (1 + math.cos(math.pi * k / max(steps, 1)))
            tot += model.train_step(X[idx], T[idx], max(lr_k, lr * 1e-2))
            nb += 1
            k += 1
        losses.append(tot / max(nb, 1))
        if log:
            log(f"epoch {ep}: loss {losses[-1]:.4e}")
    return losses
