"""Learned-stop sweeps: stop each structure early and continue its port
signals with damped modes whose poles the extrapolator predicts, instead of
running to -50 dB.

  plain run    every structure runs until its energy is 50 dB below its
               whole-run peak (or max_steps); S11 from the port's running DFT
  learned run  all structures run together to a checkpoint P (first
               window + check, then every `stride` steps).  At P the network
               maps the decimated V and I of [P - check - window, P - check)
               to damped modes; their poles are kept and the residues of V
               and I refitted by least squares to the simulated samples.  A
               structure stops when the refitted modes reproduce the held-out
               steps [P - check, P), the energy they predict after P is small,
               and S11 hardly moves when the predicted stop moves; its S11 is
               the simulated DFT over [0, P) plus the closed-form DFT of the
               modes from P to the step at which the plain run's decay test
               is predicted to stop it (the plain run's own horizon).  The
               others continue from their state at P (same fields, DFT sums,
               energy peak and decay-check steps as the plain run), so a
               structure that is never accepted ends exactly as in the plain
               run, with the same S11.

The DFT of a mode sum is a geometric series per mode (U sampled at (k + 1)
dt, I at (k + 1/2) dt, as the engine samples E and H), so no tail samples
are synthesised for S11.
"""
from __future__ import annotations

import contextlib
import math
import os
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np

from . import workloads as W
from .engine import ScheduledExcitation, Sweep
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


def shifted(exc, k: int):
    """Excitations of a run that starts k steps into the schedule."""
    return [ScheduledExcitation(e.position, e.direction, e.samples[k:], e.soft_source) for e in exc]


def build(structs: List[W.Microstrip], max_steps: int, record_steps: int, freqs):
    return [W.build_microstrip_hip(ms, max_steps, record_steps=record_steps, freqs=freqs) for ms in structs]


def run_sweep(structs: List[W.Microstrip], max_steps: int, decay_db: Optional[float], record_steps: int,
              freqs=None, check: int = 100, autotune: bool = True) -> SweepRun:
    fr = np.linspace(1e9, 20e9, 64) if freqs is None else np.asarray(freqs)
    built = build(structs, max_steps, record_steps, fr)
    sw = Sweep()
    for op, eng, cp, port, exc in built:
        sw.add(eng, cp, port, exc)
    t0 = time.perf_counter()
    res = sw.run(max_steps, energy_decay_db=decay_db, check_interval=check, autotune=autotune)
    wall = time.perf_counter() - t0
    u, i, s11, dft = [], [], [], []
    for op, eng, cp, port, exc in built:
        ser = port.series()
        u.append(ser[:, 0].copy())
        i.append(ser[:, 1].copy())
        s11.append(port.s11(50.0))
        dft.append(port.dft_values())
    return SweepRun(np.array([r.timesteps_executed for r in res]), u, i, np.array(s11), np.array(dft), wall,
                    built[0][0].dt, fr)


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


def refit_residues(x: np.ndarray, s: np.ndarray, k0: int, k1: int, ref: int, decimate: int, stride: int = 1):
    """Least-squares residues for given poles: the complex a [M] minimising
    sum_k (x[k] - Re sum_m a_m exp(s_m (k - ref) / decimate))^2 over
    k = k0, k0 + stride, ... < k1 (s per decimated sample).  Columns are
    scaled to unit norm before the solve (poles of very different damping
    make them span many decades); returns (a, relative residual).  x may
    also stack several signals [c, n] that share the poles (port V and I):
    one factorisation then serves all of them, and a is [c, M], the
    residual [c]."""
    x = np.asarray(x)
    if x.ndim == 2:
        a, r = _refit(x, s, k0, k1, ref, decimate, stride)
        return a, r
    a, r = _refit(x[None], s, k0, k1, ref, decimate, stride)
    return a[0], float(r[0])


def _refit(x: np.ndarray, s: np.ndarray, k0: int, k1: int, ref: int, decimate: int, stride: int):
    k = np.arange(k0, k1, stride)
    t = (k - ref) / decimate
    # each column divided by its largest modulus first (exp of the largest
    # real exponent over the range), so no column overflows
    shift = np.maximum(s.real * t[0], s.real * t[-1])
    A = np.exp(np.outer(t, s) - shift)                            # [K, M] complex
    D = np.concatenate([A.real, -A.imag], axis=1)               # x = D [Re a; Im a]
    sc = np.linalg.norm(D, axis=0)
    sc[sc == 0] = 1.0
    y = x[:, k0:k1:stride].astype(np.float64).T                  # [K, c]
    sol, *_ = np.linalg.lstsq(D / sc, y, rcond=1e-12)
    sol = sol / sc[:, None]
    fit = D @ sol
    M = len(s)
    with np.errstate(over="ignore", under="ignore"):
        g = np.exp(-shift)
    a = (sol[:M].T + 1j * sol[M:].T) * g                        # [c, M]
    a[~np.isfinite(a)] = 0.0
    r = np.sum((fit - y) ** 2, axis=0) / np.maximum(np.sum(y * y, axis=0), 1e-300)
    return a, r


def modes_eval(s: np.ndarray, a: np.ndarray, k: np.ndarray, ref: int, decimate: int) -> np.ndarray:
    """Re sum_m a_m exp(s_m (k - ref) / decimate) at the steps k."""
    return np.real(np.exp(np.outer((np.asarray(k) - ref) / decimate, s)) @ a)


def calc_s11(U, I, z_ref=50.0):
    inc = 0.5 * (U + I * z_ref)
    return (U - inc) / inc


@dataclass
class LearnedResult:
    s11: np.ndarray
    accepted: np.ndarray            # bool per structure
    residual: np.ndarray            # validation error / window energy at the last checkpoint seen
    wall_learned: float             # all sweep segments + inference
    wall_inference: float
    steps_learned: np.ndarray       # steps each structure ran
    checkpoints: List[int] = field(default_factory=list)
    tail_energy: Optional[np.ndarray] = None  # predicted tail energy / energy so far, last checkpoint seen


def _window(u: np.ndarray, i: np.ndarray, a: int, window: int, decimate: int, n_chk: int):
    """Network input [2, window/decimate] of V, I over [a, a + window), the
    next n_chk decimated samples (the check), and the window peaks."""
    L0 = window // decimate
    sv = max(np.abs(u[a:a + window]).max(), 1e-30)
    si = max(np.abs(i[a:a + window]).max(), 1e-30)
    x = np.stack([u[a:a + window:decimate][:L0] / sv, i[a:a + window:decimate][:L0] / si])
    b = a + window
    chk = np.stack([u[b:b + n_chk * decimate:decimate] / sv, i[b:b + n_chk * decimate:decimate] / si])
    return x.astype(np.float32), chk, (sv, si)


def predict_stop(timesteps: np.ndarray, energies: np.ndarray, pos: int, max_steps: int, decay_db: float = 50.0,
                 check: int = 100, last: int = 10) -> int:
    """The step count at which the plain run's decay test
    (src/fdtd/basic_engine.rs:128-141, whole-run peak) is predicted to stop:
    log-energy over the last `last` checks before `pos` extrapolated as a
    line to decay_db below the peak, rounded up to the next check (the run
    then holds that check's step, hence + 1); max_steps if the energy does
    not fall."""
    m = timesteps < pos
    ts, en = timesteps[m].astype(np.float64), energies[m]
    if len(ts) < 3 or not np.all(en > 0):
        return max_steps
    thr = en.max() * 10.0 ** (-decay_db / 10.0)
    b, a = np.polyfit(ts[-last:], np.log(en[-last:]), 1)
    if not b < 0:
        return max_steps
    t_cross = (math.log(thr) - a) / b
    c = max(int(math.ceil(t_cross / check)) * check, (pos // check + 1) * check)
    return int(min(c + 1, max_steps))


def pencil_poles(u: np.ndarray, i: np.ndarray, k0: int, k1: int, decimate: int, order: int) -> np.ndarray:
    """Poles per decimated sample of the decimated V and I over [k0, k1)
    by the matrix pencil method (Hua and Sarkar, IEEE Trans. ASSP 38(5),
    1990): a Hankel matrix of the samples with pencil parameter N/3, its
    `order` leading right singular vectors, and the eigenvalues z of the
    shifted pencil; s = log z, decaying ones with Im s >= 0 kept.  A
    classical stand-in for the network's poles, for comparison."""
    out = []
    for x in (u, i):
        y = x[k0:k1:decimate].astype(np.float64)
        y = y / max(np.abs(y).max(), 1e-300)
        N = len(y)
        L = max(N // 3, order + 1)
        Y = np.lib.stride_tricks.sliding_window_view(y, L + 1)
        _, _, Vh = np.linalg.svd(Y, full_matrices=False)
        V = Vh[:min(order, Vh.shape[0])].conj().T
        z = np.linalg.eigvals(np.linalg.pinv(V[:-1]) @ V[1:])
        out.append(np.log(z.astype(np.complex128)))
    s = np.concatenate(out)
    return s[(s.real < 0) & (s.real > -1.0) & (s.imag >= 0)]


def _passes(value: float, bar: float) -> bool:
    return bar == np.inf or (np.isfinite(value) and value <= bar)


def _one_thread_blas():
    """BLAS limited to one thread per caller while host threads each solve
    small least-squares problems (threads inside a 512 x 64 solve only slow
    it down)."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        return contextlib.nullcontext()
    return threadpool_limits(1)


@dataclass
class _Checkpoint:
    """What the assessment of the structures at checkpoint pos reads: the
    port series and DFT sums as they were at pos, the energy history up to
    pos, the poles (network) and the gate settings."""
    pos: int
    still: List[int]
    sers: list
    dfts: list
    hist: list                     # (timesteps, energies) per structure of still
    s_all: Optional[np.ndarray]    # network poles per structure, or None (pencil)
    window: int
    check: int
    decimate: int
    max_steps: int
    dt: float
    freqs: np.ndarray
    max_residual: float
    max_tail_energy: float
    horizon_window: int
    max_horizon_change: float
    decay_db: float
    decay_check: int
    pencil_order: int
    fit_stride: int


def _assess(c: _Checkpoint, k: int):
    """(accepted, S11, residual, tail energy fraction) of structure
    c.still[k] at checkpoint c.pos: the gates of learned_sweep."""
    pos, dec = c.pos, c.decimate
    a0, a1 = pos - c.window - c.check, pos - c.check
    u, i = c.sers[k][:, 0], c.sers[k][:, 1]
    if c.s_all is not None:
        s = c.s_all[k]
    else:
        try:
            s = pencil_poles(u, i, a0, a1, dec, c.pencil_order)
        except np.linalg.LinAlgError:
            return False, None, np.nan, np.nan
        if len(s) == 0:
            return False, None, np.nan, np.nan
    aUI, _ = refit_residues(np.stack([u, i]), s, a0, a1, a1, dec, c.fit_stride)
    kk = np.arange(a1, pos)
    eu = float(np.sum(u[a0:a1] ** 2))
    ei = float(np.sum(i[a0:a1] ** 2))
    ru = float(np.sum((modes_eval(s, aUI[0], kk, a1, dec) - u[a1:pos]) ** 2)) / max(eu, 1e-300)
    ri = float(np.sum((modes_eval(s, aUI[1], kk, a1, dec) - i[a1:pos]) ** 2)) / max(ei, 1e-300)
    res_q = max(ru, ri)
    if not _passes(res_q, c.max_residual):
        return False, None, res_q, np.nan
    aUI, _ = refit_residues(np.stack([u, i]), s, pos - c.window, pos, pos, dec, c.fit_stride)
    aU, aI = aUI[0], aUI[1]
    ts, en = c.hist[k]
    stop = predict_stop(ts, en, pos, c.max_steps, c.decay_db, c.decay_check)
    sv = max(float(np.abs(u[a0:pos]).max()), 1e-300)
    si = max(float(np.abs(i[a0:pos]).max()), 1e-300)
    kt = np.arange(pos, stop, dec)
    e_tail = float(np.sum((modes_eval(s, aU, kt, pos, dec) / sv) ** 2)
                   + np.sum((modes_eval(s, aI, kt, pos, dec) / si) ** 2))
    e_run = float(np.sum((u[:pos:dec] / sv) ** 2) + np.sum((i[:pos:dec] / si) ** 2))
    frac = e_tail / max(e_run, 1e-300)
    if not _passes(frac, c.max_tail_energy):
        return False, None, res_q, frac

    def s11_to(end):
        U = c.dfts[k][:, 0] + tail_dft(s, aU, pos, pos, end, dec, c.dt, 1.0, c.freqs)
        I = c.dfts[k][:, 1] + tail_dft(s, aI, pos, pos, end, dec, c.dt, 0.5, c.freqs)
        return calc_s11(U, I)

    s_pred = s11_to(stop)
    change = max(float(np.abs(s11_to(max(pos, stop - c.horizon_window)) - s_pred).max()),
                 float(np.abs(s11_to(min(c.max_steps, stop + c.horizon_window)) - s_pred).max()))
    if not _passes(change, c.max_horizon_change):
        return False, None, res_q, frac
    return True, s_pred, res_q, frac


def _assess_all(c: _Checkpoint, workers: int):
    """_assess over the checkpoint's structures, which are independent, on
    host threads (LAPACK releases the interpreter lock), each with a
    one-thread BLAS."""
    with _one_thread_blas(), ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(lambda k: _assess(c, k), range(len(c.still))))


def learned_sweep(model: Extrapolator, structs, max_steps: int, window: int, decimate: int, check: int,
                  max_residual: float = 1e-6, freqs=None, decay_db: float = 50.0, stride: Optional[int] = None,
                  decay_check: int = 100, max_tail_energy: float = 1e-4, horizon_window: int = 300,
                  max_horizon_change: float = 3e-4, poles: str = "network", pencil_order: int = 24,
                  min_decay_db: float = 20.0, fit_stride: int = 4, workers: Optional[int] = None,
                  overlap: bool = True, log=None) -> LearnedResult:
    """All structures run together from checkpoint to checkpoint (the first
    at window + check steps, then every `stride`).  At a checkpoint P the
    network maps the decimated V and I of [P - check - window, P - check) to
    damped modes; their poles are kept and the residues of V and I refitted
    by least squares to the simulated samples of that range
    (refit_residues).  A structure stops when
      (1) the refitted modes reproduce the held-out steps [P - check, P) to
          max_residual of the window's energy (V and I each),
      (2) the energy they predict from P to the predicted stop is at most
          max_tail_energy of the energy recorded so far, and
      (3) S11 moves by at most max_horizon_change when the predicted stop
          moves by horizon_window steps either way;
    the predicted stop is where the plain run's decay test would end it
    (predict_stop on the energies of every check so far).  Its S11 is the
    simulated DFT over [0, P) plus the closed-form DFT of the modes
    (residues refitted over [P - window, P)) from P to the predicted stop,
    the plain run's own horizon.  The other structures continue from their
    state at P with the plain run's fields, DFT sums, energy peak and check
    steps, so one that never stops ends exactly as the plain run does.
    poles="pencil" takes the poles from the matrix pencil of the same window
    instead of the network (pencil_poles), the classical comparison.  Only
    structures whose energy has already fallen min_decay_db below its peak
    are evaluated at a checkpoint (a tail of at most max_tail_energy of the
    signal energy needs most of the decay behind it), and the least-squares
    fits use every fit_stride-th step.  The structures of a checkpoint are
    assessed on `workers` host threads (default min(8, cpus)).  With
    overlap, that assessment runs while the GPU steps the next segment, in
    which the candidates take part: one accepted at P keeps the S11 formed
    from its state at P, and the segment it ran beyond P counts in its steps
    (overlap=False: it stops at P)."""
    d = model.dims
    workers = workers or min(8, os.cpu_count() or 1)
    early = window + check
    stride = stride or window
    fr = np.linspace(1e9, 20e9, 64) if freqs is None else np.asarray(freqs)
    built = build(structs, max_steps, max_steps, fr)
    dt = built[0][0].dt
    n = len(structs)
    s11 = np.zeros((n, len(fr)), np.complex128)
    accepted = np.zeros(n, bool)
    resid = np.full(n, np.nan)
    tail_frac = np.full(n, np.nan)
    steps = np.zeros(n, np.int64)
    hist_t = [[] for _ in range(n)]
    hist_e = [[] for _ in range(n)]
    live = list(range(n))
    pos = 0
    ci = None
    checkpoints = []
    wall = 0.0
    t_inf = 0.0
    pending = None  # the last checkpoint, assessed while the next segment runs

    def apply(c: _Checkpoint, outcomes):
        rejected = []
        for k, q in enumerate(c.still):
            ok, s_q, resid[q], tail_frac[q] = outcomes[k]
            if ok:
                accepted[q] = True
                s11[q] = s_q
            else:
                rejected.append(q)
        if log:
            log(f"checkpoint {c.pos}: {int(accepted.sum())} accepted")
        return rejected

    while live:
        target = min(max_steps, early if pos < early else pos + stride)
        sw = Sweep()
        for q in live:
            op, eng, cp, port, exc = built[q]
            sw.add(eng, cp, port, shifted(exc, pos))
        if ci is not None:
            sw.set_ci(ci)
        seg = {}

        def run_segment(sw=sw, n_steps=target - pos, tune=ci is None):
            s0 = time.perf_counter()
            try:
                seg["res"] = sw.run(n_steps, energy_decay_db=decay_db, check_interval=decay_check, autotune=tune)
            except BaseException as err:  # re-raised on the calling thread
                seg["err"] = err
            seg["t"] = time.perf_counter() - s0

        t0 = time.perf_counter()
        outcomes = None
        if pending is not None:
            th = threading.Thread(target=run_segment)
            th.start()
            try:
                outcomes = _assess_all(pending, workers)
            finally:
                th.join()
        else:
            run_segment()
        if "err" in seg:
            raise seg["err"]
        wall += seg["t"]
        t_inf += time.perf_counter() - t0 - seg["t"]  # host work the segment did not hide
        res = seg["res"]
        if ci is None:
            ci = sw.tuning()["ci"]
        if pending is not None:
            apply(pending, outcomes)
            pending = None
        pos = target
        still, waiting = [], []
        for k, q in enumerate(live):
            steps[q] += res[k].timesteps_executed
            ts, en = sw.energy_history(k)
            hist_t[q].append(ts)
            hist_e[q].append(en)
            if accepted[q]:
                continue  # stopped at the last checkpoint; this segment's state is not used
            if res[k].termination_reason != 0 or pos >= max_steps:
                s11[q] = built[q][3].s11(50.0)
                continue
            e_all = np.concatenate(hist_e[q])
            if len(e_all) and e_all.max() > 0 and e_all[-1] <= e_all.max() * 10.0 ** (-min_decay_db / 10.0):
                still.append(q)
            else:
                waiting.append(q)
        if not still:
            live = waiting
            if live:
                checkpoints.append(pos)
            continue
        checkpoints.append(pos)
        t1 = time.perf_counter()
        X = np.zeros((len(still), 2, d.L0), np.float32)
        sers, dfts = [], []
        for k, q in enumerate(still):
            port = built[q][3]
            ser = port.series()
            X[k], _, _ = _window(ser[:, 0], ser[:, 1], pos - early, window, decimate, 0)
            sers.append(ser)
            dfts.append(port.dft_values())
        s_all = None
        if poles == "network":
            modes = np.zeros((len(still), 6 * d.modes), np.float32)
            for b0 in range(0, len(still), d.batch):
                m, _ = model.forward(X[b0:b0 + d.batch], tails=False)
                modes[b0:b0 + d.batch] = m
            s_all, _, _ = mode_params(modes, d.modes)
        ctx = _Checkpoint(pos, list(still), sers, dfts,
                          [(np.concatenate(hist_t[q]), np.concatenate(hist_e[q])) for q in still], s_all,
                          window, check, decimate, max_steps, dt, fr, max_residual, max_tail_energy,
                          horizon_window, max_horizon_change, decay_db, decay_check, pencil_order, fit_stride)
        if overlap:
            pending = ctx
            live = still + waiting  # the candidates run on while they are assessed
        else:
            rejected = apply(ctx, _assess_all(ctx, workers))
            live = rejected + waiting
        t_inf += time.perf_counter() - t1
    return LearnedResult(s11, accepted, resid, wall + t_inf, t_inf, steps, checkpoints, tail_frac)


def refit_tail_error(model: Extrapolator, X: np.ndarray, T: np.ndarray, poles: Optional[np.ndarray] = None):
    """Tail error of modes whose residues are refitted to the window
    (refit_residues on the decimated window samples), divided by that of
    predicting zero, over the whole tail.  The poles are the network's for
    each window, or the given [M] poles for every window (for instance the
    untrained frequency-spread start).  X [B, 2, L0] and T [B, 2, tail] in
    the network's normalised units."""
    B = model.dims.batch
    err = zero = 0.0
    L0, tail = X.shape[2], T.shape[2]
    n_in = np.arange(L0)
    n_out = np.arange(L0, L0 + tail)
    for b0 in range(0, len(X), B):
        xb = X[b0:b0 + B]
        nb = len(xb)
        if poles is None:
            pad = xb if nb == B else np.concatenate([xb, np.zeros((B - nb,) + xb.shape[1:], xb.dtype)])
            m, _ = model.forward(pad, tails=False)
            s_all, _, _ = mode_params(m[:nb], model.dims.modes)
        else:
            s_all = np.broadcast_to(poles, (nb, len(poles)))
        for k in range(nb):
            for ch in range(2):
                a, _ = refit_residues(xb[k, ch].astype(np.float64), s_all[k], 0, L0, L0, 1)
                pred = modes_eval(s_all[k], a, n_out, L0, 1)
                err += float(np.sum((pred - T[b0 + k, ch]) ** 2))
                zero += float(np.sum(T[b0 + k, ch].astype(np.float64) ** 2))
    return err / max(zero, 1e-300)


def spread_poles(model: Extrapolator) -> np.ndarray:
    """The poles the output layer starts from (spread_modes): the network's
    modes for a zero input, before training."""
    d = model.dims
    m, _ = model.forward(np.zeros((d.batch, 2, d.L0), np.float32), tails=False)
    return mode_params(m[:1], d.modes)[0][0]


def dataset(run: SweepRun, window: int, decimate: int, tail: int, offsets=(0,)):
    """Training pairs from recorded runs: for every offset o, the decimated
    window [o, o + window) and the next `tail` decimated samples (past the
    end of a record they count as zero, so records should cover
    max(offsets) + window + tail * decimate steps)."""
    X, T = [], []
    for q in range(len(run.u)):
        for o in offsets:
            x, t, _ = training_pair(run.u[q][o:], run.i[q][o:], window, decimate, tail)
            X.append(x)
            T.append(t)
    return np.array(X), np.array(T)


def spread_modes(model: Extrapolator, f_lo: float = 0.004, f_hi: float = 0.2, damping: float = -3.0,
                 shrink: float = 0.1):
    """Training start for the output layer: mode m's frequency output at
    f_lo..f_hi cycles per decimated sample (evenly spaced; 0.2 is 26 GHz at
    8 steps of 0.95 ps), its damping output at `damping` (softplus(-3) *
    A_SCALE = 2.4e-3 per sample) and the layer's weights scaled by `shrink`.
    From the default start every mode sits near F_MAX / 2 and gradient
    descent shrinks the residues to zero rather than moving frequencies."""
    d = model.dims
    M = d.modes
    p = model.params()
    b5 = p[-6 * M:].reshape(M, 6)
    f = np.linspace(f_lo, f_hi, M)
    b5[:, 0] = damping
    b5[:, 1] = np.log(f / (F_MAX - f))       # F_MAX * sigmoid(o) = f
    p[-6 * M:] = b5.reshape(-1)
    p[-6 * M - 6 * M * d.hidden:-6 * M] *= shrink
    model.set_params(p)


def tail_schedule(epochs: int, full: int, first: int = 16, share: float = 0.7) -> List[int]:
    """Tail length per epoch: first, 2 first, 4 first, ... below `full` over
    the first `share` of the epochs (equal parts), then the full tail.  A
    short horizon keeps a frequency error's phase drift small, so the modes
    find their frequencies before the long tail weights their damping."""
    stages = []
    n = first
    while n < full:
        stages.append(n)
        n *= 2
    if not stages:
        return [full] * epochs
    n_cur = int(round(share * epochs))
    out = []
    for ep in range(epochs):
        out.append(stages[min(len(stages) - 1, ep * len(stages) // max(n_cur, 1))] if ep < n_cur else full)
    return out


def train(model: Extrapolator, X: np.ndarray, T: np.ndarray, epochs: int, lr: float, seed: int = 0,
          log=None, curriculum: bool = True) -> List[float]:
    """Adam with a cosine learning rate; with `curriculum`, the tail grows
    over the epochs (tail_schedule) on models of the same parameters and
    shorter tails, and the last epochs train `model` itself."""
    rng = np.random.default_rng(seed)
    d = model.dims
    B = d.batch
    losses = []
    steps = epochs * max(1, len(X) // B)
    sched = tail_schedule(epochs, d.tail) if curriculum else [d.tail] * epochs
    k = 0
    cur = model
    for ep in range(epochs):
        n = sched[ep]
        if n != cur.dims.tail:
            p = cur.params()
            cur = model if n == d.tail else Extrapolator(replace(d, tail=n), params=p, seed=seed)
            if cur is model:
                model.set_params(p)
        Tn = np.ascontiguousarray(T[:, :, :n])
        perm = rng.permutation(len(X))
        tot = 0.0
        nb = 0
        for b0 in range(0, len(X) - B + 1 if len(X) >= B else 1, B):
            idx = perm[b0:b0 + B]
            lr_k = lr * 0.5 * (1 + math.cos(math.pi * k / max(steps, 1)))
            tot += cur.train_step(X[idx], Tn[idx], max(lr_k, lr * 1e-2))
            nb += 1
            k += 1
        losses.append(tot / max(nb, 1))
        if log:
            log(f"epoch {ep}: tail {n}, loss {losses[-1]:.4e}")
    if cur is not model:
        model.set_params(cur.params())
    return losses


def tail_error(model: Extrapolator, X: np.ndarray, T: np.ndarray, horizons=(32, 128, 512)):
    """Mean squared tail error of the network over the first n decimated
    samples, divided by that of predicting zero (stopping without
    extrapolating), for each n in horizons."""
    B = model.dims.batch
    err = np.zeros(len(horizons))
    zero = np.zeros(len(horizons))
    for b0 in range(0, len(X), B):
        xb = X[b0:b0 + B]
        nb = len(xb)
        if nb < B:
            xb = np.concatenate([xb, np.zeros((B - nb,) + xb.shape[1:], xb.dtype)])
        _, tl = model.forward(xb)
        for h, n in enumerate(horizons):
            err[h] += float(((tl[:nb, :, :n] - T[b0:b0 + nb, :, :n]) ** 2).sum())
            zero[h] += float((T[b0:b0 + nb, :, :n].astype(np.float64) ** 2).sum())
    return {int(n): float(err[h] / max(zero[h], 1e-300)) for h, n in enumerate(horizons)}
