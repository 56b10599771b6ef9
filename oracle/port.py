"""Lumped port, running DFT and S-parameters, restated for the oracle.

TEST INFRASTRUCTURE ONLY (see oracle/oe_oracle.cpp header).

Sources (reference root):
  LumpedPort geometry, excitation and probes  python/openEMS/ports.py:155-209
  CalcPort (incident/reflected waves)         python/openEMS/ports.py:115-152
  DFT_time2freq ('pulse': sum * dt * 2)       python/openEMS/utilities.py:21-35
The reference's own LumpedRlc (src/extensions/lumped_rlc.rs:165-313) has no
port excitation, probes or S-parameters, so openEMS is followed.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np


def lumped_port_terms(shape, delta, direction: int, region, sign: int = 1):
    """U and I signals of a lumped port over region [i0,i1,j0,j1,k0,k1) with
    edges along `direction`.

    U: the port voltage, -sum(E_dir * d_dir) along the centre line of the
       cross-section (openEMS voltage probe with weight -1; for sign = -1 the
       line runs the other way).
    I: the current through the mid-plane of the port, the loop integral of H
       around the dir-edges of the cross-section at the middle edge index
       (openEMS current probe, weight = sign).
    Terms are (component 0..5, i, j, k, weight)."""
    lo = [region[0], region[2], region[4]]
    hi = [region[1], region[3], region[5]]
    a, b = (direction + 1) % 3, (direction + 2) % 3
    ca = lo[a] + (hi[a] - lo[a] - 1) // 2
    cb = lo[b] + (hi[b] - lo[b] - 1) // 2
    u = []
    for q in range(lo[direction], hi[direction]):
        pos = [0, 0, 0]
        pos[direction], pos[a], pos[b] = q, ca, cb
        u.append((direction, pos[0], pos[1], pos[2], -sign * delta[direction]))
    mid = lo[direction] + (hi[direction] - lo[direction]) // 2
    # curl_dir H through edge (.., q_a, q_b) = (H_b(q_a) - H_b(q_a - 1)) d_b - (H_a(q_b) - H_a(q_b - 1)) d_a
    cur = []
    for qa in range(lo[a], hi[a]):
        for qb in range(lo[b], hi[b]):
            def at(da, db):
                pos = [0, 0, 0]
                pos[direction], pos[a], pos[b] = mid, qa + da, qb + db
                return tuple(pos)
            cur.append((3 + b, *at(0, 0), sign * delta[b]))
            if qa - 1 >= 0:
                cur.append((3 + b, *at(-1, 0), -sign * delta[b]))
            cur.append((3 + a, *at(0, 0), -sign * delta[a]))
            if qb - 1 >= 0:
                cur.append((3 + a, *at(0, -1), sign * delta[a]))
    return [u, cur]


def port_excitations(direction: int, region, sign: int = 1, amplitude: float = 1.0):
    """Soft E sources of the port's excitation box: every dir-edge in the box
    gets -sign * amplitude * signal (python/openEMS/ports.py:183-188)."""
    out = []
    for i in range(region[0], region[1]):
        for j in range(region[2], region[3]):
            for k in range(region[4], region[5]):
                out.append(((i, j, k), direction, -sign * amplitude))
    return out


def dft(series: np.ndarray, t: np.ndarray, freqs: Sequence[float]) -> np.ndarray:
    """Running-DFT definition, X(f) = 2 dt sum_n x_n exp(-j 2 pi f t_n)
    ('pulse' signals, single-sided; python/openEMS/utilities.py:21-35).
    series: [steps, m] -> [n_f, m] complex."""
    dt = t[1] - t[0] if len(t) > 1 else 1.0
    out = np.empty((len(freqs), series.shape[1]), np.complex128)
    for q, f in enumerate(freqs):
        ph = np.exp(-2j * math.pi * f * t)
        out[q] = 2.0 * dt * (ph[:, None] * series).sum(0)
    return out


def calc_port(uf_tot, if_tot, z_ref):
    """CalcPort (python/openEMS/ports.py:136-149) -> dict with s11 etc."""
    uf_inc = 0.5 * (uf_tot + if_tot * z_ref)
    if_inc = 0.5 * (if_tot + uf_tot / z_ref)
    uf_ref = uf_tot - uf_inc
    if_ref = if_inc - if_tot
    return dict(uf_inc=uf_inc, if_inc=if_inc, uf_ref=uf_ref, if_ref=if_ref, s11=uf_ref / uf_inc,
                p_inc=0.5 * np.real(uf_inc * np.conj(if_inc)), p_ref=0.5 * np.real(uf_ref * np.conj(if_ref)),
                p_acc=0.5 * np.real(uf_tot * np.conj(if_tot)))
