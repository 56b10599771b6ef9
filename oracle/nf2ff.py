"""Near-field to far-field transform, restated for the oracle (float64 numpy).

TEST INFRASTRUCTURE ONLY (see oracle/oe_oracle.cpp header).

Follows matlab/AnalyzeNF2FF.m (the reference's Rust integrate_surfaces,
src/nf2ff/mod.rs:157-170, returns zeros: a placeholder):
  outward normals n (:76-81), J = n x H, M = -n x E (:186-191),
  e_fkt = exp(+j k r_hat . r') (:192-193), N_theta, N_phi, L_theta, L_phi by
  2-D trapezoid (:194-197, dbltrapz :214), E_theta = -jk e^{-jkr}/(4 pi r)
  (L_phi + Z0 N_theta), E_phi = +jk e^{-jkr}/(4 pi r)(L_theta - Z0 N_phi)
  (:104-105), U = r^2/(2 Z0)(|E_theta|^2 + |E_phi|^2) (:112), Prad = 1/2
  dbltrapz(Re((E x H*) . n)) (:201-209), Dmax = 4 pi Umax / Prad (:125).
The phase reference point `center` is an argument (AnalyzeNF2FF uses 0); it
changes only a per-angle phase of E_theta/E_phi, not |E| or D.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np

C0 = 299792458.0
Z0 = 376.730313668


def trapz_weights(c: np.ndarray) -> np.ndarray:
    """Weights w with sum(w * f) == trapz(f, c)."""
    w = np.zeros(len(c))
    if len(c) > 1:
        d = np.diff(c)
        w[:-1] += d / 2
        w[1:] += d / 2
    return w


@dataclass
class Face:
    """One face of the NF2FF box: fields at the mesh nodes of the face.
    axis: normal axis; sign: +1/-1 outward; c0: coordinate of the face;
    c1, c2: node coordinates along the two in-plane axes (axis+1, axis+2 mod 3);
    E, H: complex arrays [3, len(c1), len(c2)] (x, y, z components)."""
    axis: int
    sign: int
    c0: float
    c1: np.ndarray
    c2: np.ndarray
    E: np.ndarray
    H: np.ndarray


def _currents(face: Face):
    n = np.zeros(3)
    n[face.axis] = face.sign
    H, E = face.H, face.E
    J = np.stack([n[1] * H[2] - n[2] * H[1], n[2] * H[0] - n[0] * H[2], n[0] * H[1] - n[1] * H[0]])
    M = np.stack([-n[1] * E[2] + n[2] * E[1], -n[2] * E[0] + n[0] * E[2], -n[0] * E[1] + n[1] * E[0]])
    return n, J, M


def _coords(face: Face, center):
    a1, a2 = (face.axis + 1) % 3, (face.axis + 2) % 3
    X = [None, None, None]
    g1, g2 = np.meshgrid(face.c1, face.c2, indexing="ij")
    X[face.axis] = np.full(g1.shape, face.c0) - center[face.axis]
    X[a1] = g1 - center[a1]
    X[a2] = g2 - center[a2]
    return X


def far_field(faces: List[Face], f: float, theta: np.ndarray, phi: np.ndarray, r: float = 1.0, center=(0, 0, 0)):
    """Returns E_theta, E_phi [n_theta, n_phi], Prad, Dmax."""
    k = 2 * math.pi * f / C0
    th = np.asarray(theta, np.float64)[:, None]
    ph = np.asarray(phi, np.float64)[None, :]
    st, ct, sp, cp = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
    Nx = Ny = Nz = Lx = Ly = Lz = 0
    for face in faces:
        n, J, M = _currents(face)
        X = _coords(face, center)
        w = trapz_weights(face.c1)[:, None] * trapz_weights(face.c2)[None, :]
        # r_hat . r' for every (theta, phi, point)
        rc = (X[0][None, None] * (st * cp)[..., None, None] + X[1][None, None] * (st * sp)[..., None, None]
              + X[2][None, None] * np.broadcast_to(ct, st.shape)[..., None, None])
        e = np.exp(1j * k * rc) * w[None, None]
        Nx = Nx + np.einsum("abij,ij->ab", e, J[0])
        Ny = Ny + np.einsum("abij,ij->ab", e, J[1])
        Nz = Nz + np.einsum("abij,ij->ab", e, J[2])
        Lx = Lx + np.einsum("abij,ij->ab", e, M[0])
        Ly = Ly + np.einsum("abij,ij->ab", e, M[1])
        Lz = Lz + np.einsum("abij,ij->ab", e, M[2])
    N_t = Nx * ct * cp + Ny * ct * sp - Nz * st
    N_p = -Nx * sp + Ny * cp
    L_t = Lx * ct * cp + Ly * ct * sp - Lz * st
    L_p = -Lx * sp + Ly * cp
    fac = 1j * k * np.exp(-1j * k * r) / (4 * math.pi * r)
    E_t = -fac * (L_p + Z0 * N_t)
    E_p = fac * (L_t - Z0 * N_p)
    U = r * r / (2 * Z0) * (np.abs(E_t) ** 2 + np.abs(E_p) ** 2)
    prad = radiated_power(faces)
    return E_t, E_p, prad, 4 * math.pi * U.max() / prad


def radiated_power(faces: List[Face]) -> float:
    P = 0.0
    for face in faces:
        n, _, _ = _currents(face)
        E, H = face.E, face.H
        S = np.cross(E, np.conj(H), axis=0)
        w = trapz_weights(face.c1)[:, None] * trapz_weights(face.c2)[None, :]
        P += 0.5 * np.sum(np.real(np.tensordot(n, S, axes=(0, 0))) * w)
    return float(P)


def hertzian_dipole_fields(p, f, X, Y, Z):
    """Exact E, H (complex phasors, e^{+jwt}) of a z-directed Hertzian dipole
    of moment p = I dl at the origin, at points (X, Y, Z)."""
    k = 2 * math.pi * f / C0
    R = np.sqrt(X * X + Y * Y + Z * Z)
    ct = Z / R
    st = np.sqrt(np.maximum(0.0, 1 - ct * ct))
    rho = np.sqrt(X * X + Y * Y)
    cp = np.where(rho > 0, X / np.where(rho > 0, rho, 1), 1.0)
    sp = np.where(rho > 0, Y / np.where(rho > 0, rho, 1), 0.0)
    g = np.exp(-1j * k * R)
    Er = Z0 * p * ct / (2 * math.pi * R * R) * (1 + 1 / (1j * k * R)) * g
    Et = 1j * Z0 * k * p * st / (4 * math.pi * R) * (1 + 1 / (1j * k * R) - 1 / (k * R) ** 2) * g
    Hp = 1j * k * p * st / (4 * math.pi * R) * (1 + 1 / (1j * k * R)) * g
    # spherical -> cartesian
    Ex = Er * st * cp + Et * ct * cp
    Ey = Er * st * sp + Et * ct * sp
    Ez = Er * ct - Et * st
    Hx = -Hp * sp
    Hy = Hp * cp
    Hz = 0 * Hp
    return np.stack([Ex, Ey, Ez]), np.stack([Hx, Hy, Hz])


def box_faces(lo, hi, steps, field_fn):
    """Faces of the box [lo, hi] (3-vectors, metres) sampled on a uniform node
    grid with `steps` nodes per axis; field_fn(X, Y, Z) -> (E[3,...], H[3,...])."""
    axes = [np.linspace(lo[a], hi[a], steps[a]) for a in range(3)]
    faces = []
    for axis in range(3):
        a1, a2 = (axis + 1) % 3, (axis + 2) % 3
        for sign, c0 in ((-1, lo[axis]), (1, hi[axis])):
            g1, g2 = np.meshgrid(axes[a1], axes[a2], indexing="ij")
            P = [None] * 3
            P[axis] = np.full(g1.shape, c0)
            P[a1], P[a2] = g1, g2
            E, H = field_fn(P[0], P[1], P[2])
            faces.append(Face(axis, sign, c0, axes[a1], axes[a2], E, H))
    return faces


def _node_terms(ax, q, a1, a2, i1, i2):
    """The four signals (E_a1, E_a2, H_a1, H_a2) at one node of a face."""
    p = [0, 0, 0]
    p[ax], p[a1], p[a2] = q, i1, i2
    out = []
    for c in (a1, a2):
        sig = []
        for d in (-1, 0):
            t = list(p)
            t[c] += d
            sig.append((c, *t, 0.5))
        out.append(sig)
    for c in (a1, a2):
        o = a2 if c == a1 else a1
        sig = []
        for dn in (-1, 0):
            for dt in (-1, 0):
                t = list(p)
                t[ax] += dn
                t[o] += dt
                sig.append((3 + c, *t, 0.25))
        out.append(sig)
    return out


def surface_faces(box, directions: int = 0x3F):
    """Faces of the recording box, (axis, sign, q_axis, c1 node indices, c2
    node indices), in the order of surface_terms."""
    lo = box[0::2]
    hi = box[1::2]
    faces = []
    for ax in range(3):
        for side in range(2):
            if not directions >> (2 * ax + side) & 1:
                continue
            a1, a2 = (ax + 1) % 3, (ax + 2) % 3
            faces.append((ax, 1 if side else -1, hi[ax] if side else lo[ax], list(range(lo[a1], hi[a1] + 1)),
                          list(range(lo[a2], hi[a2] + 1))))
    return faces


def surface_terms(shape, delta, box, directions: int = 0x3F):
    """DFT signal terms of an NF2FF recording box through the mesh nodes
    box = (i0, i1, j0, j1, k0, k1) (inclusive), restating openEMS's
    interpolated field dump (python/openEMS/nf2ff.py:62-72, dump_mode 1):
    at a node, a tangential E component is the mean of the two edges of that
    component meeting there (E_c lives at half-integer positions along c), a
    tangential H component the mean of the four faces around it (H_c lives at
    half-integer positions along the two axes other than c).  Faces in the
    order (axis 0..2) x (lower, upper), nodes [c1][c2] over the in-plane axes
    (axis+1)%3, (axis+2)%3, per node the signals E_a1, E_a2, H_a1, H_a2.
    Terms are (component 0..5, i, j, k, weight).  Returns (terms, faces) where
    faces are (axis, sign, q_axis, c1 node indices, c2 node indices)."""
    faces = surface_faces(box, directions)
    terms = []
    for ax, sign, q, r1, r2 in faces:
        a1, a2 = (ax + 1) % 3, (ax + 2) % 3
        for i1 in r1:
            for i2 in r2:
                terms.extend(_node_terms(ax, q, a1, a2, i1, i2))
    return terms, faces


def surface_terms_subset(box, signals, directions: int = 0x3F):
    """The terms of the given signal indices (positions in surface_terms'
    order) only, for boxes too large to list every signal."""
    faces = surface_faces(box, directions)
    starts = np.cumsum([0] + [4 * len(r1) * len(r2) for _, _, _, r1, r2 in faces])
    out = []
    for s in signals:
        f = int(np.searchsorted(starts, s, side="right")) - 1
        ax, _, q, r1, r2 = faces[f]
        node, which = divmod(int(s) - int(starts[f]), 4)
        n1, n2 = divmod(node, len(r2))
        out.append(_node_terms(ax, q, (ax + 1) % 3, (ax + 2) % 3, r1[n1], r2[n2])[which])
    return out


def faces_from_phasors(values, delta, faces):
    """Face list for far_field from the phasors of surface_terms' signals
    (values: complex [n_signals] in that order).  Node coordinates are
    index * delta."""
    out, m = [], 0
    for ax, sign, q, r1, r2 in faces:
        a1, a2 = (ax + 1) % 3, (ax + 2) % 3
        n = len(r1) * len(r2)
        v = np.asarray(values[m:m + 4 * n]).reshape(len(r1), len(r2), 4)
        m += 4 * n
        E = np.zeros((3, len(r1), len(r2)), np.complex128)
        H = np.zeros_like(E)
        E[a1], E[a2], H[a1], H[a2] = v[..., 0], v[..., 1], v[..., 2], v[..., 3]
        out.append(Face(ax, sign, q * delta[ax], np.asarray(r1) * delta[a1], np.asarray(r2) * delta[a2], E, H))
    return out
