"""Pins the NF2FF restatement of the oracle (oracle/nf2ff.py).

- the Hertzian dipole: closed-form radiated power, sin(theta) pattern,
  far-field magnitude and D_max = 1.5;
- a brute-force per-angle, per-point sum with explicit cross products on a
  tiny box of random phasors (catches a wrong sign or component in the
  theta/phi projections of far_field);
- the phase-centre rule E(c) = E(0) exp(-j k r_hat . c);
- surface_terms: on fields linear in position, sampled at each component's
  Yee position, every node signal equals the linear function at the node
  (a wrong edge or face offset shifts it by half a cell).
"""
import math

import numpy as np

from oracle import nf2ff as N


def _dipole_faces(f, steps=(41, 41, 41), p=1e-3):
    lam = N.C0 / f
    lo, hi = [-0.5 * lam] * 3, [0.5 * lam] * 3
    return N.box_faces(lo, hi, steps, lambda X, Y, Z: N.hertzian_dipole_fields(p, f, X, Y, Z))


def test_hertzian_dipole_power_pattern_directivity():
    f, p = 3e9, 1e-3
    faces = _dipole_faces(f, p=p)
    th = np.linspace(0, math.pi, 37)
    ph = np.linspace(0, 2 * math.pi, 73)[:-1]
    Et, Ep, prad, dmax = N.far_field(faces, f, th, ph)
    k = 2 * math.pi * f / N.C0
    exact = N.Z0 * (k * p) ** 2 / (12 * math.pi)
    assert abs(prad / exact - 1) < 2e-3
    assert abs(dmax - 1.5) < 5e-3
    pat = np.abs(Et[:, 0])
    assert np.max(np.abs(pat / pat.max() - np.sin(th))) < 2e-3
    assert np.abs(Ep).max() < 1e-3 * np.abs(Et).max()
    assert abs(np.abs(Et).max() / (N.Z0 * k * p / (4 * math.pi)) - 1) < 2e-3
    # rotational symmetry about z, up to the box's own discretisation
    assert np.max(np.abs(np.abs(Et) - np.abs(Et[:, :1]))) < 2e-3 * np.abs(Et).max()


def _brute_force(faces, f, theta, phi, r):
    k = 2 * math.pi * f / N.C0
    Et = np.zeros((len(theta), len(phi)), np.complex128)
    Ep = np.zeros_like(Et)
    for it, t in enumerate(theta):
        for ip, q in enumerate(phi):
            rh = np.array([math.sin(t) * math.cos(q), math.sin(t) * math.sin(q), math.cos(t)])
            th_h = np.array([math.cos(t) * math.cos(q), math.cos(t) * math.sin(q), -math.sin(t)])
            ph_h = np.array([-math.sin(q), math.cos(q), 0.0])
            Nv = np.zeros(3, np.complex128)
            Lv = np.zeros(3, np.complex128)
            for fc in faces:
                n = np.zeros(3)
                n[fc.axis] = fc.sign
                a1, a2 = (fc.axis + 1) % 3, (fc.axis + 2) % 3
                for i1, x1 in enumerate(fc.c1):
                    for i2, x2 in enumerate(fc.c2):
                        w1 = ((fc.c1[i1] - fc.c1[i1 - 1]) if i1 > 0 else 0) + ((fc.c1[i1 + 1] - x1) if i1 + 1 < len(fc.c1) else 0)
                        w2 = ((fc.c2[i2] - fc.c2[i2 - 1]) if i2 > 0 else 0) + ((fc.c2[i2 + 1] - x2) if i2 + 1 < len(fc.c2) else 0)
                        pos = np.zeros(3)
                        pos[fc.axis], pos[a1], pos[a2] = fc.c0, x1, x2
                        E = fc.E[:, i1, i2]
                        H = fc.H[:, i1, i2]
                        e = np.exp(1j * k * rh.dot(pos)) * 0.25 * w1 * w2
                        Nv += np.cross(n, H) * e
                        Lv += -np.cross(n, E) * e
            fac = 1j * k * np.exp(-1j * k * r) / (4 * math.pi * r)
            Et[it, ip] = -fac * (Lv.dot(ph_h) + N.Z0 * Nv.dot(th_h))
            Ep[it, ip] = fac * (Lv.dot(th_h) - N.Z0 * Nv.dot(ph_h))
    return Et, Ep


def _random_faces(rng, lo, hi, steps):
    axes = [np.sort(rng.uniform(lo[a], hi[a], steps[a])) for a in range(3)]
    for a in range(3):
        axes[a][0], axes[a][-1] = lo[a], hi[a]
    faces = []
    for axis in range(3):
        a1, a2 = (axis + 1) % 3, (axis + 2) % 3
        for sign, c0 in ((-1, lo[axis]), (1, hi[axis])):
            shape = (3, steps[a1], steps[a2])
            E = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
            H = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / N.Z0
            faces.append(N.Face(axis, sign, c0, axes[a1], axes[a2], E, H))
    return faces


def test_far_field_matches_brute_force_sum():
    rng = np.random.default_rng(3)
    faces = _random_faces(rng, [-0.02, -0.03, -0.015], [0.025, 0.02, 0.03], (4, 5, 3))
    f = 5e9
    th = np.array([0.0, 0.4, 1.3, 2.9])
    ph = np.array([0.0, 1.1, 3.7])
    Et, Ep, _, _ = N.far_field(faces, f, th, ph, r=2.5)
    Bt, Bp = _brute_force(faces, f, th, ph, 2.5)
    assert np.allclose(Et, Bt, rtol=1e-12, atol=1e-12 * np.abs(Bt).max())
    assert np.allclose(Ep, Bp, rtol=1e-12, atol=1e-12 * np.abs(Bp).max())


def test_phase_centre_shift():
    rng = np.random.default_rng(4)
    faces = _random_faces(rng, [-0.02] * 3, [0.02] * 3, (5, 4, 6))
    f = 7e9
    k = 2 * math.pi * f / N.C0
    th = np.linspace(0.1, 3.0, 5)
    ph = np.linspace(0, 6, 4)
    c = np.array([0.003, -0.007, 0.011])
    E0t, E0p, P0, D0 = N.far_field(faces, f, th, ph)
    E1t, E1p, P1, D1 = N.far_field(faces, f, th, ph, center=c)
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    rc = np.sin(T) * np.cos(Ph) * c[0] + np.sin(T) * np.sin(Ph) * c[1] + np.cos(T) * c[2]
    ph_f = np.exp(-1j * k * rc)
    assert np.allclose(E1t, E0t * ph_f, rtol=1e-11, atol=1e-12 * np.abs(E0t).max())
    assert np.allclose(E1p, E0p * ph_f, rtol=1e-11, atol=1e-12 * np.abs(E0p).max())
    assert P1 == P0 and abs(D1 - D0) < 1e-12 * abs(D0)


def test_radiated_power_of_plane_wave_slab():
    # a +z plane wave through the top face only: P = |E|^2 / (2 Z0) * area
    c1 = np.linspace(0, 0.02, 5)
    c2 = np.linspace(0, 0.03, 7)
    E = np.zeros((3, 5, 7), np.complex128)
    H = np.zeros_like(E)
    E[0] = 2.0
    H[1] = 2.0 / N.Z0
    P = N.radiated_power([N.Face(2, 1, 0.0, c1, c2, E, H)])
    assert abs(P - 4.0 / (2 * N.Z0) * 0.02 * 0.03) < 1e-15
    assert abs(N.radiated_power([N.Face(2, -1, 0.0, c1, c2, E, H)]) + P) < 1e-15


def test_surface_terms_interpolate_linear_fields_to_nodes():
    n = (9, 10, 11)
    d = (1.0, 2.0, 0.5)
    rng = np.random.default_rng(5)
    alpha = rng.standard_normal((6, 3))
    beta = rng.standard_normal(6)
    F = np.zeros((6,) + n)
    I, J, K = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    idx = [I, J, K]
    for c in range(6):
        pos = [idx[a].astype(float) for a in range(3)]
        for a in range(3):
            # E_c sits half a cell along c; H_c half a cell along the two other axes
            if (c < 3 and a == c) or (c >= 3 and a != c - 3):
                pos[a] = pos[a] + 0.5
        F[c] = sum(alpha[c, a] * pos[a] * d[a] for a in range(3)) + beta[c]
    box = (2, 6, 1, 8, 3, 9)
    terms, faces = N.surface_terms(n, d, box)
    vals = np.array([sum(w * F[c, i, j, k] for (c, i, j, k, w) in sig) for sig in terms])
    m = 0
    for ax, sign, q, r1, r2 in faces:
        a1, a2 = (ax + 1) % 3, (ax + 2) % 3
        for i1 in r1:
            for i2 in r2:
                node = [0, 0, 0]
                node[ax], node[a1], node[a2] = q, i1, i2
                x = np.array(node, float) * d
                for c in (a1, a2, 3 + a1, 3 + a2):
                    assert abs(vals[m] - (alpha[c] @ x + beta[c])) < 1e-12
                    m += 1
    assert m == len(terms)
    # 4 signals per node, every face of the closed box
    assert len(terms) == 4 * 2 * (8 * 7 + 7 * 5 + 5 * 8)
    built = N.faces_from_phasors(vals.astype(np.complex128), d, faces)
    assert built[0].axis == 0 and built[0].sign == -1 and built[0].c0 == 2.0
    assert np.array_equal(built[5].c1, np.arange(2, 7) * 1.0)


def test_surface_terms_subset_equals_full_list():
    box = (3, 9, 2, 8, 4, 11)
    terms, faces = N.surface_terms((12, 12, 14), (1e-3,) * 3, box)
    assert [f[:3] for f in N.surface_faces(box)] == [f[:3] for f in faces]
    idx = np.random.default_rng(0).choice(len(terms), 60, replace=False)
    idx = np.concatenate([idx, [0, len(terms) - 1]])
    assert N.surface_terms_subset(box, idx) == [terms[i] for i in idx]
