"""Pins the port / DFT restatement of the oracle (oracle/port.py and the
lumped element + signal recording of oracle/oe_oracle.cpp)."""
import math

import numpy as np
import pytest

from oracle import excitation as X
from oracle import oracle as O
from oracle import port as P

EPS0 = 8.854187817e-12


def test_dft_gaussian_closed_form():
    # 2 dt sum g(t) e^{-jwt} -> 2 * sigma sqrt(2 pi) exp(-(sigma w)^2/2) e^{-j w t0}
    fc = 15e9
    t0, tau = X.openems_gaussian_params(fc)
    dt = 2e-13
    t = np.arange(20000) * dt
    g = X.gaussian(t, t0, tau)[:, None]
    freqs = [1e9, 5e9, 12e9]
    F = P.dft(g, t, freqs)[:, 0]
    sigma = tau / math.sqrt(2)
    for q, f in enumerate(freqs):
        w = 2 * math.pi * f
        ref = 2 * sigma * math.sqrt(2 * math.pi) * math.exp(-0.5 * (sigma * w) ** 2) * np.exp(-1j * w * t0)
        assert abs(F[q] - ref) < 1e-4 * abs(ref)  # the pulse starts at t = 0 (exp(-9) cut-off)


def test_calc_port_limits():
    # matched load U = R I -> S11 = 0; short -> -1; open -> +1; general (Z-R)/(Z+R)
    R = 50.0
    I = np.array([1 + 0.5j, -2j])
    assert np.allclose(P.calc_port(R * I, I, R)["s11"], 0)
    assert np.allclose(P.calc_port(0 * I, I, R)["s11"], -1)
    assert np.allclose(P.calc_port(I, 0 * I, R)["s11"], 1)
    Z = 20 - 35j
    assert np.allclose(P.calc_port(Z * I, I, R)["s11"], (Z - R) / (Z + R))
    d = P.calc_port(Z * I, I, R)
    assert np.allclose(d["p_acc"], d["p_inc"] - d["p_ref"])


def test_lumped_element_coefficients():
    # 50 ohm over a 1x1x4 z-port: 4 edges in series, one line -> R_edge = 12.5 ohm,
    # sigma = dz / (R_edge dx dy)
    d = 0.5e-3
    s = O.OracleSim(10, 10, 12, d, 64)
    s.set_material(3.0, 1.0, 0.0, 0.0, (0, 10, 0, 10, 0, 12))
    s.add_lumped(2, (4, 5, 4, 5, 4, 8), 50.0)
    c = s.coefficients()
    dt = s.dt
    sig = d / (12.5 * d * d)
    eps = 3.0 * EPS0
    loss = sig * dt / (2 * eps)
    for k in range(4, 8):
        assert c[2, 4, 4, k] == pytest.approx((1 - loss) / (1 + loss), rel=1e-14)
        assert c[5, 4, 4, k] == pytest.approx((dt / eps) / (1 + loss) / d, rel=1e-14)
    # only the port component and cells change
    assert c[0, 4, 4, 5] == 1.0 and c[2, 4, 4, 8] == 1.0 and c[2, 4, 4, 3] == 1.0
    # two parallel lines of 2 series edges: R_edge = 50 * 2 / 2
    s2 = O.OracleSim(10, 10, 12, d, 64)
    s2.add_lumped(0, (3, 5, 4, 6, 5, 6), 50.0)
    c2 = s2.coefficients()
    sig2 = d / (50.0 * d * d)
    loss2 = sig2 * dt / (2 * EPS0)
    assert c2[0, 3, 4, 5] == pytest.approx((1 - loss2) / (1 + loss2), rel=1e-14)


def test_port_terms_geometry():
    # U: -d along the centre line; I: loop of H around each edge at the middle index
    u, cur = P.lumped_port_terms((10, 10, 12), (1e-3, 2e-3, 3e-3), 2, (4, 5, 4, 5, 4, 8))
    assert [t[:4] for t in u] == [(2, 4, 4, k) for k in range(4, 8)]
    assert all(t[4] == -3e-3 for t in u)
    # Ez curl at (4,4,6): (Hy(4) - Hy(3)) dy - (Hx(j=4) - Hx(j=3)) dx
    assert sorted(cur) == sorted([(4, 4, 4, 6, 2e-3), (4, 3, 4, 6, -2e-3), (3, 4, 4, 6, -1e-3), (3, 4, 3, 6, 1e-3)])


def test_signals_match_direct_sums():
    s = O.OracleSim(8, 8, 8, 1e-3, 64)
    rng = np.random.default_rng(0)
    f = np.zeros((6, 8, 8, 8))
    f[:3, 1:, 1:, 1:] = rng.standard_normal((3, 7, 7, 7))
    s.set_fields(f)
    terms = [[(2, 3, 3, 3, 0.5), (2, 3, 3, 4, -1.5)], [(4, 2, 3, 4, 2.0)]]
    s.add_signals(terms)
    s.run_batch(1)
    F = s.fields()
    ser = s.signal_series(0, 2)
    assert ser.shape == (1, 2)
    assert ser[0, 0] == pytest.approx(0.5 * F[2, 3, 3, 3] - 1.5 * F[2, 3, 3, 4], rel=1e-15)
    assert ser[0, 1] == pytest.approx(2.0 * F[4, 2, 3, 4], rel=1e-15)
