"""The composite workload (CPML on five faces, Mur on +z, lumped port, Drude
and Lorentz spheres in the physical ADE mode, port running DFT, NF2FF box)
at a reduced size, against the oracle running the same extensions.

Bars: fields and port/surface signal series bit-identical; FP64 running DFTs
within 1e-11 of a float64 DFT of the oracle's series; S11 within 1e-9 of
CalcPort on those spectra; far field within rel L2 1e-5 of the float64
transform of the same phasors.
"""
import math

import numpy as np
import pytest

from openems_amd import EngineBatch
from openems_amd import workloads as W
from openems_amd.engine import port_excitations
from oracle import nf2ff as N
from oracle import port as P
from tests.parity_util import port_spheres_oracle, say

pytestmark = pytest.mark.gpu


def _oracle(wl, samples):
    surf_terms, faces = N.surface_terms(wl.shape, (wl.delta,) * 3, wl.nf2ff_box)
    sim, port_terms = port_spheres_oracle(wl, samples, surf_terms)
    return sim, port_terms, surf_terms, faces


def test_composite_workload_against_oracle():
    wl = W.port_spheres(shape=(56, 56, 48), steps=400, n_spheres=2, rmin=3, rmax=5)
    say(f"[composite] {wl.shape}, {len(wl.spheres)} spheres, port {wl.port_region}, nf2ff {wl.nf2ff_box}, "
        f"{wl.steps} steps")
    op, eng, exts = W.build_port_spheres_hip(wl, precision=0, record_steps=wl.steps)
    smp = wl.excitation_samples(op.dt, wl.steps)
    eng.run_batch(EngineBatch(num_steps=wl.steps, excitations=port_excitations(2, wl.port_region, smp),
                              extensions=list(exts.values())))
    sim, port_terms, surf_terms, faces = _oracle(wl, smp)
    E, H = eng.read_fields()
    F = sim.fields()
    diff = np.abs(np.concatenate([E, H]) - F).max()
    say(f"[composite] fields max |gpu - oracle| = {diff:.3e} (scale {np.abs(F).max():.3e})")
    assert np.array_equal(np.concatenate([E, H]), F)
    # port series and spectra
    ser = sim.signal_series(0, 2 + len(surf_terms))
    assert np.array_equal(exts["port"].series(), ser[:, :2])
    st = np.arange(wl.steps)
    U = P.dft(ser[:, :1], (st + 1) * op.dt, wl.dft_freqs)[:, 0]
    I = P.dft(ser[:, 1:2], (st + 0.5) * op.dt, wl.dft_freqs)[:, 0]
    got = exts["port"].dft_values()
    assert np.allclose(got[:, 0], U, rtol=1e-11, atol=1e-11 * np.abs(U).max())
    assert np.allclose(got[:, 1], I, rtol=1e-11, atol=1e-11 * np.abs(I).max())
    s11 = exts["port"].s11(50.0)
    assert np.allclose(s11, P.calc_port(U, I, 50.0)["s11"], rtol=1e-9, atol=1e-9)
    # NF2FF surface phasors and far field
    surf = ser[:, 2:]
    is_h = np.array([sig[0][0] >= 3 for sig in surf_terms])
    ref = np.empty((len(wl.nf2ff_freqs), len(surf_terms)), np.complex128)
    ref[:, ~is_h] = P.dft(surf[:, ~is_h], (st + 1) * op.dt, wl.nf2ff_freqs)
    ref[:, is_h] = P.dft(surf[:, is_h], (st + 0.5) * op.dt, wl.nf2ff_freqs)
    ph = exts["nf2ff"].dft_values()
    err = np.abs(ph - ref).max() / np.abs(ref).max()
    say(f"[composite] nf2ff phasors: max rel {err:.2e} over {len(surf_terms)} signals")
    assert err < 1e-11
    th = np.linspace(0, math.pi, 19)
    phi = np.linspace(0, 2 * math.pi, 37)[:-1]
    for q, f in enumerate(wl.nf2ff_freqs):
        g = exts["nf2ff"].far_field(q, th, phi)
        Et, Ep, prad, dmax = N.far_field(N.faces_from_phasors(ph[q], (wl.delta,) * 3, faces), f, th, phi)
        e = math.sqrt((np.linalg.norm(g.e_theta - Et) ** 2 + np.linalg.norm(g.e_phi - Ep) ** 2)
                      / (np.linalg.norm(Et) ** 2 + np.linalg.norm(Ep) ** 2))
        say(f"[composite] {f / 1e9:.0f} GHz far field rel L2 {e:.2e}, P_rad {g.prad:.4e}, D_max {g.dmax:.3f}")
        assert e <= 1e-5
        assert abs(g.prad - prad) <= 1e-12 * abs(prad)
    # the FP32-Kahan port spectra of the same run
    op2, eng2, exts2 = W.build_port_spheres_hip(wl, precision=1)
    eng2.run_batch(EngineBatch(num_steps=wl.steps, excitations=port_excitations(2, wl.port_region, smp),
                               extensions=list(exts2.values())))
    g32 = exts2["port"].dft_values()
    e32 = np.abs(g32 - got).max(axis=0) / np.abs(got).max(axis=0)
    say(f"[composite] port spectra fp32-kahan vs fp64: U {e32[0]:.2e}, I {e32[1]:.2e}; "
        f"|dS11| max {np.abs(exts2['port'].s11(50.0) - s11).max():.2e}")
    assert e32.max() < 1e-5
