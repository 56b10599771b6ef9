"""The workloads at their full sizes, through the C ABI, against the oracle,
with the error against the float64 oracle reported for every result.

- cpml_boxes: 256^3, 0.5 mm cells, 8-cell CPML (m = 3, R0 = 1e-6), 200
  seeded dielectric boxes, openEMS Gaussian (fc = 15 GHz) soft Ez at the
  centre, the whole 4000 steps, physical and compat CPML: bit-identical to
  the float32 oracle; fields and probe against the float64 oracle reported.
- port_spheres: 384 x 384 x 256 with 8 Drude and 8 Lorentz spheres, CPML on
  five faces, Mur on +z, the 50-ohm port, a 64-frequency port DFT (float64
  and float32-Kahan) and the NF2FF box, 400 steps: fields and port series
  bit-identical, spectra and S11 against the oracle's, a seeded sample of the
  2.2 M surface phasors against the oracle's DFT, the far field on the full
  181 x 360 grid against the float64 transform on the theta/phi rows it can
  afford, and the float64 oracle's fields and port series reported.
- microstrip sweep: all 256 seeded structures to -50 dB or 20000 steps in one
  sweep: every field and S11 finite (|S11| reported), the smallest, median and
  largest structure bit-identical to single engines, the smallest against the
  oracle.
- extrapolator: loss and gradients of the full-size network (1.87 M
  parameters) on one batch against the float64 oracle.
"""
import math
import time

import numpy as np
import pytest

from openems_amd import MODE_COMPAT, MODE_PHYSICAL, EngineBatch, Sweep
from openems_amd import workloads as W
from openems_amd.engine import port_excitations
from oracle import excitation as X
from oracle import nf2ff as N
from oracle import oracle as O
from oracle import port as P
from tests.parity_util import Scenario, assert_parity, port_spheres_oracle, rel_l2, run_pair, say

pytestmark = pytest.mark.gpu

FP64_FIELD_TOL = 1e-3  # float32 engine against the float64 oracle (reported; a loose sanity bound)


def _cpml_boxes_scenario(mode):
    wl = W.cpml_boxes()
    assert wl.shape == (256, 256, 256) and wl.steps == 4000 and len(wl.boxes) == 200
    t0, tau = wl.gaussian_t0_tau()
    return Scenario(f"cpml_boxes256_{wl.steps}_{'phys' if mode == MODE_PHYSICAL else 'compat'}", wl.shape, wl.delta,
                    wl.steps, materials=wl.boxes, cpml=dict(layers=(8,) * 6, mode=mode),
                    excitations=[(wl.source, 2, True, lambda t: X.gaussian(t, t0, tau))], probes=[wl.probe])


@pytest.mark.timeout(600)
def test_cpml_boxes_full_run_physical_and_fp64_error():
    sc = _cpml_boxes_scenario(MODE_PHYSICAL)
    t = time.perf_counter()
    res = run_pair(sc)
    assert_parity(res, sc.name)
    g, gp = res["gpu"], res["gpu_probes"][0]
    say(f"[config] {sc.name}: bit-identical to the float32 oracle ({time.perf_counter() - t:.0f} s)")
    del res
    res64 = run_pair(sc, prec=64)
    e = rel_l2(g, res64["oracle"])
    ep = rel_l2(gp, res64["oracle_probes"][0])
    say(f"[config] {sc.name}: rel L2 vs float64 oracle: fields {e:.3e}, probe {ep:.3e}")
    assert e < FP64_FIELD_TOL and ep < FP64_FIELD_TOL
    # the pulse has crossed the grid: the CPML and the boxes took part
    assert np.abs(gp[:, 2]).max() > 0 and np.linalg.norm(g) > 0


@pytest.mark.timeout(600)
def test_cpml_boxes_full_run_compat():
    sc = _cpml_boxes_scenario(MODE_COMPAT)
    res = run_pair(sc)
    assert_parity(res, sc.name)
    cp, sim = res["exts"][0], res["sim"]
    for face in (0, 5):
        for field in ("E", "H"):
            assert np.array_equal(cp.psi(face, field, "pos"), sim.cpml_array(face, field, "pos").astype(np.float32))


@pytest.mark.timeout(900)
def test_port_spheres_full_size():
    wl = W.port_spheres(steps=400)
    assert wl.shape == (384, 384, 256) and len(wl.spheres) == 16
    say(f"[config] port_spheres {wl.shape}, {len(wl.spheres)} spheres, port {wl.port_region}, "
        f"nf2ff {wl.nf2ff_box}, {wl.steps} steps")
    op, eng, exts = W.build_port_spheres_hip(wl, precision=0, record_steps=wl.steps)
    smp = wl.excitation_samples(op.dt, wl.steps)
    exc = port_excitations(2, wl.port_region, smp)
    eng.run_batch(EngineBatch(num_steps=wl.steps, excitations=exc, extensions=list(exts.values())))
    E, H = eng.read_fields()
    G = np.concatenate([E, H])
    del E, H
    assert np.all(np.isfinite(G))
    port_ser = exts["port"].series()
    port_dft = exts["port"].dft_values()
    s11 = exts["port"].s11(50.0)
    ph = exts["nf2ff"].dft_values()
    faces = N.surface_faces(wl.nf2ff_box)
    n_sig = ph.shape[1]
    assert n_sig == 4 * sum(len(r1) * len(r2) for _, _, _, r1, r2 in faces)
    sample = np.sort(np.random.default_rng(7).choice(n_sig, 4000, replace=False))
    surf_terms = N.surface_terms_subset(wl.nf2ff_box, sample)

    t = time.perf_counter()
    sim, _ = port_spheres_oracle(wl, smp, surf_terms)
    F = sim.fields()
    diff = int(np.count_nonzero(G != F))
    say(f"[config] port_spheres fields: {diff} differing of {G.size} (oracle {time.perf_counter() - t:.0f} s)")
    assert diff == 0
    del F
    ser = sim.signal_series(0, 2 + len(surf_terms))
    assert np.array_equal(port_ser, ser[:, :2])
    st = np.arange(wl.steps)
    U = P.dft(ser[:, :1], (st + 1) * op.dt, wl.dft_freqs)[:, 0]
    I = P.dft(ser[:, 1:2], (st + 0.5) * op.dt, wl.dft_freqs)[:, 0]
    assert np.allclose(port_dft[:, 0], U, rtol=1e-11, atol=1e-11 * np.abs(U).max())
    assert np.allclose(port_dft[:, 1], I, rtol=1e-11, atol=1e-11 * np.abs(I).max())
    assert np.allclose(s11, P.calc_port(U, I, 50.0)["s11"], rtol=1e-9, atol=1e-9)
    surf = ser[:, 2:]
    is_h = np.array([sig[0][0] >= 3 for sig in surf_terms])
    ref = np.empty((len(wl.nf2ff_freqs), len(surf_terms)), np.complex128)
    ref[:, ~is_h] = P.dft(surf[:, ~is_h], (st + 1) * op.dt, wl.nf2ff_freqs)
    ref[:, is_h] = P.dft(surf[:, is_h], (st + 0.5) * op.dt, wl.nf2ff_freqs)
    err = np.abs(ph[:, sample] - ref).max() / np.abs(ref).max()
    say(f"[config] nf2ff phasors: max rel {err:.2e} over a sample of {len(sample)} of {n_sig} signals")
    assert err < 1e-11
    del sim, ser, surf

    # far field: the full 181 x 360 grid on the device, the float64 transform
    # of the same phasors on every 30th theta row and every 10th phi column
    th = np.linspace(0, math.pi, 181)
    phi = np.linspace(0, 2 * math.pi, 361)[:-1]
    ti, pj = np.arange(0, 181, 30), np.arange(0, 360, 10)
    for q, f in enumerate(wl.nf2ff_freqs):
        g = exts["nf2ff"].far_field(q, th, phi)
        fl = N.faces_from_phasors(ph[q], (wl.delta,) * 3, faces)
        Et = np.empty((len(ti), len(pj)), np.complex128)
        Ep = np.empty_like(Et)
        for r, i in enumerate(ti):
            Et[r], Ep[r], prad, _ = N.far_field(fl, f, th[i:i + 1], phi[pj])
        gt, gph = g.e_theta[np.ix_(ti, pj)], g.e_phi[np.ix_(ti, pj)]
        e = math.sqrt((np.linalg.norm(gt - Et) ** 2 + np.linalg.norm(gph - Ep) ** 2)
                      / (np.linalg.norm(Et) ** 2 + np.linalg.norm(Ep) ** 2))
        say(f"[config] {f / 1e9:.0f} GHz far field ({len(th)}x{len(phi)} on the device): rel L2 {e:.2e} on "
            f"{Et.size} angles, P_rad {g.prad:.4e} (float64 {prad:.4e}), D_max {g.dmax:.3f}")
        assert e <= 1e-5
        assert abs(g.prad - prad) <= 1e-12 * abs(prad)
        assert np.all(np.isfinite(g.e_theta)) and np.all(np.isfinite(g.e_phi))

    # float32-Kahan port spectra and surface phasors of the same run
    op2, eng2, exts2 = W.build_port_spheres_hip(wl, precision=1)
    eng2.run_batch(EngineBatch(num_steps=wl.steps, excitations=exc, extensions=list(exts2.values())))
    g32 = exts2["port"].dft_values()
    e32 = np.abs(g32 - port_dft).max(axis=0) / np.abs(port_dft).max(axis=0)
    ph32 = exts2["nf2ff"].dft_values()
    en = np.abs(ph32 - ph).max() / np.abs(ph).max()
    say(f"[config] float32-Kahan vs float64: port U {e32[0]:.2e}, I {e32[1]:.2e}, "
        f"|dS11| max {np.abs(exts2['port'].s11(50.0) - s11).max():.2e}, nf2ff phasors {en:.2e}")
    assert e32.max() < 1e-5 and en < 1e-5
    del eng2, exts2

    # the float64 oracle (port signals only)
    t = time.perf_counter()
    sim64, _ = port_spheres_oracle(wl, smp, prec=64)
    e64 = rel_l2(G, sim64.fields())
    ser64 = sim64.signal_series(0, 2)
    eu, ei = rel_l2(port_ser[:, 0], ser64[:, 0]), rel_l2(port_ser[:, 1], ser64[:, 1])
    U64 = P.dft(ser64[:, :1], (st + 1) * op.dt, wl.dft_freqs)[:, 0]
    I64 = P.dft(ser64[:, 1:2], (st + 0.5) * op.dt, wl.dft_freqs)[:, 0]
    es = np.abs(s11 - P.calc_port(U64, I64, 50.0)["s11"]).max()
    say(f"[config] port_spheres vs float64 oracle ({time.perf_counter() - t:.0f} s): fields {e64:.3e}, "
        f"port U {eu:.3e}, I {ei:.3e}, |dS11| max {es:.3e}")
    assert e64 < FP64_FIELD_TOL and max(eu, ei) < FP64_FIELD_TOL


SWEEP_MAX = 20000


def _sweep_single(ms, record=0):
    op, eng, cp, port, exc = W.build_microstrip_hip(ms, SWEEP_MAX, record_steps=record)
    r = eng.run_batch(EngineBatch(max_timesteps=SWEEP_MAX, excitations=exc, extensions=[cp, port],
                                  energy_decay_db=50.0, check_interval=100, sample_interval=100, track_peak=True,
                                  persistent_peak=True))
    return eng, port, r, exc


@pytest.mark.timeout(900)
def test_microstrip_sweep_full():
    structs = W.microstrip_sweep(256, seed=4)
    cells = np.array([m.cells for m in structs])
    order = np.argsort(cells)
    pick = {"smallest": int(order[0]), "median": int(order[len(order) // 2]), "largest": int(order[-1])}
    say(f"[config] sweep: {len(structs)} structures, {cells.sum() / 1e6:.1f} M cells, "
        f"checked {[(k, structs[v].shape) for k, v in pick.items()]}")
    sw = Sweep()
    keep = []
    for q, ms in enumerate(structs):
        op, eng, cp, port, exc = W.build_microstrip_hip(ms, SWEEP_MAX,
                                                         record_steps=SWEEP_MAX if q == pick["smallest"] else 0)
        sw.add(eng, cp, port, exc)
        keep.append((eng, port, exc))
    t = time.perf_counter()
    res = sw.run(SWEEP_MAX, energy_decay_db=50.0, check_interval=100, use_graph=True, autotune=True)
    wall = time.perf_counter() - t
    steps = np.array([r.timesteps_executed for r in res])
    say(f"[config] sweep ran in {wall:.1f} s: steps min/median/max {steps.min()}/{int(np.median(steps))}/"
        f"{steps.max()}, {sum(r.termination_reason == 1 for r in res)} decayed, tuning {sw.tuning()}")
    s11max = 0.0
    for q, (eng, port, _) in enumerate(keep):
        E, H = eng.read_fields()
        assert np.all(np.isfinite(E)) and np.all(np.isfinite(H)), structs[q].shape
        s = port.s11(50.0)
        assert np.all(np.isfinite(s))
        s11max = max(s11max, float(np.abs(s).max()))
    say(f"[config] sweep: every field finite, max |S11| over 1-20 GHz and all structures {s11max:.4f}")
    assert s11max < 1.5  # passive structures: a blown-up run shows here
    for name, q in pick.items():
        eng1, port1, r1, exc = _sweep_single(structs[q], record=SWEEP_MAX if name == "smallest" else 0)
        eng2, port2, _ = keep[q]
        say(f"[config] sweep {name} {structs[q].shape}: sweep {res[q].timesteps_executed} steps "
            f"(reason {res[q].termination_reason}), single engine {r1.timesteps_executed} "
            f"(reason {r1.termination_reason})")
        assert (res[q].timesteps_executed, res[q].termination_reason) == (r1.timesteps_executed,
                                                                           r1.termination_reason)
        assert np.array_equal(np.concatenate(eng1.read_fields()), np.concatenate(eng2.read_fields()))
        assert np.array_equal(port1.dft_values(), port2.dft_values())
    # the smallest structure against the oracle, for the steps the sweep ran
    q = pick["smallest"]
    ms = structs[q]
    eng, port, exc = keep[q]
    n = int(res[q].timesteps_executed)
    sim = O.OracleSim(*ms.shape, ms.delta)
    ms.apply(sim)
    sim.add_cpml((ms.layers,) * 6, mode=O.MODE_PHYSICAL)
    sim.add_signals(P.lumped_port_terms(ms.shape, (ms.delta,) * 3, 2, ms.port_region))
    sim.run_batch(n, [(e.position[0], e.position[1], e.position[2], e.direction, True) for e in exc],
                  np.array([e.samples[:n].astype(np.float64) for e in exc]))
    assert np.array_equal(np.concatenate(eng.read_fields()), sim.fields())
    ser = sim.signal_series(0, 2)
    assert np.array_equal(port.series(), ser)
    st = np.arange(n)
    U = P.dft(ser[:, :1], (st + 1) * eng.dt, np.linspace(1e9, 20e9, 64))[:, 0]
    I = P.dft(ser[:, 1:2], (st + 0.5) * eng.dt, np.linspace(1e9, 20e9, 64))[:, 0]
    s_ref = P.calc_port(U, I, 50.0)["s11"]
    assert np.allclose(port.s11(50.0), s_ref, rtol=1e-9, atol=1e-9)
    sim64 = O.OracleSim(*ms.shape, ms.delta, 64)
    ms.apply(sim64)
    sim64.add_cpml((ms.layers,) * 6, mode=O.MODE_PHYSICAL)
    sim64.add_signals(P.lumped_port_terms(ms.shape, (ms.delta,) * 3, 2, ms.port_region))
    sim64.run_batch(n, [(e.position[0], e.position[1], e.position[2], e.direction, True) for e in exc],
                    np.array([e.samples[:n].astype(np.float64) for e in exc]))
    s64 = sim64.signal_series(0, 2)
    U64 = P.dft(s64[:, :1], (st + 1) * eng.dt, np.linspace(1e9, 20e9, 64))[:, 0]
    I64 = P.dft(s64[:, 1:2], (st + 0.5) * eng.dt, np.linspace(1e9, 20e9, 64))[:, 0]
    e64 = rel_l2(np.concatenate(eng.read_fields()), sim64.fields())
    ds = float(np.abs(port.s11(50.0) - P.calc_port(U64, I64, 50.0)["s11"]).max())
    say(f"[config] sweep smallest {ms.shape}, {n} steps: bit-identical to the float32 oracle; vs float64: "
        f"fields {e64:.3e}, |dS11| max {ds:.3e}")
    assert e64 < FP64_FIELD_TOL and ds < 1e-3


@pytest.mark.timeout(300)
def test_extrapolator_full_size_gradients():
    from openems_amd.extrapolator import Extrapolator, ExtrapolatorDims, pack, unpack
    from oracle import extrapolator as XO

    B = 16
    dims = ExtrapolatorDims(batch=B)
    od = XO.Dims(L0=dims.L0, c=(dims.c1, dims.c2, dims.c3), hidden=dims.hidden, modes=dims.modes)
    p = XO.init_params(od, seed=11)
    rng = np.random.default_rng(11)
    for k in p:
        if k.startswith("b"):
            p[k] = 0.05 * rng.standard_normal(p[k].shape)
    p = {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}
    n_par = sum(v.size for v in p.values())
    assert 1.8e6 < n_par < 1.95e6
    # damped multi-mode windows and continuations, normalised to the window peak
    t = np.arange(dims.L0 + dims.tail)
    x = np.zeros((B, 2, dims.L0), np.float32)
    y = np.zeros((B, 2, dims.tail), np.float32)
    for b in range(B):
        for ch in range(2):
            s = sum(np.real((rng.normal() + 1j * rng.normal()) *
                            np.exp((-rng.uniform(0.002, 0.02) + 2j * np.pi * rng.uniform(0.01, 0.2)) * t))
                    for _ in range(3))
            sc = np.abs(s[:dims.L0]).max()
            x[b, ch], y[b, ch] = s[:dims.L0] / sc, s[dims.L0:] / sc
    ex = Extrapolator(dims, params=pack(p, dims))
    loss, g = ex.gradients(x, y)
    ref_loss, ref_g = XO.loss_and_grads(p, x.astype(np.float64), y.astype(np.float64), od, use_bf16=True)
    loss64, ref64 = XO.loss_and_grads(p, x.astype(np.float64), y.astype(np.float64), od, use_bf16=False)
    say(f"[config] extrapolator {n_par} parameters, batch {B}: loss {loss:.6e} vs oracle {ref_loss:.6e} "
        f"(float64 without bf16 rounding {loss64:.6e})")
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    gd = unpack(g, dims)
    for k in XO.ORDER:
        e = rel_l2(gd[k], ref_g[k])
        e64 = rel_l2(gd[k], ref64[k])
        say(f"[config] extrapolator grad {k} {gd[k].shape}: rel L2 {e:.2e} (bf16 oracle), {e64:.2e} (float64)")
        assert e < 2e-2, k
        assert e64 < 0.2, k
