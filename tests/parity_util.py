"""Run one scenario on the HIP engine (through the C ABI) and on the CPU oracle
with identical inputs, and compare.  Mirrors run_scenario_with_engine and
compare_simulation_results_allclose (src/fdtd/engine_testing.rs:401-475,
670-720)."""
from __future__ import annotations

import sys
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from openems_amd import CpmlBoundaries, EngineBatch, HipEngine, Operator, ScheduledExcitation
from oracle import excitation as X
from oracle import oracle as O

REL_L2_TOL = 1e-5  # field parity bar: relative L2 of all six arrays


@dataclass
class Scenario:
    name: str
    shape: Tuple[int, int, int]
    delta: float
    steps: int
    materials: list = field(default_factory=list)  # (eps, mu, se, sm, region)
    pecs: list = field(default_factory=list)
    excitations: list = field(default_factory=list)  # (pos, dir, soft, fn(t))
    cpml: Optional[dict] = None  # kwargs of CpmlBoundaries
    mur: Optional[Sequence[bool]] = None
    upml: Optional[dict] = None  # layers, grading_order, reflection_coeff
    ades: list = field(default_factory=list)  # (kind, params, positions, mode)
    probes: list = field(default_factory=list)
    tiles: Optional[dict] = None
    fused: Optional[int] = None  # engine step scheme (HipEngine.fused_steps); None: the default
    table: Optional[int] = None  # material table on (1) or off (0); None: the default
    graph: Optional[int] = None  # steps per captured graph (0: direct launches); None: the default


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_pair(sc: Scenario, prec: int = 32, batches: Optional[List[int]] = None):
    """Returns dict with gpu/oracle fields, probes, extension states."""
    say(f"[parity] {sc.name}: shape={sc.shape} steps={sc.steps}")
    nx, ny, nz = sc.shape
    op = Operator.uniform(nx, ny, nz, sc.delta)
    sim = O.OracleSim(nx, ny, nz, sc.delta, prec)
    for m in sc.materials:
        op.set_material(*m)
        sim.set_material(*m)
    for r in sc.pecs:
        op.set_pec(r)
        sim.set_pec(r)
    eng = HipEngine(op)
    if sc.tiles:
        eng.set_tiles(**sc.tiles)
    if sc.fused is not None:
        eng.fused_steps(sc.fused)
        if sc.fused:
            eng.set_profiling(1)  # launch counts show which scheme ran
    if sc.table is not None:
        eng.material_table(sc.table)
    if sc.graph is not None:
        eng.graph_steps(sc.graph)
    dt = op.dt
    assert dt == sim.dt
    exts = []
    if sc.cpml is not None:
        cfg = CpmlBoundaries(**sc.cpml)
        exts.append(eng.cpml(cfg))
        sim.add_cpml(cfg.layers, cfg.grading_order, cfg.reflection_coeff, cfg.alpha_max, cfg.kappa_max, cfg.mode)
    ade_ext = []
    for kind, params, pos, mode in sc.ades:
        x = eng.ade(kind, params, pos, mode)
        exts.append(x)
        ade_ext.append(x)
        sim.add_ade(kind, params, pos, mode)
    if sc.mur is not None:
        exts.append(eng.mur(sc.mur))
        sim.add_mur(sc.mur)
    upml_ext = None
    if sc.upml is not None:
        upml_ext = eng.upml(**sc.upml)
        exts.append(upml_ext)
        sim.add_upml(**sc.upml)
    probe_ext = []
    for p in sc.probes:
        x = eng.field_probe(p)
        exts.append(x)
        probe_ext.append(x)
        sim.add_field_probe(p)

    batches = batches or [sc.steps]
    start = 0
    for nb in batches:
        g_exc, o_exc, o_smp = [], [], []
        for pos, d, soft, fn in sc.excitations:
            smp = X.schedule(fn, start, nb, dt)
            g_exc.append(ScheduledExcitation(tuple(pos), d, smp, soft))
            o_exc.append((pos[0], pos[1], pos[2], d, soft))
            o_smp.append(smp.astype(np.float64))
        o_smp = np.array(o_smp) if o_smp else None
        eng.run_batch(EngineBatch(num_steps=nb, excitations=g_exc, extensions=exts))
        sim.run_batch(nb, o_exc, o_smp)
        start += nb
    E, H = eng.read_fields()
    out = dict(gpu=np.concatenate([E, H]), oracle=sim.fields(), eng=eng, sim=sim, exts=exts)
    out["gpu_probes"] = [x.data() for x in probe_ext]
    out["oracle_probes"] = [sim.probe_data(q) for q in range(len(sc.probes))]
    out["gpu_ade"] = [x.state() for x in ade_ext]
    out["oracle_ade"] = [sim.ade_state(q) for q in range(len(sc.ades))]
    out["upml"] = upml_ext
    return out


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def assert_parity(res, name="", bit_exact=True):
    g, o = res["gpu"], res["oracle"]
    assert np.all(np.isfinite(g)), f"{name}: non-finite GPU field"
    r = rel_l2(g, o)
    say(f"[parity] {name}: rel_l2={r:.3e} max_abs={np.max(np.abs(g.astype(np.float64) - o)):.3e} "
        f"|o|={np.linalg.norm(o.astype(np.float64)):.3e}")
    assert r <= REL_L2_TOL, f"{name}: rel L2 {r}"
    if bit_exact:
        bad = np.count_nonzero(g != o)
        assert bad == 0, f"{name}: {bad} elements differ bitwise (rel L2 {r})"
    for gp, op_ in zip(res["gpu_probes"], res["oracle_probes"]):
        assert gp.shape == op_.shape
        assert np.array_equal(gp, op_) if bit_exact else rel_l2(gp, op_) <= REL_L2_TOL


def gauss_ref(center_freq, bw=0.5):
    t0, tau = X.reference_gaussian_params(center_freq, bw)
    return lambda t: X.gaussian(t, t0, tau)


def port_spheres_oracle(wl, samples, surf_terms=(), prec: int = 32):
    """The composite workload (openems_amd.workloads.PortSpheres) on the
    oracle: Lorentz background material, port resistor, CPML on five faces,
    Mur on +z, Drude and Lorentz ADE, the port's U/I signals followed by
    surf_terms, driven by the port's soft sources.  Returns (sim, port_terms)."""
    from oracle import port as P

    sim = O.OracleSim(*wl.shape, wl.delta, prec)
    drude, lorentz = [], []
    for kind, c, r in wl.spheres:
        cells = wl.sphere_cells(c, r)
        (drude if kind == "drude" else lorentz).append(cells)
        if kind == "lorentz":
            for i, j, k0, k1 in wl.columns(cells):
                sim.set_material(wl.LORENTZ_EPS_INF, 1.0, 0.0, 0.0, (i, i + 1, j, j + 1, k0, k1))
    sim.add_lumped(2, wl.port_region, wl.port_R)
    sim.add_cpml(wl.cpml_layers, mode=O.MODE_PHYSICAL)
    sim.add_mur(wl.mur_faces)
    if drude:
        sim.add_ade(O.ADE_DRUDE, wl.DRUDE, np.concatenate(drude), O.MODE_PHYSICAL)
    if lorentz:
        sim.add_ade(O.ADE_LORENTZ, wl.LORENTZ, np.concatenate(lorentz), O.MODE_PHYSICAL)
    port_terms = P.lumped_port_terms(wl.shape, (wl.delta,) * 3, 2, wl.port_region)
    sim.add_signals(list(port_terms) + list(surf_terms))
    exc = P.port_excitations(2, wl.port_region)
    sim.run_batch(wl.steps, [(p[0], p[1], p[2], d, True) for p, d, a in exc],
                  np.array([a * samples.astype(np.float64) for _, _, a in exc]))
    return sim, port_terms
