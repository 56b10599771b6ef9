#!/usr/bin/env python3
"""Throughput of the FDTD timestep path on one MI355X.

Workload (default): 256^3 grid, 0.5 mm cells, 8-cell CPML on every face
(physical mode), 200 seeded dielectric boxes, soft Ez Gaussian (fc = 15 GHz)
at the centre; energy sampled every 100 steps as the reference's
Simulation::run does.  A "step" is one full timestep (H update with fused
CPML, E update with fused CPML, excitation, energy monitoring).  --cavity
runs the 100^3 plain vacuum cavity instead (1 mm cells, the reference's PEC
faces), the size the reference publishes its own per-step times for; its
vs_baseline is against the reference's best published 100^3 figure.

Prints one JSON line:
  value      MCells/s = cells * steps / wall over the timed steps (inputs and
             state resident in HBM; the timed region is bracketed by device
             synchronisation)
  roofline   the dominant kernel's algorithmic bytes per launch divided by its
             mean launch time, measured with HIP events on the engine's stream
             inside the timed region, against 8 TB/s
  hbm_frac_step       the whole step's algorithmic bytes (this layout) / wall
                      / 8 TB/s
  array_layout_equivalent_step_gbs
                      the metric's byte convention (120 B per cell-update,
                      the reference's 12 coefficient arrays, plus 32 B per
                      CPML slab cell and direction) / wall: a bandwidth in
                      that layout's bytes, not a fraction of HBM, since the
                      material table does not move those coefficient bytes
  stream_copy         a 1 GiB device-to-device copy measured after the timed
                      region, and the step and kernel rates as fractions of it
  cpu_baseline  the oracle's OpenMP restatement of the reference's
             ParallelEngine (same arithmetic, AVX-512 when present) timed on
             the host cores on a bounded sample of the same workload
  perf_point_512  the same workload at 512^3 (200 timed steps, its own
             roofline record), after the headline and outside its timing
  sweep      the 256-structure microstrip sweep to -50 dB or 20000 steps:
             wall time and cell-updates per second
"""
from __future__ import annotations

import argparse
import contextlib
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MCells/s & %HBM peak; sweep time-to-|ΔS11|≤1e-3, learned-stop vs plain"
HBM_PEAK_GBS = 8000.0
REF_MCELLS_100 = 718.0  # reference WGPU engine, 100^3 vacuum Yee, Apple M2 Pro


def log(msg):
    print(msg, file=sys.stderr, flush=True)


@contextlib.contextmanager
def quiet_gc():
    """Collect Python garbage before a timed region and pause the collector
    inside it: a full collection over the interpreter's heap (torch leaves a
    large one) takes tens of milliseconds and lands at random."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def slab_cells(shape, layers):
    """Cells carrying CPML psi per direction, summed over directions."""
    nx, ny, nz = shape
    tot = 0
    for d, n in enumerate(shape):
        other = nx * ny * nz // n
        L = 2 * layers if 2 * layers < n else 0
        tot += L * other
    return tot


CELL_BYTES = {
    # per cell and half-step: the own 3 components read and written, the other
    # field's 3 components read, and the coefficients
    "arrays": 60,   # 6 coefficient floats (the reference's 12-array layout)
    "table": 38,    # one 2-byte material-table row index
    "table8": 37,   # a table of 2 to 256 rows: a 1-byte row index
    "table1": 36,   # a table of one row: the kernels read no index
}


def half_coef(coef, rows, h):
    """Byte class of one half ('update_h' when h) from the engine's material
    table rows (E, H): a tabulated half of one row reads no row index, one
    of at most 256 rows reads it as a byte (unless OE_BYTE_INDEX=0)."""
    r = rows[1 if h else 0]
    if coef != "table":
        return coef
    if r == 1:
        return "table1"
    if r <= 256 and os.environ.get("OE_BYTE_INDEX", "1") != "0":
        return "table8"
    return "table"


def kernel_bytes(shape, layers, coef="arrays"):
    """Algorithmic HBM bytes of one H (or E) half-step over the whole grid:
    CELL_BYTES[coef] per cell plus 2 psi reads + 2 psi writes per slab cell
    per direction."""
    n = shape[0] * shape[1] * shape[2]
    return CELL_BYTES[coef] * n + 16 * slab_cells(shape, layers)


FUSED_CELL_BYTES = {
    # per cell and fused step: E and H read and written (48 B), and the
    # coefficients of both halves
    "arrays": 96,   # 12 coefficient floats
    "table": 48,    # plus the row indices (fused_index_bytes)
}


def fused_index_bytes(rows):
    """Row-index bytes per cell of the fused pass: none for a half of one
    row; the E index as a byte for 2 to 256 rows when the default kernel
    (OE_FUSED_KIND 2) reads the byte-wide copy (OE_BYTE_INDEX), else 2 B;
    the H index always 2 B."""
    e_rows, h_rows = rows
    byte_e = (os.environ.get("OE_BYTE_INDEX", "1") != "0" and os.environ.get("OE_FUSED_KIND", "2") != "1"
              and 2 <= e_rows <= 256)
    e = 0 if e_rows == 1 else (1 if byte_e else 2)
    return e + (0 if h_rows == 1 else 2)


def launch_bytes(name, shape, layers, coef="arrays", rows=(0, 0)):
    """Algorithmic bytes of one launch of the plan: 'grid' (no CPML), or
    'interior' (x/y slabs excluded; z-face psi at the row ends) and 'slabs'
    (the x/y face slabs with all faces) when a CPML is attached.  A fused
    launch ('fused:interior', 'fused:grid') covers both half-steps of its
    box: FUSED_CELL_BYTES per cell, 32 B per z-slab cell (H and E psi).
    rows: the material table's rows of E and H (a half of one row reads no
    index)."""
    nx, ny, nz = shape
    kind, part = name.split(":")
    if kind == "fused":
        ix, jy = (nx - 2 * layers, ny - 2 * layers) if part == "interior" else (nx, ny)
        zl = 2 * layers if layers and 2 * layers < nz else 0
        fb = FUSED_CELL_BYTES[coef] + (fused_index_bytes(rows) if coef == "table" else 0)
        return fb * ix * jy * nz + 32 * ix * jy * zl
    coef = half_coef(coef, rows, kind == "update_h")
    if part == "grid" or not layers:
        return kernel_bytes(shape, 0, coef)
    ix, jy = nx - 2 * layers, ny - 2 * layers
    interior = CELL_BYTES[coef] * ix * jy * nz + 16 * ix * jy * 2 * layers
    if part == "interior":
        return interior
    return kernel_bytes(shape, layers, coef) - interior


def summarize_rocprof(paths):
    """Markdown summary of rocprofv3 result databases (kernel statistics and
    per-dispatch counters).  Only the standard library is used."""
    import sqlite3

    out = []
    for path in paths:
        con = sqlite3.connect(path)
        views = {r[0] for r in con.execute("select name from sqlite_master")}
        out.append(f"### `{path}`\n")
        if "top_kernels" in views:
            rows = con.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
            if rows:
                out.append("| kernel | calls | total us | mean us | % |\n|---|---|---|---|---|")
                for r in rows[:12]:
                    out.append(f"| `{r[0][:90]}` | {r[1]} | {r[2]:.1f} | {r[3]:.2f} | {r[4]:.2f} |")
                out.append("")
        if "kernels" in views:
            out.append(busy_summary(con.execute("select start, end from kernels order by start").fetchall()))
            out.append(grid_split(con))
        if "counters_collection" in views:
            rows = con.execute("select kernel_name, counter_name, count(*), avg(value), min(value), max(value) "
                               "from counters_collection group by kernel_name, counter_name "
                               "order by avg(value) desc").fetchall()
            if rows:
                out.append("| kernel | counter | dispatches | mean KiB | min | max |\n|---|---|---|---|---|---|")
                for r in rows[:10]:
                    out.append(f"| `{r[0][:90]}` | {r[1]} | {r[2]} | {r[3]:.0f} | {r[4]:.0f} | {r[5]:.0f} |")
                out.append("")
            names = {r[1] for r in rows}
            if {"SQ_VALU_MFMA_BUSY_CYCLES", "GRBM_GUI_ACTIVE"} <= names:
                out.append(mfma_summary(con))
    return "\n".join(out)


def grid_split(con) -> str:
    """Kernels launched with more than one grid size (one kernel serving
    several extensions, e.g. a port's DFT and a recording surface's), one
    row per grid size (total work-items along x, y, z)."""
    cols = [r[1] for r in con.execute("pragma table_info(kernels)")]
    grid = [c for c in cols if c.startswith("grid")][:3]
    if "name" not in cols or not grid:
        return ""
    rows = con.execute(f"select name, {', '.join(grid)}, count(*), sum(end - start), avg(end - start) from kernels "
                       f"group by name, {', '.join(grid)}").fetchall()
    names = [r[0] for r in rows]
    rows = [r for r in rows if names.count(r[0]) > 1]
    if not rows:
        return ""
    out = ["| kernel | grid | calls | total us | mean us |", "|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: (r[0], -r[-2])):
        g = "x".join(str(v) for v in r[1:1 + len(grid)])
        out.append(f"| `{r[0][:90]}` | {g} | {r[-3]} | {r[-2] / 1e3:.1f} | {r[-1] / 1e3:.2f} |")
    return "\n".join(out) + "\n"


def busy_summary(intervals, host_gap_ns: float = 1e6):
    """Device busy time of a kernel trace: the union of the dispatches'
    [start, end) intervals (ns) against the span from the first start to the
    last end.  Gaps longer than host_gap_ns (host phases) are reported
    apart from the short gaps between launches."""
    if not intervals:
        return ""
    busy = 0
    gaps = []
    cur_s, cur_e = intervals[0]
    for s, e in intervals[1:]:
        if s > cur_e:
            busy += cur_e - cur_s
            gaps.append(s - cur_e)
            cur_s, cur_e = s, e
        else:
            cur_e = max(cur_e, e)
    busy += cur_e - cur_s
    span = max(e for _, e in intervals) - intervals[0][0]
    big = sum(g for g in gaps if g > host_gap_ns)
    small = [g for g in gaps if g <= host_gap_ns]
    return (f"Dispatches {len(intervals)}, span {span / 1e9:.3f} s, busy {busy / 1e9:.3f} s "
            f"({busy / span:.3f}); gaps over {host_gap_ns / 1e6:g} ms {big / 1e9:.3f} s; "
            f"busy / (span - those gaps) {busy / (span - big):.3f}; "
            f"{len(small)} shorter gaps, {sum(small) / 1e9:.3f} s.\n")


def mfma_summary(con, xcds: int = 8, simds: int = 1024):
    """Matrix-core busy fraction per kernel and grid size: the sum of
    SQ_VALU_MFMA_BUSY_CYCLES over all SIMDs divided by (GRBM_GUI_ACTIVE per
    XCD x SIMDs); GRBM_GUI_ACTIVE as stored is summed over the 8 XCDs.  FLOP
    from SQ_INSTS_VALU_MFMA_MOPS_F32 x 512."""
    rows = con.execute("select dispatch_id, kernel_name, grid_size, counter_name, sum(value) from counters_collection "
                       "group by dispatch_id, counter_name").fetchall()
    per = {}
    for disp, name, grid, cn, v in rows:
        d = per.setdefault(disp, {"name": name, "grid": grid})
        d[cn] = v
    groups = {}
    for d in per.values():
        if "SQ_VALU_MFMA_BUSY_CYCLES" not in d or d.get("SQ_VALU_MFMA_BUSY_CYCLES", 0) == 0:
            continue
        groups.setdefault((d["name"][:70], d["grid"]), []).append(d)
    out = ["| kernel | work-items | dispatches | MFMA busy % of SIMD-cycles | MFMA GFLOP |", "|---|---|---|---|---|"]
    for (name, grid), ds in sorted(groups.items()):
        busy = sum(d["SQ_VALU_MFMA_BUSY_CYCLES"] for d in ds)
        gui = sum(d["GRBM_GUI_ACTIVE"] for d in ds) / xcds
        fl = sum(d.get("SQ_INSTS_VALU_MFMA_MOPS_F32", 0) for d in ds) * 512 / 1e9
        out.append(f"| `{name}` | {grid} | {len(ds)} | {100 * busy / (gui * simds):.1f} | {fl:.2f} |")
    return "\n".join(out) + "\n"


def measured_traffic(kernel_key):
    """HBM bytes per launch from the committed counter summary, if present."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(kernel_key)
    except (OSError, ValueError):
        return None


def build_engine(wl, mode):
    import numpy as np

    from openems_amd import CpmlBoundaries, HipEngine, Operator
    from openems_amd import workloads as W

    op = Operator.uniform(*wl.shape, wl.delta)
    W.apply_materials(op, wl)
    eng = HipEngine(op)
    exts = []
    if wl.cpml_layers:
        exts.append(eng.cpml(CpmlBoundaries(layers=(wl.cpml_layers,) * 6, grading_order=wl.cpml_grading,
                                            reflection_coeff=wl.cpml_R0, mode=mode)))
    return op, eng, exts


def run_steps(eng, exts, wl, dt, start, n, sample_interval=100):
    from openems_amd import EngineBatch, ScheduledExcitation, Waveform

    t0, tau = wl.gaussian_t0_tau()
    smp = Waveform.gaussian(t0, tau).sample(start, n, dt)
    return eng.run_batch(EngineBatch(num_steps=n, excitations=[ScheduledExcitation(wl.source, wl.source_dir, smp)],
                                     extensions=exts, sample_interval=sample_interval, track_peak=True))


def stream_copy_gbs(nbytes=1 << 30, reps=20):
    """Device-to-device copy of nbytes (read + write counted) on the current
    stream, mean over reps after two warm-up copies: the streaming ceiling a
    read/write kernel can reach on this device."""
    import torch

    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(wl, budget_s=12.0, max_steps=4000):
    """Time the oracle (OpenMP, all host threads) on the same workload."""
    import numpy as np

    from oracle import oracle as O
    from openems_amd import workloads as W

    sim = O.OracleSim(*wl.shape, wl.delta, 32)
    W.apply_materials(sim, wl)
    if wl.cpml_layers:
        sim.add_cpml((wl.cpml_layers,) * 6, wl.cpml_grading, wl.cpml_R0, 0.0, 1.0, O.MODE_PHYSICAL)
    sim.yee_steps(1)  # first touch
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and time.perf_counter() - t0 < budget_s:
        sim.yee_steps(2)
        steps += 2
    wall = time.perf_counter() - t0
    threads = O.max_threads()
    isa = "AVX-512" if O.library_path().endswith("v4.so") else "AVX2"
    return {
        "value": wl.cells * steps / wall / 1e6,
        "unit": "MCells/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} timesteps (H+E{' with physical CPML' if wl.cpml_layers else ''}) of the "
                  f"{wl.shape[0]}x{wl.shape[1]}x{wl.shape[2]} "
                  f"workload on {threads} OpenMP threads ({isa}), {wall:.1f} s",
    }


def bench_nf2ff(args):
    """NF2FF transform of the 384x384x256 workload's recording box: nodes
    18..366 x 18..366 x 18..238 (10 cells inside the 8-cell CPML), 0.5 mm
    cells, 181 x 360 (theta, phi), frequencies 5, 10, 15 GHz, seeded
    synthetic surface phasors already resident on the device."""
    import numpy as np
    import torch

    from openems_amd import Nf2ff

    d = 0.5e-3
    lo, hi = (18, 18, 18), (366, 366, 238)
    axes = [np.arange(lo[a], hi[a] + 1) * d for a in range(3)]
    rng = np.random.default_rng(7)
    th = np.linspace(0, np.pi, 181)
    ph = np.linspace(0, 2 * np.pi, 361)[:-1]
    freqs = (5e9, 10e9, 15e9)
    nfs = []
    for f in freqs:
        nf = Nf2ff(f)
        for ax in range(3):
            a1, a2 = (ax + 1) % 3, (ax + 2) % 3
            for side in (0, 1):
                shape = (2, axes[a1].size, axes[a2].size)
                e = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
                h = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / 376.73
                nf.add_surface(ax, 1 if side else -1, (hi if side else lo)[ax] * d, axes[a1], axes[a2], e, h)
        nfs.append(nf)
    log(f"[bench] nf2ff: box {[a.size for a in axes]} nodes, {th.size}x{ph.size} angles, {len(freqs)} frequencies")
    for _ in range(args.warmup):
        for nf in nfs:
            nf.calculate(th, ph)
    torch.cuda.synchronize()
    gemm_ms = total_ms = flop = split_ms = split_flop = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for nf in nfs:
            r = nf.calculate(th, ph)
            gemm_ms += r.gemm_ms
            total_ms += r.total_ms
            flop += r.gemm_flop
            split_ms += r.gemm_split_ms
            split_flop += r.gemm_split_flop
    wall = time.perf_counter() - t0
    n_calc = args.steps * len(freqs)
    # The GEMMs' own rate, for the roofline record: the same transforms with
    # every face on one stream (by default the x/y faces run on a side
    # stream beside the z faces' GEMMs, and the GEMM events then also time
    # the kernels that share the device with them).
    prev = os.environ.get("OE_NF2FF_STREAMS")
    os.environ["OE_NF2FF_STREAMS"] = "1"
    try:
        gemm_ms = flop = split_ms = split_flop = 0.0
        for _ in range(args.steps):
            for nf in nfs:
                r = nf.calculate(th, ph)
                gemm_ms += r.gemm_ms
                flop += r.gemm_flop
                split_ms += r.gemm_split_ms
                split_flop += r.gemm_split_flop
    finally:
        if prev is None:
            del os.environ["OE_NF2FF_STREAMS"]
        else:
            os.environ["OE_NF2FF_STREAMS"] = prev
    tflops = flop / (gemm_ms * 1e-3) / 1e12
    out = {
        "metric": "NF2FF far-field transform, ms per frequency (181x360 angles)",
        "value": round(wall / n_calc * 1e3, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 operands split into two fp16 halves on the matrix cores (fp32 accumulate) / f64 (phases, "
                 "reduction, P_rad)",
        "data": "synthetic",
        "config": {"workload": "nf2ff box of the 384x384x256 CPML workload: 349x349x221 surface nodes, "
                               "0.5 mm cells, 181x360 angles, 5/10/15 GHz",
                   "device_ms_per_frequency": round(total_ms / n_calc, 3)},
        "roofline": ({"bound": "mfma",
                      "kernel": "k_hgemm3 (3 x v_mfma_f32_32x32x16_f16 per FP32 product, operands split hi + lo)",
                      # matrix-core work: three fp16 products per useful FP32 multiply-add
                      "achieved": round(3 * split_flop / (split_ms * 1e-3) / 1e12, 2), "peak": 2500.0,
                      "unit": "TFLOP/s", "frac": round(3 * split_flop / (split_ms * 1e-3) / 1e12 / 2500.0, 4),
                      "traffic": None,
                      "fp32_equivalent_tflops": round(split_flop / (split_ms * 1e-3) / 1e12, 2),
                      "flop_per_frequency": flop / n_calc,
                      "gemm_ms_per_frequency": round(gemm_ms / n_calc, 3),
                      "split_gemm_ms_per_frequency": round(split_ms / n_calc, 3),
                      "all_gemms_fp32_equivalent_tflops": round(tflops, 2),
                      "timed": "GEMM launches of a separate pass with every face on one stream"}
                     if split_ms > 0 else
                     {"bound": "mfma", "kernel": "k_sgemm_mfma (v_mfma_f32_32x32x2f32)",
                      "achieved": round(tflops, 2), "peak": 157.3, "unit": "TFLOP/s",
                      "frac": round(tflops / 157.3, 4), "traffic": None,
                      "flop_per_frequency": flop / n_calc,
                      "gemm_ms_per_frequency": round(gemm_ms / n_calc, 3)}),
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_sweep(args):
    """The microstrip sweep: seeded structures (grids U{60..220} x U{60..120}
    x U{30..48}, nz rounded up to a multiple of 4; 0.5 mm cells; PEC ground,
    substrate eps_r U[2.2,10.2] 1-4 mm, PEC trace polyline of 2-6 segments
    1-4 mm wide; 8-cell CPML; 50-ohm port at the trace start; 64-frequency
    port DFT) run to -50 dB energy decay (whole-run peak) or max steps in one
    sweep: one launch per kernel variant per half-step for all structures,
    graph-replayed between energy checks every 100 steps, tile depth
    autotuned first."""
    import numpy as np
    import torch

    from openems_amd import Sweep
    from openems_amd import workloads as W

    structs = W.microstrip_sweep(args.sweep_count, seed=4)
    log(f"[bench] sweep: building {len(structs)} structures, {sum(m.cells for m in structs) / 1e6:.1f} M cells")
    t_build = time.perf_counter()
    sw = Sweep()
    keep = []
    for ms in structs:
        op, eng, cp, port, exc = W.build_microstrip_hip(ms, args.sweep_max_steps)
        sw.add(eng, cp, port, exc)
        keep.append((ms, eng, port))
    torch.cuda.synchronize()
    log(f"[bench] sweep: built in {time.perf_counter() - t_build:.1f} s; running")
    with quiet_gc():
        t0 = time.perf_counter()
        res = sw.run(args.sweep_max_steps, energy_decay_db=50.0, check_interval=100, use_graph=True, autotune=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    steps = np.array([r.timesteps_executed for r in res])
    cells = np.array([ms.cells for ms, _, _ in keep])
    updates = float((steps * cells).sum())
    decayed = int(sum(r.termination_reason == 1 for r in res))
    s11 = [port.s11(50.0) for _, _, port in keep]
    finite = bool(all(np.isfinite(x).all() for x in s11))
    tune = sw.tuning()
    out = {
        "metric": "microstrip sweep: wall time to -50 dB decay (or max steps) for every structure, plain method",
        "value": round(wall, 3),
        "unit": "s",
        "n_gpus": 1,
        "steps": int(steps.max()),
        "warmup": 0,
        "ms_per_step": round(wall / max(int(steps.max()), 1) * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"microstrip sweep: {len(structs)} structures (seed 4), "
                               f"{cells.sum() / 1e6:.1f} M cells, max {args.sweep_max_steps} steps, -50 dB decay",
                   "tiles": tune},
        "cell_updates_per_s": round(updates / wall / 1e6, 1),
        "cell_updates_unit": "MCells/s",
        "steps_min_median_max": [int(steps.min()), int(np.median(steps)), int(steps.max())],
        "decayed": decayed,
        "s11_finite": finite,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_composite(args):
    """The composite workload at full size: 384x384x256 cells of 0.5 mm, CPML
    (8 cells, physical) on five faces, first-order Mur on +z, a 50-ohm port
    (1x1x4 cells) with an openEMS Gaussian (fc = 20 GHz), 8 Drude and 8
    Lorentz spheres (seeded), port DFT over 64 frequencies (1-20 GHz, FP64),
    NF2FF box 10 cells inside the CPML recording 5/10/15 GHz every
    --nf2ff-stride steps; then the far field at 181x360 angles for the three
    frequencies."""
    import numpy as np
    import torch

    from openems_amd import EngineBatch
    from openems_amd import workloads as W
    from openems_amd.engine import port_excitations

    wl = W.port_spheres(steps=args.steps)
    op, eng, exts = W.build_port_spheres_hip(wl, precision=0, nf2ff_stride=args.nf2ff_stride)
    if args.tiles:
        bx, by, ci, vec = (int(x) for x in args.tiles.split(","))
        eng.set_tiles(bx, by, ci, vec)
    smp = wl.excitation_samples(op.dt, args.warmup + args.steps)
    exc = port_excitations(2, wl.port_region, smp)
    log(f"[bench] composite {wl.shape}: {len(wl.spheres)} spheres, warmup {args.warmup}")
    if args.warmup:
        eng.run_batch(EngineBatch(num_steps=args.warmup, excitations=exc, extensions=list(exts.values())))
    torch.cuda.synchronize()
    exc = port_excitations(2, wl.port_region, smp[args.warmup:])
    with quiet_gc():
        t0 = time.perf_counter()
        eng.run_batch(EngineBatch(num_steps=args.steps, excitations=exc, extensions=list(exts.values())))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    th = np.linspace(0, np.pi, 181)
    ph = np.linspace(0, 2 * np.pi, 361)[:-1]
    t1 = time.perf_counter()
    ffs = [exts["nf2ff"].far_field(q, th, ph) for q in range(len(wl.nf2ff_freqs))]
    t_ff = time.perf_counter() - t1
    s11 = exts["port"].s11(50.0)
    out = {
        "metric": "MCells/s, composite workload (every extension on)",
        "value": round(wl.cells * args.steps / wall / 1e6, 1),
        "unit": "MCells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields, f64 DFT accumulators",
        "data": "synthetic",
        "config": {"workload": f"port_spheres {wl.shape[0]}x{wl.shape[1]}x{wl.shape[2]}: CPML on 5 faces, Mur +z, "
                               f"50-ohm port, {len(wl.spheres)} Drude/Lorentz spheres (ADE, physical), 64-frequency "
                               f"port DFT, NF2FF box {wl.nf2ff_box} every {args.nf2ff_stride} steps",
                   "nf2ff_far_field_s": round(t_ff, 3)},
        "dmax": [round(f.dmax, 4) for f in ffs],
        "s11_finite": bool(np.isfinite(s11).all()),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_learned(args):
    """Learned-stop vs plain microstrip sweeps.  Training data: runs of
    --learned-train structures (seed 5) recorded for --learned-data-steps
    steps; each gives pairs (2048-step window, next 4096 steps) at offsets
    0, 512, ..., recorded length - 6144; training starts from spread mode
    frequencies and grows the tail over the epochs.  Held-out:
    --learned-test structures (seed 6) run plain (to -50 dB or
    --learned-max-steps) and learned: checkpoints at 2048 + 256 steps, then
    every --learned-stride; at a checkpoint the network's poles with
    least-squares residues must reproduce the last 256 steps to 1e-6 of the
    window energy, predict at most 1e-4 of the energy so far after it, and
    keep S11 within 3e-4 when the predicted plain stop moves by 300 steps;
    an accepted structure's tail DFT runs to that predicted stop.  Reports
    wall times, the accepted count, the |dS11| distribution over 1-20 GHz
    against the plain runs, and on held-out windows of the plain runs the
    tail error (relative to predicting zero) of the network's own modes, of
    its poles with refitted residues, and of the untrained spread poles with
    refitted residues."""
    import numpy as np

    from openems_amd import learned_stop as LS
    from openems_amd import workloads as W
    from openems_amd.extrapolator import Extrapolator, ExtrapolatorDims

    window, dec, check, T = 2048, 8, 256, args.learned_max_steps
    tail_steps = 4096
    TD = args.learned_data_steps
    offsets = list(range(0, TD - window - tail_steps + 1, 512))
    tail = tail_steps // dec
    rng_tr = np.random.default_rng(5)
    rng_te = np.random.default_rng(6)
    train_s = [W.microstrip(rng_tr) for _ in range(args.learned_train)]
    test_s = [W.microstrip(rng_te) for _ in range(args.learned_test)]
    X, Y = [], []
    t_data = time.perf_counter()
    for b0 in range(0, len(train_s), 256):
        run = LS.run_sweep(train_s[b0:b0 + 256], TD, None, record_steps=TD)
        x, y = LS.dataset(run, window, dec, tail, offsets)
        X.append(x.astype(np.float32))
        Y.append(y.astype(np.float32))
        log(f"[bench] learned: training data {b0 + len(run.u)}/{len(train_s)}, sweep {run.wall:.1f} s")
        del run
    X = np.concatenate(X)
    Y = np.concatenate(Y)
    t_data = time.perf_counter() - t_data
    dims = ExtrapolatorDims(L0=window // dec, batch=min(256, len(X)), tail=tail)
    model = Extrapolator(dims, seed=0)
    LS.spread_modes(model)
    start_poles = LS.spread_poles(model)
    t_train = time.perf_counter()
    losses = LS.train(model, X, Y, args.learned_epochs, 1e-3, log=lambda m: log(f"[bench] learned: {m}"))
    t_train = time.perf_counter() - t_train
    plain = LS.run_sweep(test_s, T, 50.0, record_steps=T)
    log(f"[bench] learned: plain held-out sweep {plain.wall:.1f} s, steps {int(plain.steps.sum())}")
    # held-out windows: offsets whose tail lies inside the plain run
    Xh, Yh = [], []
    for q in range(len(test_s)):
        for o in offsets:
            if o + window + tail_steps <= plain.steps[q]:
                x, y, _ = LS.training_pair(plain.u[q][o:], plain.i[q][o:], window, dec, tail)
                Xh.append(x)
                Yh.append(y)
    Xh, Yh = np.array(Xh), np.array(Yh)
    gain = None
    if len(Xh):
        sub = slice(0, min(len(Xh), 1024))
        gain = {"network_modes": LS.tail_error(model, Xh[sub], Yh[sub], horizons=(tail,))[tail],
                "network_poles_refit": LS.refit_tail_error(model, Xh[sub], Yh[sub]),
                "untrained_poles_refit": LS.refit_tail_error(model, Xh[sub], Yh[sub], poles=start_poles),
                "windows": int(len(Xh[sub]))}
    log(f"[bench] learned: held-out tail error / zero-prediction error {gain}")
    learned = LS.learned_sweep(model, test_s, T, window, dec, check, stride=args.learned_stride,
                               log=lambda m: log(f"[bench] learned (network): {m}"))
    ds11 = np.abs(learned.s11 - plain.s11).max(axis=1)
    acc = learned.accepted
    log(f"[bench] learned: checkpoints {learned.checkpoints}; accepted {int(acc.sum())}; "
        f"steps plain {int(plain.steps.sum())} learned {int(learned.steps_learned.sum())}")
    qs = (lambda v: [float(np.median(v)), float(np.percentile(v, 90)), float(v.max())] if len(v) else None)
    # the same gates and horizon with classical matrix-pencil poles: what an
    # exact-structure extrapolator reaches against the -50 dB plain run
    pen = LS.learned_sweep(model, test_s, T, window, dec, check, stride=args.learned_stride, poles="pencil",
                           log=lambda m: log(f"[bench] learned (pencil): {m}"))
    dp = np.abs(pen.s11 - plain.s11).max(axis=1)
    pencil = {"wall_ratio": round(pen.wall_learned / plain.wall, 4), "accepted": int(pen.accepted.sum()),
              "within_1e-3": int((dp[pen.accepted] <= 1e-3).sum()),
              "dS11_accepted_median_p90_max": qs(dp[pen.accepted]),
              "steps_learned": int(pen.steps_learned.sum())}
    log(f"[bench] learned: pencil poles {pencil}")
    out = {
        "metric": "sweep time-to-|dS11|<=1e-3, learned-stop vs plain",
        "value": round(learned.wall_learned / plain.wall, 4),
        "unit": "learned wall / plain wall",
        "n_gpus": 1,
        "steps": int(plain.steps.max()),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields, bf16 network (fp32 accumulate)",
        "data": "synthetic",
        "config": {"workload": f"microstrip sweep, {len(train_s)} training runs (seed 5) of {TD} steps x "
                               f"{len(offsets)} window offsets, {len(test_s)} held-out (seed 6), max {T} steps, "
                               f"window {window}, check {check}, checkpoint stride {args.learned_stride}, "
                               f"decimation {dec}",
                   "params": int(model.n_params)},
        "plain_wall_s": round(plain.wall, 3),
        "learned_wall_s": round(learned.wall_learned, 3),
        "inference_wall_s": round(learned.wall_inference, 3),
        "steps_plain": int(plain.steps.sum()),
        "steps_learned": int(learned.steps_learned.sum()),
        "accepted": int(acc.sum()),
        "acceptance": "refit residual <= 1e-6 of window energy, predicted tail energy <= 1e-4 of energy so far, "
                      "S11 change <= 3e-4 for +-300 steps of predicted stop",
        "dS11_accepted_median_p90_max": qs(ds11[acc]),
        "within_1e-3": int((ds11[acc] <= 1e-3).sum()),
        "continued_equal_plain": bool(np.array_equal(learned.s11[~acc], plain.s11[~acc])),
        "train_loss_first_last": [losses[0], losses[-1]],
        "heldout_tail_error_vs_zero": gain,
        "pencil_poles_same_gates": pencil,
        "data_s": round(t_data, 1),
        "train_s": round(t_train, 1),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def yee_point(wl, mode, steps, warmup, period, tiles="", coef_opt="auto", fused=False, mode_name="physical"):
    """Build the workload's engine, run `warmup` untimed steps, then time
    `steps` steps (HIP events around every launch of every `period`-th step)
    and return the measurements: MCells/s, the dominant Yee launch against
    the HBM roofline, the whole step's algorithmic bytes per second."""
    import numpy as np
    import torch

    log(f"[bench] building {wl.shape} with {len(wl.boxes)} boxes, CPML {wl.cpml_layers} ({mode_name})")
    op, eng, exts = build_engine(wl, mode)
    if tiles:
        bx, by, ci, vec = (int(x) for x in tiles.split(","))
        eng.set_tiles(bx, by, ci, vec)
    rows = eng.material_table(0 if coef_opt == "arrays" else -1)
    eng.fused_steps(1 if fused else 0)
    coef = "table" if rows[0] and rows[1] else "arrays"
    if coef_opt == "table" and coef != "table":
        raise SystemExit(f"material table unavailable for this grid (rows {rows})")
    dt = op.dt
    log(f"[bench] tiles {eng.tiles()}; coefficients: {coef} (rows E/H {rows}); warmup {warmup} steps; "
        f"launches timed every {period} steps")
    # the warm-up runs with the timed region's launch timing, so that the
    # timing events' first use falls outside the timed region
    eng.set_profiling(period)
    if warmup:
        run_steps(eng, exts, wl, dt, 0, warmup)
    kernel_names = ("update_h:grid", "update_e:grid", "update_h:interior", "update_e:interior", "update_h:slabs",
                    "update_e:slabs", "fused:interior", "fused:grid", "point_ops", "energy")
    before = {k: eng.kernel_time(k) for k in kernel_names}
    with quiet_gc():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = run_steps(eng, exts, wl, dt, warmup, steps)
        t1 = time.perf_counter()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    eng.set_profiling(0)
    assert res.timesteps_executed == steps
    log(f"[bench] timed {steps} steps in {wall:.3f} s (batch {res.elapsed_seconds:.3f} s, "
        f"call {t1 - t0:.3f} s, device sync {wall - (t1 - t0):.4f} s)")

    kern = {}
    for name in kernel_names:  # launches timed inside the timed region only
        n, ms = eng.kernel_time(name)
        n, ms = n - before[name][0], ms - before[name][1]
        if n:
            kern[name] = (n, ms / n)
    yee = [k for k in kern if k.startswith("update_") or k.startswith("fused:")]
    dom = max(yee, key=lambda k: kern[k][0] * kern[k][1])
    avg_ms = kern[dom][1]
    kb = launch_bytes(dom, wl.shape, wl.cpml_layers, coef, rows)
    # the reference's layout and scheme: 12 coefficient arrays, two in-place
    # half-steps (a fused launch counts as both halves of its box)
    kb_ref = launch_bytes(dom.replace("fused:", "update_h:"), wl.shape, wl.cpml_layers, "arrays")
    if dom.startswith("fused:"):
        kb_ref *= 2
    achieved = kb / (avg_ms * 1e-3) / 1e9
    for k, (n, ms) in kern.items():
        log(f"[bench] {k}: {n} launches, mean {ms * 1e3:.1f} us")
    dom_key = f"{dom}@{wl.name}{wl.shape[0]}" + ("" if coef == "arrays" else "/table")
    E, H = eng.read_fields()
    finite = bool(np.isfinite(E).all() and np.isfinite(H).all())
    del E, H
    # the metric's own convention: 120 B per cell-update (the reference's 12
    # coefficient arrays, two half-steps) plus 32 B per CPML slab cell and
    # slab direction: a bandwidth in that layout's bytes, not a fraction of
    # HBM (the material table moves far fewer coefficient bytes)
    nx, ny, nz = wl.shape
    L = wl.cpml_layers
    conv_bytes = 120 * wl.cells + 32 * 2 * L * (ny * nz + nx * nz + nx * ny)
    step_bytes = (kernel_bytes(wl.shape, wl.cpml_layers, half_coef(coef, rows, True))
                  + kernel_bytes(wl.shape, wl.cpml_layers, half_coef(coef, rows, False)))
    if "fused:interior" in kern:
        step_bytes = (launch_bytes("fused:interior", wl.shape, wl.cpml_layers, coef, rows)
                      + launch_bytes("update_h:slabs", wl.shape, wl.cpml_layers, coef, rows)
                      + launch_bytes("update_e:slabs", wl.shape, wl.cpml_layers, coef, rows))
    elif "fused:grid" in kern:
        step_bytes = launch_bytes("fused:grid", wl.shape, wl.cpml_layers, coef, rows)
    step_gbs = step_bytes * steps / wall / 1e9
    config = {
        "workload": (f"plain_cavity: {wl.shape[0]}^3 FP32 Yee, 1 mm cells, vacuum, the reference's PEC faces "
                     f"(no boundary extension), soft Ez Gaussian fc=15 GHz at the centre, energy sampled every "
                     f"100 steps" if wl.name == "plain_cavity" else
                     f"cpml_boxes: {wl.shape[0]}x{wl.shape[1]}x{wl.shape[2]} FP32 Yee, 0.5 mm cells, "
                     f"{wl.cpml_layers}-cell CPML ({mode_name}) on all faces, {len(wl.boxes)} seeded dielectric "
                     f"boxes, soft Ez Gaussian fc=15 GHz, energy sampled every 100 steps"),
        "grid": list(wl.shape),
        "tiles": eng.tiles(),
        "coefficients": coef,
        "step_scheme": "fused" if any(k.startswith("fused:") for k in kern) else "two half-steps",
    }
    roofline = {
        "bound": "hbm",
        "kernel": {"update_h:interior": "k_update_h<4,AX_Z,phys> (interior box, z-face CPML)",
                   "update_e:interior": "k_update_e<4,AX_Z,phys> (interior box, z-face CPML)",
                   "fused:interior": "k_fused_rows<AX_Z,phys> (H and E of the interior box in one pass)"}.get(dom, dom),
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(dom_key),
        "bytes_per_launch": kb,
        "mean_launch_us": round(avg_ms * 1e3, 2),
        "launches_timed": kern[dom][0],
        # the same launch priced at the reference's 12-array layout (60 B per
        # cell): the bandwidth that layout would need for this launch time
        "bytes_per_launch_array_layout": kb_ref,
        "array_layout_equivalent_gbs": round(kb_ref / (avg_ms * 1e-3) / 1e9, 1),
    }
    eng.close()
    return {"config": config, "steps": steps, "ms_per_step": round(wall / steps * 1e3, 4),
            "mcells": wl.cells * steps / wall / 1e6, "hbm_frac_step": round(step_gbs / HBM_PEAK_GBS, 4),
            "step_gbs": step_gbs,
            "array_layout_equivalent_step_gbs": round(conv_bytes * steps / wall / 1e9, 1),
            "roofline": roofline, "fields_finite": finite}


def sweep_record(count, max_steps, autotune=True):
    """The microstrip sweep (see bench_sweep): wall time and throughput."""
    import numpy as np
    import torch

    from openems_amd import Sweep
    from openems_amd import workloads as W

    structs = W.microstrip_sweep(count, seed=4)
    log(f"[bench] sweep: building {len(structs)} structures, {sum(m.cells for m in structs) / 1e6:.1f} M cells")
    t_build = time.perf_counter()
    sw = Sweep()
    keep = []
    for ms in structs:
        op, eng, cp, port, exc = W.build_microstrip_hip(ms, max_steps)
        sw.add(eng, cp, port, exc)
        keep.append((ms, eng, port))
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build
    log(f"[bench] sweep: built in {t_build:.1f} s; running")
    with quiet_gc():
        t0 = time.perf_counter()
        res = sw.run(max_steps, energy_decay_db=50.0, check_interval=100, use_graph=True, autotune=autotune)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    steps = np.array([r.timesteps_executed for r in res])
    cells = np.array([ms.cells for ms, _, _ in keep])
    updates = float((steps * cells).sum())
    s11 = [port.s11(50.0) for _, _, port in keep]
    rec = {
        "workload": f"microstrip sweep: {len(structs)} structures (seed 4), {cells.sum() / 1e6:.1f} M cells, "
                    f"max {max_steps} steps, -50 dB decay (whole-run peak), CPML 8, 64-frequency port DFT",
        "wall_s": round(wall, 3),
        "build_s": round(t_build, 1),
        "cell_updates_per_s": round(updates / wall / 1e6, 1),
        "cell_updates_unit": "MCells/s",
        "steps_min_median_max": [int(steps.min()), int(np.median(steps)), int(steps.max())],
        "steps_total": int(steps.sum()),
        "decayed": int(sum(r.termination_reason == 1 for r in res)),
        "s11_finite": bool(all(np.isfinite(x).all() for x in s11)),
        "tiles": sw.tuning(),
    }
    log(f"[bench] sweep: {rec}")
    del sw
    for _, eng, _ in keep:
        eng.close()
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=256, help="grid edge of the CPML/boxes workload")
    ap.add_argument("--mode", choices=["physical", "compat"], default="physical")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--prof-period", type=int, default=0,
                    help="HIP events around every launch of every n-th timed step (1 = all; 0: every step up to "
                         "64 timed steps, then at most every 64th, so that the steps between replay captured graphs)")
    ap.add_argument("--no-subrecords", action="store_true",
                    help="skip the 512^3 perf point and the sweep that the default line carries beside the headline")
    ap.add_argument("--perf-point-steps", type=int, default=200)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--tiles", type=str, default="", help="bx,by,ci,vec override")
    ap.add_argument("--fused", action="store_true",
                    help="one fused pass per timestep over the CPML interior (slower than two half-steps on MI355X)")
    ap.add_argument("--coef", choices=["auto", "table", "arrays"], default="auto",
                    help="coefficient layout of the Yee updates (auto: the material table when the grid fits it)")
    ap.add_argument("--summarize-rocprof", nargs="+", metavar="DB",
                    help="print a markdown summary of rocprofv3 result databases and exit")
    ap.add_argument("--sweep", action="store_true", help="run the microstrip parameter sweep instead")
    ap.add_argument("--sweep-count", type=int, default=256)
    ap.add_argument("--sweep-max-steps", type=int, default=20000)
    ap.add_argument("--composite", action="store_true", help="run the 384x384x256 composite workload instead")
    ap.add_argument("--cavity", action="store_true",
                    help="the 100^3 plain vacuum cavity (the reference's published size) instead of the CPML workload")
    ap.add_argument("--nf2ff-stride", type=int, default=4)
    ap.add_argument("--learned", action="store_true", help="learned-stop vs plain sweep (trains the extrapolator)")
    ap.add_argument("--learned-train", type=int, default=4096)
    ap.add_argument("--learned-test", type=int, default=512)
    ap.add_argument("--learned-max-steps", type=int, default=20000)
    ap.add_argument("--learned-data-steps", type=int, default=12000)
    ap.add_argument("--learned-epochs", type=int, default=40)
    ap.add_argument("--learned-stride", type=int, default=1024)
    ap.add_argument("--nf2ff", action="store_true",
                    help="time the NF2FF transform of the 384x384x256 workload's box instead (steps = repetitions)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 10 device copies of 1 GiB (torch copy_), to scale HBM byte counters")
    args = ap.parse_args()

    if args.summarize_rocprof:
        print(summarize_rocprof(args.summarize_rocprof))
        return

    if args.copy_calibration:
        import torch

        a = torch.ones(1 << 28, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 10 copies of 1 GiB (read 1 GiB + write 1 GiB each)")
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the engine has no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.nf2ff:
        bench_nf2ff(args)
        return
    if args.sweep:
        bench_sweep(args)
        return
    if args.learned:
        bench_learned(args)
        return
    if args.composite:
        bench_composite(args)
        return

    import torch

    from openems_amd import MODE_COMPAT, MODE_PHYSICAL
    from openems_amd import workloads as W

    wl = W.plain_cavity(steps=args.steps) if args.cavity else W.cpml_boxes(n=args.n, steps=args.steps)
    mode = MODE_PHYSICAL if args.mode == "physical" else MODE_COMPAT
    # launches of every period-th step are timed with events; the steps in
    # between run from the engine's captured graphs (16 steps each)
    period = args.prof_period if args.prof_period > 0 else max(1, min(64, args.steps // 64))
    pt = yee_point(wl, mode, args.steps, args.warmup, period, args.tiles, args.coef, args.fused, args.mode)

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle, OpenMP) ...")
        cpu = cpu_baseline(wl, budget_s=args.cpu_budget)
        log(f"[bench] CPU baseline {cpu['value']:.1f} MCells/s")

    copy_gbs = stream_copy_gbs()
    log(f"[bench] stream copy (1 GiB device-to-device): {copy_gbs:.1f} GB/s")
    out = {
        "metric": METRIC,
        "value": round(pt["mcells"], 1),
        "unit": "MCells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": pt["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        # the reference's published 100^3 plain-Yee figure (its WGPU engine,
        # 1.39 ms per step = 718 Mcells/s, BENCHMARK_RESULTS.md:14-17): the
        # cavity workload only
        "vs_baseline": round(pt["mcells"] / REF_MCELLS_100, 2) if args.cavity else None,
        "dtype": "f32",
        "data": "synthetic",
        "config": pt["config"],
        "hbm_frac_step": pt["hbm_frac_step"],
        "array_layout_equivalent_step_gbs": pt["array_layout_equivalent_step_gbs"],
        "stream_copy": {"gbs": round(copy_gbs, 1), "bytes": 2 << 30,
                        "step_frac_of_copy": round(pt["step_gbs"] / copy_gbs, 4),
                        "kernel_frac_of_copy": round(pt["roofline"]["achieved"] / copy_gbs, 4)},
        "roofline": pt["roofline"],
        "fields_finite": pt["fields_finite"],
        "cpu_baseline": cpu,
    }
    if not args.no_subrecords and not args.cavity and args.n == 256:
        # the other figures of the metric beside the headline: the 512^3
        # perf point of the same workload, and the 256-structure sweep
        big = W.cpml_boxes(n=512, steps=args.perf_point_steps)
        p5 = yee_point(big, mode, args.perf_point_steps, 20, 4, "", "auto", False, args.mode)
        out["perf_point_512"] = {k: p5[k] for k in ("config", "steps", "ms_per_step", "mcells", "hbm_frac_step",
                                                    "array_layout_equivalent_step_gbs", "roofline",
                                                    "fields_finite")}
        out["perf_point_512"]["mcells"] = round(p5["mcells"], 1)
        del p5
        gc.collect()
        out["sweep"] = sweep_record(args.sweep_count, args.sweep_max_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
