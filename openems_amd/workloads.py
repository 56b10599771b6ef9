"""Synthetic workloads of the FDTD timestep path, shared by bench.py and tests.

plain_cavity   100^3 grid, 1 mm cells, vacuum, reference boundary semantics
               ("PEC" on all faces), soft Ez Gaussian (openEMS f0 = 0,
               fc = 15 GHz) at the centre, 1000 steps, field probe 20 cells
               off-centre along x.
cpml_boxes     256^3 grid, 0.5 mm cells, 8-cell CPML on every face (grading
               m = 3, R0 = 1e-6), 200 seeded axis-aligned dielectric boxes
               (sides U{5..40} cells, eps_r U[1,10], sigma U[0,0.05] S/m,
               placed inside the CPML-free interior, applied in order), same
               source at the centre, 4000 steps.

Both operator kinds (openems_amd.Operator and the test oracle) expose
set_material(eps_r, mu_r, sigma_e, sigma_m, region), so a workload is applied
to either through apply_materials().
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np


@dataclass
class Workload:
    name: str
    shape: Tuple[int, int, int]
    delta: float
    steps: int
    source: Tuple[int, int, int]
    source_dir: int
    fc: float
    probe: Tuple[int, int, int]
    cpml_layers: int = 0
    cpml_grading: float = 3.0
    cpml_R0: float = 1e-6
    boxes: List[Tuple[float, float, float, float, Tuple[int, int, int, int, int, int]]] = field(default_factory=list)

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))

    def gaussian_t0_tau(self):
        """openEMS Gaussian (f0 = 0, cutoff fc) as the reference's Gaussian{t0, tau}."""
        return 9.0 / (2.0 * math.pi * self.fc), 3.0 / (2.0 * math.pi * self.fc)


def plain_cavity(n: int = 100, steps: int = 1000) -> Workload:
    c = n // 2
    return Workload("plain_cavity", (n, n, n), 1e-3, steps, (c, c, c), 2, 15e9, (c + 20 if c + 20 < n else n - 1, c, c))


def seeded_boxes(n: int, layers: int, count: int, seed: int):
    """count boxes: (eps_r, mu_r, sigma_e, sigma_m, region) inside [layers, n - layers)."""
    rng = np.random.default_rng(seed)
    lo, hi = layers, n - layers
    out = []
    for _ in range(count):
        side = rng.integers(5, 41, size=3)
        side = np.minimum(side, hi - lo)
        org = [int(rng.integers(lo, hi - side[a] + 1)) for a in range(3)]
        eps = float(rng.uniform(1.0, 10.0))
        sig = float(rng.uniform(0.0, 0.05))
        reg = (org[0], org[0] + int(side[0]), org[1], org[1] + int(side[1]), org[2], org[2] + int(side[2]))
        out.append((eps, 1.0, sig, 0.0, reg))
    return out


def cpml_boxes(n: int = 256, steps: int = 4000, count: int = 200, seed: int = 2, layers: int = 8) -> Workload:
    c = n // 2
    return Workload("cpml_boxes", (n, n, n), 0.5e-3, steps, (c, c, c), 2, 15e9, (c + 20 if c + 20 < n else n - 1, c, c),
                    cpml_layers=layers, boxes=seeded_boxes(n, layers, count, seed))


def apply_materials(op, wl: Workload):
    for eps, mu, se, sm, reg in wl.boxes:
        op.set_material(eps, mu, se, sm, reg)


@dataclass
class PortSpheres:
    """The composite workload: CPML on five faces and first-order Mur on +z,
    a 50-ohm z-directed lumped port (1 x 1 x 4 cells) driven by an openEMS
    Gaussian, seeded Drude spheres (omega_p = 1e15 rad/s, gamma = 1e13 /s)
    and Lorentz spheres (eps_inf = 2 as the background material of the
    sphere, delta_eps = 3, omega_0 = 2 pi 10 GHz, gamma = 1e9 /s) in the
    physical ADE mode, port V/I running DFT over 64 frequencies (1-20 GHz)
    and an NF2FF recording box 10 cells inside the CPML at 5, 10 and 15 GHz.
    The cell size is not fixed by the configuration; 0.5 mm is used."""
    shape: Tuple[int, int, int]
    delta: float
    steps: int
    layers: int
    port_region: Tuple[int, int, int, int, int, int]
    port_R: float
    fc: float
    spheres: List[Tuple[str, Tuple[int, int, int], int]]  # ("drude" | "lorentz", centre, radius)
    dft_freqs: np.ndarray
    nf2ff_box: Tuple[int, int, int, int, int, int]
    nf2ff_freqs: Tuple[float, ...]

    DRUDE = (1e15, 1e13)
    LORENTZ = (2 * math.pi * 10e9, 1e9, 3.0)
    LORENTZ_EPS_INF = 2.0

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))

    @property
    def cpml_layers(self):
        return (self.layers,) * 5 + (0,)

    @property
    def mur_faces(self):
        return (False,) * 5 + (True,)

    def sphere_cells(self, centre, radius) -> np.ndarray:
        c = np.asarray(centre)
        r = int(radius)
        g = np.mgrid[-r:r + 1, -r:r + 1, -r:r + 1].reshape(3, -1).T
        g = g[(g * g).sum(1) <= r * r]
        return (g + c).astype(np.int64)

    def columns(self, cells: np.ndarray):
        """z-runs (i, j, k0, k1) covering the cells (one set_material box each)."""
        out = []
        order = np.lexsort((cells[:, 2], cells[:, 1], cells[:, 0]))
        cs = cells[order]
        start = 0
        for q in range(1, len(cs) + 1):
            if q == len(cs) or cs[q, 0] != cs[start, 0] or cs[q, 1] != cs[start, 1] or cs[q, 2] != cs[q - 1, 2] + 1:
                out.append((int(cs[start, 0]), int(cs[start, 1]), int(cs[start, 2]), int(cs[q - 1, 2]) + 1))
                start = q
        return out

    def excitation_samples(self, dt: float, n: int) -> np.ndarray:
        """openEMS Gaussian (f0 = 0, cutoff fc) sampled at s * dt, float32."""
        t0, tau = 9.0 / (2.0 * math.pi * self.fc), 3.0 / (2.0 * math.pi * self.fc)
        t = np.arange(n, dtype=np.float64) * dt
        return np.exp(-((t - t0) / tau) ** 2).astype(np.float32)


def port_spheres(shape=(384, 384, 256), steps: int = 2000, n_spheres: int = 8, rmin: int = 4, rmax: int = 12,
                 seed: int = 3, layers: int = 8, delta: float = 0.5e-3) -> PortSpheres:
    nx, ny, nz = shape
    cx, cy, cz = nx // 2, ny // 2, nz // 2
    port = (cx, cx + 1, cy, cy + 1, cz - 2, cz + 2)
    rng = np.random.default_rng(seed)
    lo = layers + 10
    spheres = []
    placed = [((cx, cy, cz), 4)]  # keep clear of the port
    for kind in ("drude", "lorentz"):
        count = 0
        tries = 0
        while count < n_spheres and tries < 10000:
            tries += 1
            r = int(rng.integers(rmin, rmax + 1))
            c = tuple(int(rng.integers(lo + r, s - lo - r)) for s in shape)
            if any(sum((a - b) ** 2 for a, b in zip(c, pc)) <= (r + pr + 2) ** 2 for pc, pr in placed):
                continue
            placed.append((c, r))
            spheres.append((kind, c, r))
            count += 1
    box = (lo, nx - lo, lo, ny - lo, lo, nz - lo)
    return PortSpheres(shape, delta, steps, layers, port, 50.0, 20e9, spheres, np.linspace(1e9, 20e9, 64), box,
                       (5e9, 10e9, 15e9))


def build_port_spheres_hip(wl: PortSpheres, precision: int = 0, nf2ff_stride: int = 1, record_steps: int = 0):
    """Operator, engine and extensions of the composite workload on the HIP
    engine.  Returns (op, eng, exts dict)."""
    from . import ADE_DRUDE, ADE_LORENTZ, MODE_PHYSICAL, CpmlBoundaries, HipEngine, Operator

    op = Operator.uniform(*wl.shape, wl.delta)
    ade_cells = {"drude": [], "lorentz": []}
    for kind, c, r in wl.spheres:
        cells = wl.sphere_cells(c, r)
        ade_cells[kind].append(cells)
        if kind == "lorentz":
            for i, j, k0, k1 in wl.columns(cells):
                op.set_material(wl.LORENTZ_EPS_INF, 1.0, 0.0, 0.0, (i, i + 1, j, j + 1, k0, k1))
    op.add_lumped_element(2, wl.port_region, wl.port_R)
    eng = HipEngine(op)
    exts = {"cpml": eng.cpml(CpmlBoundaries(layers=wl.cpml_layers, mode=MODE_PHYSICAL)),
            "mur": eng.mur(wl.mur_faces)}
    if ade_cells["drude"]:
        exts["drude"] = eng.ade(ADE_DRUDE, wl.DRUDE, np.concatenate(ade_cells["drude"]), MODE_PHYSICAL)
    if ade_cells["lorentz"]:
        exts["lorentz"] = eng.ade(ADE_LORENTZ, wl.LORENTZ, np.concatenate(ade_cells["lorentz"]), MODE_PHYSICAL)
    exts["port"] = eng.lumped_port(2, wl.port_region, wl.dft_freqs, precision=precision, record_steps=record_steps)
    exts["nf2ff"] = eng.nf2ff_box(wl.nf2ff_box, wl.nf2ff_freqs, precision=precision, time_stride=nf2ff_stride)
    return op, eng, exts


@dataclass
class Microstrip:
    """One structure of the microstrip sweep, in cells of 0.5 mm: PEC ground
    slab at k = 8 (just inside the -z CPML), substrate of eps_r over
    k = 9 .. 9 + h, PEC trace polyline one cell thick at k = 9 + h, 8-cell
    CPML on all faces, a 50-ohm z-directed lumped port from the ground to the
    trace start, driven by an openEMS Gaussian (f0 = 0, fc = 20 GHz)."""
    shape: Tuple[int, int, int]
    eps_r: float
    h: int
    width: int
    segments: List[Tuple[int, int, int, int]]  # trace boxes (i0, i1, j0, j1)
    port: Tuple[int, int]                      # (i, j) of the port edge column
    delta: float = 0.5e-3
    layers: int = 8
    ground_k: int = 8
    R: float = 50.0
    fc: float = 20e9

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))

    @property
    def port_region(self):
        i, j = self.port
        return (i, i + 1, j, j + 1, self.ground_k + 1, self.ground_k + 1 + self.h)

    def excitation_samples(self, dt: float, n: int) -> np.ndarray:
        t0, tau = 9.0 / (2.0 * math.pi * self.fc), 3.0 / (2.0 * math.pi * self.fc)
        t = np.arange(n, dtype=np.float64) * dt
        return np.exp(-((t - t0) / tau) ** 2).astype(np.float32)

    def apply(self, op):
        """Ground, substrate, trace and port resistor on an operator (the HIP
        Operator or the oracle, whose lumped call is add_lumped)."""
        nx, ny, nz = self.shape
        g = self.ground_k
        op.set_pec((0, nx, 0, ny, g, g + 1))
        op.set_material(self.eps_r, 1.0, 0.0, 0.0, (0, nx, 0, ny, g + 1, g + 1 + self.h))
        for i0, i1, j0, j1 in self.segments:
            op.set_pec((i0, i1, j0, j1, g + 1 + self.h, g + 2 + self.h))
        lumped = getattr(op, "add_lumped_element", None) or getattr(op, "add_lumped")
        lumped(2, self.port_region, self.R)


def microstrip(rng: np.random.Generator, layers: int = 8) -> Microstrip:
    """Grid U{60..220} x U{60..120} x U{30..48} cells (nz rounded up to a
    multiple of 4 so that every row holds whole float4 lanes), eps_r
    U[2.2, 10.2], substrate 1-4 mm, trace width 1-4 mm, 2-6 Manhattan
    segments kept 6 cells clear of the CPML."""
    nx = int(rng.integers(60, 221))
    ny = int(rng.integers(60, 121))
    nz = int(rng.integers(30, 49))
    nz = (nz + 3) // 4 * 4
    eps = float(rng.uniform(2.2, 10.2))
    h = int(rng.integers(2, 9))
    w = int(rng.integers(2, 9))
    lo = layers + 6
    hi = (nx - lo - w, ny - lo - w)
    x, y = int(rng.integers(lo, hi[0])), int(rng.integers(lo, hi[1]))
    port = (x + w // 2, y + w // 2)  # middle of the trace's first cell row
    segs = []
    n_seg = int(rng.integers(2, 7))
    along_x = bool(rng.integers(0, 2))
    for _ in range(n_seg):
        if along_x:
            nxp = int(rng.integers(lo, hi[0]))
            a, b = sorted((x, nxp))
            segs.append((a, b + w, y, y + w))
            x = nxp
        else:
            nyp = int(rng.integers(lo, hi[1]))
            a, b = sorted((y, nyp))
            segs.append((x, x + w, a, b + w))
            y = nyp
        along_x = not along_x
    return Microstrip((nx, ny, nz), eps, h, w, segs, port)


def microstrip_sweep(count: int = 256, seed: int = 4) -> List[Microstrip]:
    rng = np.random.default_rng(seed)
    return [microstrip(rng) for _ in range(count)]


def build_microstrip_hip(ms: Microstrip, max_steps: int, precision: int = 0, record_steps: int = 0,
                         freqs=None):
    """(op, eng, cpml, port, excitations) of one sweep structure on the HIP engine."""
    from . import MODE_PHYSICAL, CpmlBoundaries, HipEngine, Operator
    from .engine import port_excitations

    op = Operator.uniform(*ms.shape, ms.delta)
    ms.apply(op)
    eng = HipEngine(op)
    cp = eng.cpml(CpmlBoundaries(layers=(ms.layers,) * 6, mode=MODE_PHYSICAL))
    fr = np.linspace(1e9, 20e9, 64) if freqs is None else freqs
    port = eng.lumped_port(2, ms.port_region, fr, precision=precision, record_steps=record_steps)
    exc = port_excitations(2, ms.port_region, ms.excitation_samples(op.dt, max_steps))
    return op, eng, cp, port, exc
