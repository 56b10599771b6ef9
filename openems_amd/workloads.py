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
