"""File input and output (src/io/mod.rs): openEMS XML configuration
(`xml`), field dumps in the reference's binary layout (`oems`) and legacy
VTK text (`vtk`), plus the glue from a parsed configuration to an engine.

Only `dump_fields` and `operator_from_config` touch the GPU library; the
readers and writers are plain host code.
"""
from __future__ import annotations

from pathlib import Path
from typing import Optional, Sequence, Union

import numpy as np

from . import oems, vtk, xml  # noqa: F401
from .oems import MeshType, OemsReader, OemsWriter  # noqa: F401
from .vtk import write_field_vtk, write_vector_field_vtk  # noqa: F401
from .xml import (  # noqa: F401
    BoundaryCondition, GridConfig, SimulationConfig, XmlError, boundary_from_code, parse_config, parse_mesh,
)

DEFAULT_GRID = (20, 20, 20, 1e-3)
"""The grid the reference's command line falls back to when the file gives
none (src/main.rs:117-128): Grid::uniform(20, 20, 20, 1 mm)."""


def grid_lines(cfg: SimulationConfig) -> Sequence[np.ndarray]:
    """Mesh lines in metres: the configuration's grid scaled by its unit, or
    the command line's 20^3 default."""
    if cfg.grid is None:
        n, d = DEFAULT_GRID[0], DEFAULT_GRID[3]
        return [np.arange(n + 1) * d for _ in range(3)]
    g = cfg.grid
    unit = g.unit if g.unit else 1.0
    return [np.asarray(l, np.float64) * unit for l in (g.x_lines, g.y_lines, g.z_lines)]


def operator_from_config(cfg: SimulationConfig):
    """Operator::new on the configuration's grid (src/main.rs:117-132)."""
    from ..engine import Operator
    return Operator.from_lines(*grid_lines(cfg))


def apply_boundaries(eng, cfg: SimulationConfig, pml: str = "cpml"):
    """Device extensions for the configuration's boundary codes.  The
    reference parses the codes but never builds boundaries from them
    (Operator::new stores BoundaryConditions unused, src/fdtd/operator.rs:59-77);
    this maps them onto the engine's extensions: PEC needs none (the engine's
    fixed boundary planes), Mur faces share one Mur extension, PML faces
    share one CPML (physical mode, the reference's CPML defaults) or, with
    pml="upml", the reference's UPML (src/extensions/pml.rs, unstable).
    PMC and periodic faces raise NotImplementedError: no engine of the
    reference implements them.  Returns the list of extensions."""
    from ..engine import MODE_PHYSICAL, CpmlBoundaries
    kinds = [boundary_from_code(c) for c in cfg.boundaries]
    bad = [b.kind for b in kinds if b.kind in ("pmc", "periodic")]
    if bad:
        raise NotImplementedError(f"boundary kinds {sorted(set(bad))} are not implemented")
    if pml not in ("cpml", "upml"):
        raise ValueError("pml must be 'cpml' or 'upml'")
    exts = []
    layers = [b.layers if b.kind == "pml" else 0 for b in kinds]
    if any(layers):
        if pml == "cpml":
            exts.append(eng.cpml(CpmlBoundaries(layers=layers, mode=MODE_PHYSICAL)))
        else:
            exts.append(eng.upml(layers=layers))
    mur = [b.kind == "mur" for b in kinds]
    if any(mur):
        exts.append(eng.mur(mur))
    return exts


def dump_fields(eng, path: Union[str, Path], lines: Optional[Sequence[Sequence[float]]] = None,
                fmt: str = "oems") -> None:
    """Write the engine's E and H (read back from the device) as vector
    fields "E" and "H": `fmt` "oems" writes the reference's binary layout
    with the mesh lines (metres; uniform lines from the engine's cell size
    when None), "vtk" writes two legacy VTK files `<path>_E.vtk` and
    `<path>_H.vtk`."""
    E, H = eng.read_fields()
    if lines is None:
        lines = [np.arange(n + 1) * d for n, d in zip(eng.shape, eng.delta)]
    if fmt == "oems":
        w = OemsWriter(path)
        w.write_rect_mesh(lines)
        w.write_vector_field("E", E)
        w.write_vector_field("H", H)
        w.close()
    elif fmt == "vtk":
        p = str(path)
        write_vector_field_vtk(p + "_E.vtk", E, lines, "E")
        write_vector_field_vtk(p + "_H.vtk", H, lines, "H")
    else:
        raise ValueError("fmt must be 'oems' or 'vtk'")


def dump_phasors(ext, path: Union[str, Path], freqs: Sequence[float],
                 lines: Optional[Sequence[Sequence[float]]] = None) -> None:
    """Write a DFT extension's accumulated phasors (one complex field per
    frequency, named "f<index>", signals in the extension's term order) and
    the frequencies as an attribute, in the reference's binary layout."""
    vals = ext.dft_values()
    w = OemsWriter(path)
    if lines is not None:
        w.write_rect_mesh(lines)
    for q in range(vals.shape[0]):
        w.write_complex_field(f"f{q}", vals[q])
    w.write_attribute("frequency", freqs)
    w.close()
