"""Legacy ASCII VTK dumps (src/io/vtk.rs:10-126).

The text is the reference's, character for character: numbers are printed
as Rust's `{}` prints them (the shortest decimal that reads back to the same
f32 for field values, or f64 for mesh lines, in positional notation:
`1`, `0.1`, `0.0000001`, `-0`, `NaN`, `inf`), mesh lines end with a space
before the newline, and the cell values follow the z-fastest order of the
arrays (i outer, k inner).  As in the reference, DIMENSIONS and the
coordinate counts are the field's cells + 1 while the coordinate lines list
every mesh line the grid holds.

Formatting runs per value in Python (about 1 µs each): a 100^3 field takes
about a second.
"""
from __future__ import annotations

import math
from pathlib import Path
from typing import Sequence, Union

import numpy as np


def _fmt(x, dtype) -> str:
    v = dtype(x)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    return np.format_float_positional(v, unique=True, trim="-")


def rust_display_f32(x) -> str:
    """`format!("{}", x as f32)`."""
    return _fmt(x, np.float32)


def rust_display_f64(x) -> str:
    """`format!("{}", x as f64)`."""
    return _fmt(x, np.float64)


def _header(f, title: str, dims, lines: Sequence[Sequence[float]]) -> None:
    nx, ny, nz = dims
    f.write("# vtk DataFile Version 3.0\n")
    f.write(f"{title}\n")
    f.write("ASCII\n")
    f.write("DATASET RECTILINEAR_GRID\n")
    f.write(f"DIMENSIONS {nx + 1} {ny + 1} {nz + 1}\n")
    for axis, n, l in zip("XYZ", dims, lines):
        f.write(f"{axis}_COORDINATES {n + 1} float\n")
        f.write("".join(rust_display_f64(x) + " " for x in l))
        f.write("\n")


def _cells(a) -> np.ndarray:
    a = np.asarray(a, dtype=np.float32)
    if a.ndim != 3:
        raise ValueError("a field is a [nx, ny, nz] array")
    return a


def write_field_vtk(path: Union[str, Path], field, lines: Sequence[Sequence[float]], field_name: str) -> None:
    """write_field_vtk (src/io/vtk.rs:10-62): `field` is [nx, ny, nz];
    `lines` are the grid's x, y and z mesh lines."""
    a = _cells(field)
    with open(path, "w", newline="\n") as f:
        _header(f, "openEMS field data", a.shape, lines)
        f.write(f"CELL_DATA {a.size}\n")
        f.write(f"SCALARS {field_name} float 1\n")
        f.write("LOOKUP_TABLE default\n")
        f.write("".join(rust_display_f32(v) + "\n" for v in a.reshape(-1)))


def write_vector_field_vtk(path: Union[str, Path], field_xyz, lines: Sequence[Sequence[float]],
                           field_name: str) -> None:
    """write_vector_field_vtk (src/io/vtk.rs:65-126): `field_xyz` holds the
    x, y and z components, each [nx, ny, nz]."""
    comps = [_cells(c) for c in field_xyz]
    if len(comps) != 3 or any(c.shape != comps[0].shape for c in comps):
        raise ValueError("a vector field has three components of one shape")
    with open(path, "w", newline="\n") as f:
        _header(f, "openEMS vector field data", comps[0].shape, lines)
        f.write(f"CELL_DATA {comps[0].size}\n")
        f.write(f"VECTORS {field_name} float\n")
        x, y, z = (c.reshape(-1) for c in comps)
        f.write("".join(f"{rust_display_f32(a)} {rust_display_f32(b)} {rust_display_f32(c)}\n"
                        for a, b, c in zip(x, y, z)))
