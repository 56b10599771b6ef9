"""Field dumps in the reference's file layout (src/io/hdf5.rs:36-466).

The reference's default build (Cargo features `default = []`, HDF5 behind
the optional `hdf5` feature) writes its "OEMS" binary layout from
`Hdf5Writer::close` (src/io/hdf5.rs:361-440) and reads it back with
`Hdf5Reader::open` (src/io/hdf5.rs:53-144).  This module writes that layout
byte for byte; all integers and floats are little-endian:

    b"OEMS", version bytes 1 0 0 0, mesh type i32 (0 Cartesian, 1 Cylindrical)
    3 x { u64 count, count x f64 }         mesh lines x, y, z (3 zero counts
                                             when no mesh was set)
    u64 0                                  timesteps (always empty)
    u64 0                                  frequencies (always empty)
    u64 n_scalar,  n_scalar  x { u64 name_len, name (UTF-8), u64 n, n x f32 }
    u64 n_vector,  n_vector  x { u64 name_len, name, 3 x { u64 n, n x f32 } }
    u64 n_complex, n_complex x { u64 name_len, name, u64 n, n x (f64 re, f64 im) }

Attributes and the current group are kept by the writer but not written,
as in the reference.  `OemsReader.open` reads what the reference's reader
reads (header, mesh, timesteps, frequencies) and refuses anything else with
the reference's message; `read_datasets` also returns the datasets, which
the reference's reader skips.  HDF5 files themselves are out of reach: the
image has no HDF5 library.
"""
from __future__ import annotations

import enum
import struct
from dataclasses import dataclass, field
from pathlib import Path
from typing import BinaryIO, Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

MAGIC = b"OEMS"
VERSION = bytes([1, 0, 0, 0])


class MeshType(enum.IntEnum):
    """src/io/hdf5.rs:15-32: any code other than 1 reads as Cartesian."""
    CARTESIAN = 0
    CYLINDRICAL = 1

    @classmethod
    def from_code(cls, v: int) -> "MeshType":
        return cls.CYLINDRICAL if v == 1 else cls.CARTESIAN


def _u64(v: int) -> bytes:
    return struct.pack("<Q", v)


def _name(f: BinaryIO, name: str) -> None:
    b = name.encode("utf-8")
    f.write(_u64(len(b)))
    f.write(b)


class OemsWriter:
    """Hdf5Writer (src/io/hdf5.rs:239-466) writing the binary layout."""

    def __init__(self, path: Union[str, Path]):
        self.path = str(path)
        self.current_group = "/"
        self.mesh_lines: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None
        self.mesh_type = MeshType.CARTESIAN
        self.scalar_fields: List[Tuple[str, np.ndarray]] = []
        self.vector_fields: List[Tuple[str, Tuple[np.ndarray, np.ndarray, np.ndarray]]] = []
        self.complex_fields: List[Tuple[str, np.ndarray]] = []
        self.attributes: Dict[str, np.ndarray] = {}

    def set_current_group(self, group: str) -> None:
        self.current_group = group

    def write_rect_mesh(self, lines: Sequence[Sequence[float]], mesh_type: MeshType = MeshType.CARTESIAN) -> None:
        if len(lines) != 3:
            raise ValueError("write_rect_mesh needs x, y and z lines")
        self.mesh_lines = tuple(np.array(l, dtype=np.float64).reshape(-1) for l in lines)
        self.mesh_type = MeshType(mesh_type)

    def write_scalar_field(self, name: str, data) -> None:
        """f32 data; float64 input is rounded to f32 as `write_scalar_field_f64` does."""
        self.scalar_fields.append((name, np.array(data, dtype=np.float32).reshape(-1)))

    write_scalar_field_f64 = write_scalar_field

    def write_vector_field(self, name: str, field_xyz) -> None:
        """Three f32 component arrays, each written in its own z-fastest order."""
        if len(field_xyz) != 3:
            raise ValueError("a vector field has three components")
        self.vector_fields.append((name, tuple(np.array(c, dtype=np.float32).reshape(-1) for c in field_xyz)))

    def write_complex_field(self, name: str, data) -> None:
        self.complex_fields.append((name, np.array(data, dtype=np.complex128).reshape(-1)))

    def write_attribute(self, name: str, values) -> None:
        self.attributes[name] = np.array(values, dtype=np.float64).reshape(-1)

    def close(self) -> None:
        """write_binary (src/io/hdf5.rs:361-440)."""
        with open(self.path, "wb") as f:
            f.write(MAGIC)
            f.write(VERSION)
            f.write(struct.pack("<i", int(self.mesh_type)))
            if self.mesh_lines is not None:
                for l in self.mesh_lines:
                    f.write(_u64(l.size))
                    f.write(l.astype("<f8").tobytes())
            else:
                for _ in range(3):
                    f.write(_u64(0))
            f.write(_u64(0))
            f.write(_u64(0))
            f.write(_u64(len(self.scalar_fields)))
            for name, data in self.scalar_fields:
                _name(f, name)
                f.write(_u64(data.size))
                data.astype("<f4").tofile(f)
            f.write(_u64(len(self.vector_fields)))
            for name, comps in self.vector_fields:
                _name(f, name)
                for c in comps:
                    f.write(_u64(c.size))
                    c.astype("<f4").tofile(f)
            f.write(_u64(len(self.complex_fields)))
            for name, data in self.complex_fields:
                _name(f, name)
                f.write(_u64(data.size))
                data.astype("<c16").tofile(f)


@dataclass
class OemsReader:
    """Hdf5Reader (src/io/hdf5.rs:35-237)."""
    path: str
    mesh_lines: Tuple[np.ndarray, np.ndarray, np.ndarray]
    mesh_type: MeshType
    timesteps: np.ndarray
    frequencies: np.ndarray
    attributes: Dict[str, np.ndarray] = field(default_factory=dict)
    _end: int = 0

    @classmethod
    def open(cls, path: Union[str, Path]) -> "OemsReader":
        """Hdf5Reader::open: the binary layout, else the reference's
        "HDF5 support not enabled" error (src/io/hdf5.rs:53-73)."""
        try:
            return cls._read_header(str(path))
        except (ValueError, EOFError) as e:
            raise OSError("HDF5 support not enabled. Enable the 'hdf5' feature or use binary format.") from e

    @classmethod
    def _read_header(cls, path: str) -> "OemsReader":
        with open(path, "rb") as f:
            if _exact(f, 4) != MAGIC:
                raise ValueError("Invalid file format")
            _exact(f, 4)                                   # version, not checked
            mesh_type = MeshType.from_code(struct.unpack("<i", _exact(f, 4))[0])
            lines = tuple(_f64s(f) for _ in range(3))
            ts = _f64s(f)
            fr = _f64s(f)
            end = f.tell()
        return cls(path, lines, mesh_type, ts, fr, {}, end)

    def num_timesteps(self) -> int:
        return len(self.timesteps)

    def num_frequencies(self) -> int:
        return len(self.frequencies)

    def is_valid(self) -> bool:
        return all(len(l) > 0 for l in self.mesh_lines)

    def dimensions(self) -> Tuple[int, int, int]:
        """Cells per axis: lines - 1, at least 1 (src/io/hdf5.rs:227-236)."""
        return tuple(max(len(l) - 1, 1) for l in self.mesh_lines)

    def read_datasets(self) -> Dict[str, Dict[str, object]]:
        """The scalar, vector and complex datasets after the header (no
        counterpart in the reference's reader): {"scalar": {name: f32[n]},
        "vector": {name: (f32[n],) * 3}, "complex": {name: c128[n]}}."""
        out: Dict[str, Dict[str, object]] = {"scalar": {}, "vector": {}, "complex": {}}
        with open(self.path, "rb") as f:
            f.seek(self._end)
            for _ in range(_count(f)):
                name = _str(f)
                out["scalar"][name] = _array(f, "<f4", _count(f))
            for _ in range(_count(f)):
                name = _str(f)
                out["vector"][name] = tuple(_array(f, "<f4", _count(f)) for _ in range(3))
            for _ in range(_count(f)):
                name = _str(f)
                out["complex"][name] = _array(f, "<c16", _count(f))
            if f.read(1):
                raise ValueError("trailing bytes after the complex fields")
        return out


def _exact(f: BinaryIO, n: int) -> bytes:
    b = f.read(n)
    if len(b) != n:
        raise EOFError("unexpected end of file")
    return b


def _count(f: BinaryIO) -> int:
    return struct.unpack("<Q", _exact(f, 8))[0]


def _str(f: BinaryIO) -> str:
    return _array(f, "u1", _count(f)).tobytes().decode("utf-8")


def _array(f: BinaryIO, dtype: str, n: int) -> np.ndarray:
    size = np.dtype(dtype).itemsize * n
    here = f.tell()
    left = f.seek(0, 2) - here
    f.seek(here)
    if size > left:
        raise EOFError("unexpected end of file")
    return np.frombuffer(_exact(f, size), dtype=dtype).copy()


def _f64s(f: BinaryIO) -> np.ndarray:
    return _array(f, "<f8", _count(f))
