"""openEMS XML input (src/io/xml.rs:12-221).

`parse_config` follows the reference reader event for event:

  * the reference scans the file with quick-xml 0.37 (`trim_text(true)`,
    end-name checking on, empty elements NOT expanded): `<FDTD ...>` is a
    Start event, `<FDTD .../>` an Empty event.  Only Start events are acted on
    (src/io/xml.rs:24-37,127-131), so a self-closing `<FDTD/>`,
    `<BoundaryCond/>` or `<ContinuousStructure/>` is ignored, exactly as there;
  * NumberOfTimesteps, endCriteria, f0 and fc are read from the attributes of
    `<FDTD>` itself (src/io/xml.rs:98-123), with the reference's fallbacks:
    an unparsable NumberOfTimesteps gives 10000, an unparsable f0/fc gives 0,
    an unparsable endCriteria gives None, an absent attribute leaves the
    default (0, None, 0, 0).  endCriteria is kept as given; the reference
    treats it as a decay in dB (src/main.rs:138-141);
  * `<BoundaryCond xmin=.. zmax=..>` codes are i32 with fallback 0
    (src/io/xml.rs:146-188), mapped by `boundary_from_code`
    (src/io/xml.rs:212-221);
  * attribute values are the raw bytes between the quotes (quick-xml does not
    unescape `Attribute::value`), and of a repeated attribute the first wins
    (quick-xml reports the repeat as an error that `flatten()` drops);
  * numbers parse by Rust's `FromStr` rules (`parse_u64`, `parse_i32`,
    `parse_f64`): no surrounding whitespace, no digit separators;
  * `<ContinuousStructure>` is skipped (src/io/xml.rs:191-209): `grid` stays
    None and the excitation type stays Gaussian.

Beyond the reference: `parse_mesh` reads the `<RectilinearGrid>` mesh lines
(XLines/YLines/ZLines, DeltaUnit) that openEMS writes inside
ContinuousStructure, and `parse_config(..., read_mesh=True)` stores them in
`SimulationConfig.grid`, the slot the reference declares (src/io/xml.rs:60-61,
84-95) but never fills.

Malformed XML raises `XmlError` (src/lib.rs Error::Xml).  The reference
returns that error from its top-level loop; inside the FDTD and
ContinuousStructure loops it drops reader errors and keeps reading, which
this reader does not imitate.
"""
from __future__ import annotations

import enum
import math
import re
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, Iterator, List, Optional, Tuple, Union

__all__ = [
    "XmlError", "ExcitationType", "GridConfig", "SimulationConfig", "BoundaryCondition",
    "parse_config", "parse_config_str", "parse_mesh", "parse_mesh_str", "boundary_from_code",
    "parse_u64", "parse_i32", "parse_f64",
]


class XmlError(ValueError):
    """Error::Xml (src/lib.rs:61-88)."""


class ExcitationType(enum.Enum):
    """src/io/xml.rs:65-79."""
    GAUSSIAN = 0
    SINUSOIDAL = 1
    DIRAC = 2
    STEP = 3
    CUSTOM = 4


@dataclass
class GridConfig:
    """src/io/xml.rs:82-95."""
    x_lines: List[float] = field(default_factory=list)
    y_lines: List[float] = field(default_factory=list)
    z_lines: List[float] = field(default_factory=list)
    unit: float = 0.0


@dataclass
class SimulationConfig:
    """src/io/xml.rs:46-62 (Default derive: 0, None, Gaussian, 0, 0, [0; 6], None)."""
    num_timesteps: int = 0
    end_criteria_db: Optional[float] = None
    excitation_type: ExcitationType = ExcitationType.GAUSSIAN
    f0: float = 0.0
    fc: float = 0.0
    boundaries: List[int] = field(default_factory=lambda: [0] * 6)
    grid: Optional[GridConfig] = None


@dataclass(frozen=True)
class BoundaryCondition:
    """BoundaryCondition (src/fdtd/mod.rs:83-100): kind in {"pec", "pmc",
    "mur", "pml", "periodic"}; `layers` is set for "pml"."""
    kind: str
    layers: int = 0


def boundary_from_code(code: int) -> BoundaryCondition:
    """src/io/xml.rs:212-221."""
    return {
        0: BoundaryCondition("pec"),
        1: BoundaryCondition("pmc"),
        2: BoundaryCondition("mur"),
        3: BoundaryCondition("pml", 8),
        4: BoundaryCondition("periodic"),
    }.get(code, BoundaryCondition("pec"))


# ----------------------------------------------------- Rust FromStr rules
_U_RE = re.compile(r"\+?[0-9]+\Z")
_I_RE = re.compile(r"[+-]?[0-9]+\Z")
_F_RE = re.compile(r"[+-]?(?:[0-9]+(?:\.[0-9]*)?|\.[0-9]+)(?:[eE][+-]?[0-9]+)?\Z")
_F_SPECIAL = {"inf": math.inf, "infinity": math.inf, "nan": math.nan}


def parse_u64(s: str) -> Optional[int]:
    """`u64::from_str`: optional '+', decimal digits, value < 2^64."""
    if not _U_RE.match(s):
        return None
    v = int(s)
    return v if v < 1 << 64 else None


def parse_i32(s: str) -> Optional[int]:
    """`i32::from_str`: optional sign, decimal digits, -2^31 <= value < 2^31."""
    if not _I_RE.match(s):
        return None
    v = int(s)
    return v if -(1 << 31) <= v < 1 << 31 else None


def parse_f64(s: str) -> Optional[float]:
    """`f64::from_str`: decimal or exponent form, or inf/infinity/nan in any
    case, each with an optional sign; no whitespace, no '_'."""
    body = s[1:] if s[:1] in "+-" else s
    special = _F_SPECIAL.get(body.lower())
    if special is not None:
        return -special if s[:1] == "-" else special
    if not _F_RE.match(s):
        return None
    return float(s)


# ------------------------------------------------------------ event stream
_NAME_END = re.compile(r"[\s/>]")
_ATTR = re.compile(r"""\s*([^\s=/>"']+)\s*=\s*("([^"]*)"|'([^']*)')""")


def _attributes(body: str) -> Dict[str, str]:
    """Raw attribute values of a tag body (the text after the name); of a
    repeated key the first is kept."""
    out: Dict[str, str] = {}
    pos = 0
    while True:
        m = _ATTR.match(body, pos)
        if not m:
            break
        key = m.group(1)
        val = m.group(3) if m.group(3) is not None else m.group(4)
        out.setdefault(key, val)
        pos = m.end()
    return out


def _events(text: str) -> Iterator[Tuple[str, str, Dict[str, str]]]:
    """quick-xml 0.37 events that the reference distinguishes: ("start",
    name, attrs), ("empty", name, attrs), ("end", name, {}), then ("eof", "",
    {}).  Declarations, processing instructions, comments, CDATA, DOCTYPE and
    text produce no event.  A closing tag that does not match the open one
    raises (check_end_names), as does a tag or comment cut off by the end of
    the input; open elements left at the end do not."""
    stack: List[str] = []
    pos = 0
    n = len(text)
    while True:
        lt = text.find("<", pos)
        if lt < 0:
            yield ("eof", "", {})
            return
        if text.startswith("<!--", lt):
            end = text.find("-->", lt + 4)
            if end < 0:
                raise XmlError("XML parse error: unclosed comment")
            pos = end + 3
            continue
        if text.startswith("<![CDATA[", lt):
            end = text.find("]]>", lt + 9)
            if end < 0:
                raise XmlError("XML parse error: unclosed CDATA")
            pos = end + 3
            continue
        if text.startswith("<?", lt):
            end = text.find("?>", lt + 2)
            if end < 0:
                raise XmlError("XML parse error: unclosed processing instruction")
            pos = end + 2
            continue
        if text.startswith("<!", lt):
            depth, q = 0, lt + 2
            while q < n:
                c = text[q]
                if c == "[":
                    depth += 1
                elif c == "]":
                    depth -= 1
                elif c == ">" and depth <= 0:
                    break
                q += 1
            if q >= n:
                raise XmlError("XML parse error: unclosed DOCTYPE")
            pos = q + 1
            continue
        # element tag: find the closing '>' outside quotes
        q, quote = lt + 1, None
        while q < n:
            c = text[q]
            if quote:
                if c == quote:
                    quote = None
            elif c in "\"'":
                quote = c
            elif c == ">":
                break
            q += 1
        if q >= n:
            raise XmlError("XML parse error: unclosed tag")
        inner = text[lt + 1:q]
        pos = q + 1
        if inner.startswith("/"):
            name = inner[1:].rstrip()
            if not stack or stack[-1] != name:
                expected = stack[-1] if stack else ""
                raise XmlError(f"XML parse error: expecting </{expected}> but found </{name}>")
            stack.pop()
            yield ("end", name, {})
            continue
        empty = inner.endswith("/")
        if empty:
            inner = inner[:-1]
        m = _NAME_END.search(inner)
        name = inner if m is None else inner[:m.start()]
        if not name:
            raise XmlError("XML parse error: empty tag name")
        attrs = _attributes(inner[len(name):])
        if empty:
            yield ("empty", name, attrs)
        else:
            stack.append(name)
            yield ("start", name, attrs)


# ------------------------------------------------------------- the reader
_FDTD_ATTRS = ("NumberOfTimesteps", "endCriteria", "f0", "fc")
_BC_KEYS = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def _fdtd_settings(ev: Iterator, attrs: Dict[str, str], cfg: SimulationConfig) -> None:
    """parse_fdtd_settings (src/io/xml.rs:98-143)."""
    for key in _FDTD_ATTRS:
        if key not in attrs:
            continue
        val = attrs[key]
        if key == "NumberOfTimesteps":
            v = parse_u64(val)
            cfg.num_timesteps = 10000 if v is None else v
        elif key == "endCriteria":
            cfg.end_criteria_db = parse_f64(val)
        else:
            v = parse_f64(val)
            setattr(cfg, key, 0.0 if v is None else v)
    for kind, name, a in ev:
        if kind == "start" and name == "BoundaryCond":
            _boundary_cond(a, cfg)
        elif (kind == "end" and name == "FDTD") or kind == "eof":
            break


def _boundary_cond(attrs: Dict[str, str], cfg: SimulationConfig) -> None:
    """parse_boundary_cond (src/io/xml.rs:146-188)."""
    for q, key in enumerate(_BC_KEYS):
        if key in attrs:
            v = parse_i32(attrs[key])
            cfg.boundaries[q] = 0 if v is None else v


def _skip(ev: Iterator) -> None:
    """parse_geometry (src/io/xml.rs:191-209): skip to the matching end."""
    depth = 1
    for kind, _, _ in ev:
        if kind == "start":
            depth += 1
        elif kind == "end":
            depth -= 1
            if depth == 0:
                return
        elif kind == "eof":
            return


def parse_config_str(text: str, read_mesh: bool = False) -> SimulationConfig:
    """parse_config (src/io/xml.rs:12-43) on XML text."""
    cfg = SimulationConfig()
    ev = _events(text)
    for kind, name, attrs in ev:
        if kind == "eof":
            break
        if kind != "start":
            continue
        if name == "FDTD":
            _fdtd_settings(ev, attrs, cfg)
        elif name == "ContinuousStructure":
            _skip(ev)
    if read_mesh:
        cfg.grid = parse_mesh_str(text)
    return cfg


def _read_text(path: Union[str, Path]) -> str:
    try:
        return Path(path).read_bytes().decode("utf-8")
    except UnicodeDecodeError as e:
        raise XmlError(f"XML parse error: {e}") from e


def parse_config(path: Union[str, Path], read_mesh: bool = False) -> SimulationConfig:
    """parse_config (src/io/xml.rs:12-43).  A missing file raises OSError
    (Error::Io)."""
    return parse_config_str(_read_text(path), read_mesh)


# ---------------------------------------------------- mesh (no counterpart)
_LINES = {"XLines": "x_lines", "YLines": "y_lines", "ZLines": "z_lines"}


def parse_mesh_str(text: str) -> Optional[GridConfig]:
    """Mesh lines of the first `<RectilinearGrid>`: comma-separated numbers in
    XLines/YLines/ZLines (sorted, duplicates removed, as openEMS uses them),
    `unit` = DeltaUnit (1 when absent).  None when the file has no grid."""
    import xml.etree.ElementTree as ET
    try:
        root = ET.fromstring(text)
    except ET.ParseError as e:
        raise XmlError(f"XML parse error: {e}") from e
    node = root if root.tag == "RectilinearGrid" else root.find(".//RectilinearGrid")
    if node is None:
        return None
    unit = parse_f64(node.get("DeltaUnit", "1"))
    if unit is None:
        raise XmlError(f"XML parse error: DeltaUnit {node.get('DeltaUnit')!r} is not a number")
    grid = GridConfig(unit=unit)
    for tag, attr in _LINES.items():
        child = node.find(tag)
        vals: List[float] = []
        for tok in (child.text or "").split(",") if child is not None else []:
            tok = tok.strip()
            if not tok:
                continue
            v = parse_f64(tok)
            if v is None or not math.isfinite(v):
                raise XmlError(f"XML parse error: {tag} entry {tok!r} is not a finite number")
            vals.append(v)
        setattr(grid, attr, sorted(set(vals)))
    return grid


def parse_mesh(path: Union[str, Path]) -> Optional[GridConfig]:
    return parse_mesh_str(_read_text(path))
