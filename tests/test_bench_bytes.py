"""The bench's algorithmic byte model (the numerator of roofline.achieved),
checked against cell counts worked out by hand for the 256^3 workload with
an 8-cell CPML on every face (CPU only).

Interior box: 240 x 240 x 256 = 14,745,600 cells, of which 240 x 240 x 16 =
921,600 lie in the two z-face slabs (16 B of psi per cell and launch).
Per cell and half-step: 36 B of fields, plus the row index of the half's
material table (none for one row, 1 B up to 256 rows, 2 B above), or 24 B of
coefficient floats in the array layout.
"""
import pytest

import bench

SHAPE, L = (256, 256, 256), 8
INTERIOR, ZSLAB = 240 * 240 * 256, 240 * 240 * 16


def test_interior_launch_bytes_by_index_width():
    rows = (201, 1)  # the workload's E and H table sizes
    assert bench.launch_bytes("update_h:interior", SHAPE, L, "table", rows) == 36 * INTERIOR + 16 * ZSLAB
    assert bench.launch_bytes("update_e:interior", SHAPE, L, "table", rows) == 37 * INTERIOR + 16 * ZSLAB
    big = (300, 300)
    for half in ("update_h", "update_e"):
        assert bench.launch_bytes(f"{half}:interior", SHAPE, L, "table", big) == 38 * INTERIOR + 16 * ZSLAB
        assert bench.launch_bytes(f"{half}:interior", SHAPE, L, "arrays", rows) == 60 * INTERIOR + 16 * ZSLAB
    # the array-layout figure the profiles quote for the same launch
    assert bench.launch_bytes("update_e:interior", SHAPE, L, "arrays") == 899_481_600


def test_byte_index_switch(monkeypatch):
    monkeypatch.setenv("OE_BYTE_INDEX", "0")
    rows = (201, 1)
    assert bench.launch_bytes("update_e:interior", SHAPE, L, "table", rows) == 38 * INTERIOR + 16 * ZSLAB
    assert bench.launch_bytes("update_h:interior", SHAPE, L, "table", rows) == 36 * INTERIOR + 16 * ZSLAB


@pytest.mark.parametrize("rows", [(201, 1), (300, 300), (1, 1)])
@pytest.mark.parametrize("half", ["update_h", "update_e"])
def test_interior_plus_slabs_is_the_whole_half_step(rows, half):
    whole = bench.kernel_bytes(SHAPE, L, bench.half_coef("table", rows, half == "update_h"))
    parts = sum(bench.launch_bytes(f"{half}:{p}", SHAPE, L, "table", rows) for p in ("interior", "slabs"))
    assert parts == whole
    # every cell once, plus 16 B per slab cell and direction (8 layers deep
    # on both sides of each of the three axes)
    per_cell = 36 + {1: 0, 201: 1, 300: 2}[rows[1 if half == "update_h" else 0]]
    assert whole == per_cell * 256 ** 3 + 16 * 3 * 16 * 256 * 256


def test_fused_interior_bytes(monkeypatch):
    # 48 B of fields per cell and step, the E index as a byte (the default
    # fused kernel reads the byte-wide copy for 2 to 256 rows), no H index,
    # 32 B of psi per z-slab cell
    monkeypatch.delenv("OE_FUSED_KIND", raising=False)
    monkeypatch.delenv("OE_BYTE_INDEX", raising=False)
    assert bench.launch_bytes("fused:interior", SHAPE, L, "table", (201, 1)) == 49 * INTERIOR + 32 * ZSLAB
    assert 49 * INTERIOR + 32 * ZSLAB == 752_025_600
    # the LDS-exchange kernel reads the 16-bit index (766,771,200 B, what
    # bench.py --fused reported while it was the fused default)
    monkeypatch.setenv("OE_FUSED_KIND", "1")
    assert bench.launch_bytes("fused:interior", SHAPE, L, "table", (201, 1)) == 50 * INTERIOR + 32 * ZSLAB == 766_771_200
    monkeypatch.setenv("OE_FUSED_KIND", "2")
    monkeypatch.setenv("OE_BYTE_INDEX", "0")
    assert bench.launch_bytes("fused:interior", SHAPE, L, "table", (201, 1)) == 50 * INTERIOR + 32 * ZSLAB
    # more than 256 rows: 2 B whatever the kernel; an H table of several rows
    # adds its 2-byte index
    monkeypatch.delenv("OE_BYTE_INDEX")
    assert bench.launch_bytes("fused:interior", SHAPE, L, "table", (300, 1)) == 50 * INTERIOR + 32 * ZSLAB
    assert bench.launch_bytes("fused:interior", SHAPE, L, "table", (201, 5)) == 51 * INTERIOR + 32 * ZSLAB
    assert bench.launch_bytes("fused:interior", SHAPE, L, "table", (1, 1)) == 48 * INTERIOR + 32 * ZSLAB
    assert bench.launch_bytes("fused:interior", SHAPE, L, "arrays", (0, 0)) == 96 * INTERIOR + 32 * ZSLAB


def test_rocprof_summary_of_a_trace():
    """busy_summary and grid_split on a synthetic kernel trace: the union of
    the dispatch intervals, the gaps between them, and one row per grid size
    for a kernel launched with two."""
    import sqlite3
    con = sqlite3.connect(":memory:")
    con.execute("create table kernels(id, name, grid_size_x, grid_size_y, grid_size_z, start, end)")
    con.executemany("insert into kernels values (?,?,?,?,?,?,?)", [
        (1, "k_dft", 256, 1, 1, 0, 5_000),
        (2, "k_dft", 2_208_512, 1, 1, 10_000, 152_000),       # 5 us gap before
        (3, "k_update", 4_096, 1, 1, 100_000, 160_000),       # overlaps the last
        (4, "k_dft", 256, 1, 1, 2_160_000, 2_165_000),        # a 2 ms host phase before
    ])
    busy = bench.busy_summary(con.execute("select start, end from kernels order by start").fetchall())
    # busy: 5 + 150 + 5 us; span 2165 us; one 2 ms gap, one 5 us gap
    assert "busy 0.000 s" in busy and "Dispatches 4" in busy
    assert "gaps over 1 ms 0.002 s" in busy and "1 shorter gaps, 0.000 s" in busy
    assert "busy / (span - those gaps) 0.970" in busy  # 160 / (2165 - 2000)
    table = bench.grid_split(con)
    rows = [r for r in table.splitlines() if r.startswith("| `")]
    assert len(rows) == 2  # k_update has one grid size and is left out
    assert "| 2208512x1x1 | 1 | 142.0 | 142.00 |" in rows[0]
    assert "| 256x1x1 | 2 | 10.0 | 5.00 |" in rows[1]
