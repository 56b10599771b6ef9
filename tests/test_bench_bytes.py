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


def test_fused_interior_bytes():
    # 48 B of fields per cell and step, the E index (2 B: the fused pass reads
    # the 16-bit form), no H index, 32 B of psi per z-slab cell; 766,771,200
    # is what bench.py --fused reported for this launch
    assert bench.launch_bytes("fused:interior", SHAPE, L, "table", (201, 1)) == 50 * INTERIOR + 32 * ZSLAB
    assert 50 * INTERIOR + 32 * ZSLAB == 766_771_200
