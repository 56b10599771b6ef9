"""Degenerate and ragged grid shapes against the oracle, bit for bit: single
cells and lines along each axis, nz below, at and just above one float4
lane, rows that leave most lanes of a wave idle, and odd sizes; with the
material table, graph replay and in-kernel sources in force (the defaults),
and once more with every one of them off."""
import pytest

from tests.parity_util import Scenario, assert_parity, gauss_ref, run_pair

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (1, 1, 8), (1, 7, 1), (9, 1, 1), (2, 3, 4), (3, 2, 5), (5, 4, 7), (4, 5, 3), (6, 6, 65),
          (7, 3, 132), (3, 70, 4), (33, 17, 9)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_shape_defaults(shape):
    nx, ny, nz = shape
    src = (nx // 2, ny // 2, nz // 2)
    mats = [(4.0, 2.0, 0.02, 0.0, (0, max(1, nx // 2), 0, ny, 0, max(1, nz // 2)))]
    sc = Scenario(f"shape_{nx}x{ny}x{nz}", shape, 1e-3, 40, materials=mats,
                  excitations=[(src, 2, True, gauss_ref(20e9)), (src, 0, False, gauss_ref(25e9))], probes=[src])
    assert_parity(run_pair(sc), sc.name)


@pytest.mark.parametrize("shape", [(1, 1, 1), (5, 4, 7), (6, 6, 65)], ids=lambda s: "x".join(map(str, s)))
def test_shape_all_switches_off(shape):
    nx, ny, nz = shape
    src = (nx // 2, ny // 2, nz // 2)
    sc = Scenario(f"shape_off_{nx}x{ny}x{nz}", shape, 1e-3, 40,
                  materials=[(4.0, 1.0, 0.0, 0.0, (0, nx, 0, max(1, ny // 2), 0, nz))],
                  excitations=[(src, 1, True, gauss_ref(20e9))], table=0, graph=0)
    assert_parity(run_pair(sc), sc.name)
