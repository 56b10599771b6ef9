"""Point sources added by the E kernels (no extension between the E update
and the sources) against the oracle, bit for bit: several sources on one
cell in list order (soft then hard, the hard one wins), sources in k = 0
cells and at the top of a float4 lane, both scalar and float4 lanes, with
and without CPML; and a batch with a source on an i = 0 plane, which keeps
the separate point kernel."""
import pytest

from openems_amd import MODE_PHYSICAL
from tests.parity_util import Scenario, assert_parity, gauss_ref, run_pair

pytestmark = pytest.mark.gpu


def _sources(i0=10):
    g1, g2 = gauss_ref(10e9), gauss_ref(14e9)
    return [((i0, 10, 11), 2, True, g1), ((i0, 10, 11), 2, False, g2), ((i0, 10, 11), 2, True, g2),
            ((12, 9, 0), 0, True, g1), ((14, 15, 15), 1, True, g2), ((16, 8, 4), 2, True, g1)]


@pytest.mark.parametrize("shape,cpml", [((28, 24, 32), None), ((30, 26, 32), dict(layers=(6,) * 6, mode=MODE_PHYSICAL)),
                                        ((26, 22, 30), None)])
def test_sources_in_e_kernel(shape, cpml):
    sc = Scenario(f"sources_{shape[2]}_{'cpml' if cpml else 'plain'}", shape, 1e-3, 120, cpml=cpml,
                  materials=[(3.0, 1.0, 0.0, 0.0, (5, 20, 5, 18, 5, 25))], excitations=_sources(),
                  probes=[(12, 9, 0), (10, 10, 11)])
    assert_parity(run_pair(sc), sc.name)


def test_source_on_boundary_plane_keeps_point_kernel():
    sc = Scenario("sources_boundary", (28, 24, 32), 1e-3, 100, excitations=_sources() + [((0, 10, 11), 1, True,
                                                                                        gauss_ref(10e9))])
    assert_parity(run_pair(sc), sc.name)
