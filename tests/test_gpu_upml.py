"""GPU parity of the UPML (reference formulation) against the oracle:
fields and both flux arrays of every region bit-identical over the first
steps of the reference's PML scenario (before its instability overflows),
and over a batch split."""
import numpy as np
import pytest

from tests.parity_util import Scenario, assert_parity, gauss_ref, run_pair, say

pytestmark = pytest.mark.gpu


def _scenario(steps, layers=(8,) * 6):
    return Scenario("pml_absorbing_boundary (UPML)", (36, 36, 36), 0.5e-3, steps,
                    excitations=[((18, 18, 18), 2, True, gauss_ref(3e9))], upml=dict(layers=layers))


@pytest.mark.parametrize("layers", [(8,) * 6, (6, 4, 0, 5, 3, 7)])
def test_upml_fields_and_flux_bit_exact(layers):
    res = run_pair(_scenario(40, layers))
    assert_parity(res, f"upml {layers}")
    x = res["upml"]
    sim = res["sim"]
    regions = sim.upml_regions()
    assert len(regions) > 0
    for q in range(len(regions)):
        for name in ("vv", "vvfo", "vvfn", "ii", "iifo", "iifn", "volt_flux", "curr_flux"):
            for c in range(3):
                g = x.upml_array(q, name, c)
                o = sim.upml_array(q, name, c).astype(np.float32)
                assert np.array_equal(g, o), (q, name, c)


def test_upml_batch_split():
    res = run_pair(_scenario(40), batches=[17, 23])
    assert_parity(res, "upml 17+23")
