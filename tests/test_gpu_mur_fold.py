"""First-order Mur on the z faces applied inside the E kernels, against the
oracle (which runs the pre, post and boundary steps as separate hooks) and
against the engine with the fold switched off (OE_FOLD_MUR=0).

Cases: both z faces with CPML on the x/y faces and a dielectric box;
sources on the i = 0 and j = 0 rows (which the E update skips) at the shift
planes, so the Mur write of those rows carries non-zero values; an ADE cell
on a Mur plane, where the engine must keep the separate Mur launches.
"""
import numpy as np
import pytest

from openems_amd import ADE_DRUDE, MODE_PHYSICAL
from tests.parity_util import Scenario, assert_parity, gauss_ref, run_pair, say

pytestmark = pytest.mark.gpu

SHAPE = (30, 26, 32)


def _scenario(name, ades=None):
    nz = SHAPE[2]
    return Scenario(name, SHAPE, 0.5e-3, 220, mur=(False, False, False, False, True, True),
                    cpml=dict(layers=(6, 6, 5, 7, 0, 0), mode=MODE_PHYSICAL),
                    materials=[(3.0, 1.0, 0.01, 0.0, (8, 22, 7, 19, 4, 28))],
                    excitations=[((15, 13, nz - 3), 2, True, gauss_ref(10e9)),
                                 ((0, 9, nz - 2), 0, True, gauss_ref(12e9)),
                                 ((11, 0, 1), 1, True, gauss_ref(9e9)),
                                 ((14, 12, 2), 0, True, gauss_ref(11e9))],
                    probes=[(15, 13, nz - 1), (0, 9, nz - 1)], ades=ades or [])


def test_folded_mur_z_faces(monkeypatch):
    out = []
    for fold in ("1", "0"):
        monkeypatch.setenv("OE_FOLD_MUR", fold)
        sc = _scenario(f"mur_fold_{fold}")
        res = run_pair(sc, batches=[57, 163])
        assert_parity(res, sc.name)
        E, H = res["eng"].read_fields()
        out.append(np.concatenate([E, H]))
        say(f"[mur] fold {fold}: |E| max on the k = nz-1 plane {np.abs(E[:, :, :, -1]).max():.3e}")
    assert np.array_equal(out[0], out[1])
    assert np.abs(out[0][:2, :, :, -1]).max() > 0 and np.abs(out[0][:2, :, :, 0]).max() > 0


def test_ade_on_mur_plane_keeps_separate_launches():
    nz = SHAPE[2]
    cells = np.array([[15, 13, nz - 2], [15, 14, nz - 2], [16, 13, nz - 1]], np.uint32)
    sc = _scenario("mur_ade_plane", ades=[(ADE_DRUDE, (1e11, 1e9, 0.0), cells, MODE_PHYSICAL)])
    res = run_pair(sc)
    assert_parity(res, sc.name)
    for g, o in zip(res["gpu_ade"], res["oracle_ade"]):
        assert np.array_equal(g, o.astype(np.float32))


def test_scalar_lanes_keep_separate_launches():
    """nz = 30 gives scalar lanes: the z-face Mur stays in its own launches
    and still matches the oracle."""
    sc = Scenario("mur_scalar_lanes", (22, 20, 30), 0.5e-3, 180, mur=(False, False, False, False, True, True),
                  excitations=[((11, 10, 27), 2, True, gauss_ref(10e9)), ((11, 10, 2), 0, True, gauss_ref(12e9))])
    assert_parity(run_pair(sc, batches=[41, 139]), sc.name)
