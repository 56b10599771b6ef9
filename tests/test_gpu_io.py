"""A configuration file drives the HIP engine, and field dumps carry the
device fields unchanged.

The XML below is read with its mesh; its boundary codes (PML on five faces,
Mur on +z) become a physical-mode CPML and a Mur extension.  The run must
match the oracle given the same grid and boundaries bit for bit, and the
OEMS and VTK dumps must hold exactly the fields read back from the device.
"""
import numpy as np
import pytest

from openems_amd import EngineBatch, HipEngine, MODE_PHYSICAL, ScheduledExcitation
from openems_amd.io import apply_boundaries, dump_fields, oems, operator_from_config, xml
from oracle import excitation as X
from oracle import oracle as O
from tests.parity_util import rel_l2, say

pytestmark = pytest.mark.gpu

CFG = """<openEMS>
  <FDTD NumberOfTimesteps="120" endCriteria="-50">
    <BoundaryCond xmin="3" xmax="3" ymin="3" ymax="3" zmin="3" zmax="2"></BoundaryCond>
  </FDTD>
  <ContinuousStructure>
    <RectilinearGrid DeltaUnit="0.0005">
      <XLines>{x}</XLines><YLines>{y}</YLines><ZLines>{z}</ZLines>
    </RectilinearGrid>
  </ContinuousStructure>
</openEMS>"""


def _text(n):
    return ",".join(str(v) for v in range(n + 1))


def test_config_drives_engine_and_dumps(tmp_path):
    cfg = xml.parse_config_str(CFG.format(x=_text(40), y=_text(36), z=_text(32)), read_mesh=True)
    op = operator_from_config(cfg)
    assert op.shape == (40, 36, 32)
    say(f"[io] config grid {op.shape}, boundaries {cfg.boundaries}, steps {cfg.num_timesteps}")
    eng = HipEngine(op)
    exts = apply_boundaries(eng, cfg)
    assert [e.kind for e in exts] == ["cpml", "mur"]
    # the oracle on the smallest mesh spacing per axis, as the operator takes it
    sim = O.OracleSim(40, 36, 32, op.delta, 32)
    assert sim.dt == op.dt
    sim.add_cpml([8, 8, 8, 8, 8, 0], 3.0, 1e-6, 0.0, 1.0, MODE_PHYSICAL)
    sim.add_mur([False] * 5 + [True])
    t0, tau = X.reference_gaussian_params(20e9, 0.5)
    smp = X.schedule(lambda t: X.gaussian(t, t0, tau), 0, cfg.num_timesteps, op.dt)
    pos = (20, 18, 16)
    eng.run_batch(EngineBatch(num_steps=cfg.num_timesteps, excitations=[ScheduledExcitation(pos, 2, smp, True)],
                              extensions=exts))
    sim.run_batch(cfg.num_timesteps, [(pos[0], pos[1], pos[2], 2, True)], smp[None].astype(np.float64))
    E, H = eng.read_fields()
    g = np.concatenate([E, H])
    o = sim.fields()
    say(f"[io] rel L2 vs oracle {rel_l2(g, o):.2e}")
    assert np.array_equal(g, o)

    dump_fields(eng, tmp_path / "f.oems", lines=[np.arange(n + 1) * 0.0005 for n in op.shape])
    r = oems.OemsReader.open(tmp_path / "f.oems")
    assert r.dimensions() == op.shape
    d = r.read_datasets()["vector"]
    for q in range(3):
        assert np.array_equal(d["E"][q], E[q].reshape(-1))
        assert np.array_equal(d["H"][q], H[q].reshape(-1))


def test_vtk_dump_of_device_fields(tmp_path):
    from openems_amd import Operator
    op = Operator.uniform(3, 2, 2, 1e-3)
    eng = HipEngine(op)
    rng = np.random.default_rng(0)
    E = rng.standard_normal((3, 3, 2, 2)).astype(np.float32)
    eng.write_fields(E=E)
    dump_fields(eng, tmp_path / "d", fmt="vtk")
    lines = (tmp_path / "d_E.vtk").read_text().splitlines()
    k = lines.index("VECTORS E float") + 1
    vals = np.array([[np.float32(v) for v in l.split()] for l in lines[k:]], np.float32)
    assert vals.shape == (12, 3)
    assert np.array_equal(vals.T, E.reshape(3, -1))
