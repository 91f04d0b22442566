"""Case field snapshots on the device path: with export_fields on, the Case
writes <output_dir>/fields.{bin,pvd} and fields_step_<n>.vtu at t = 0, every
save_frequency steps and at the last step (case.cpp:438-464, 603-625), with
the state downloaded for those steps only; the archive's last record is the
final state exactly."""
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from nekwave_amd import Case, Config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_case_writes_snapshots_at_save_frequency(tmp_path, precision):
    out = str(tmp_path / "run")
    c = Case(Config(order=4, nelx=2, nely=2, nelz=1, xmin=0, xmax=1, ymin=0, ymax=1, zmin=0, zmax=0.5,
                    num_steps=7, output_frequency=3, save_frequency=2, export_fields=True, output_dir=out,
                    precision=precision))
    c.setInitialConditionHook(lambda x, y, z: (0 * x, 0 * x, np.sin(np.pi * x) * np.sin(np.pi * y), 0 * x, 0 * x,
                                               0 * x))
    c.preprocess()
    st0 = c.state.copy()
    c.simulate()
    c.postprocess()
    steps = [0, 2, 4, 6, 7]
    ds = ET.parse(os.path.join(out, "fields.pvd")).getroot().find("Collection").findall("DataSet")
    assert [d.get("file") for d in ds] == ["fields_step_%d.vtu" % s for s in steps]
    for s in steps:
        assert os.path.exists(os.path.join(out, "fields_step_%d.vtu" % s))
    npts = c.mesh.getTotalPoints()
    raw = open(os.path.join(out, "fields.bin"), "rb").read()
    off = 32 + 24 * npts
    recs = []
    while off < len(raw):
        step, _ = struct.unpack_from("<2i", raw, off)
        t = struct.unpack_from("<d", raw, off + 8)[0]
        E = np.frombuffer(raw, "<f8", 3 * npts, off + 16).reshape(-1, 3).T
        H = np.frombuffer(raw, "<f8", 3 * npts, off + 16 + 24 * npts).reshape(-1, 3).T
        recs.append((step, t, np.concatenate([E.ravel(), H.ravel()])))
        off += 16 + 48 * npts + 4
    assert [r[0] for r in recs] == steps
    assert recs[0][1] == 0.0 and np.array_equal(recs[0][2], st0)
    assert recs[-1][1] == pytest.approx(7 * c.dt, rel=1e-14)
    assert np.array_equal(recs[-1][2], c.state)
