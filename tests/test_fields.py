"""Field snapshots (nw_fields_*, the files Hdf5Writer writes without the HDF5
library: hdf5_writer.cpp:98-119 .pvd, :121-470 .vtu, :711-733 and :1002-1033
.bin).  Where the reference tree is present, the reference's own writer
(oracle/_ref/ref_vtk_dump, src/hdf5_writer.cpp compiled in place) writes the
same mesh and states and every file must be byte-identical.  Without it, the
files are checked for their structure and values."""
import os
import struct
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import nekwave_amd as nw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_VTK = os.path.join(ROOT, "oracle", "_ref", "ref_vtk_dump")

CASES = [  # n1, nel, bounds, periodic, continuous, save fields
    (3, (2, 2, 1), (0, 1, 0, 1, 0, 0.5), (0, 0, 0), False, ["E", "H"]),
    (4, (2, 1, 2), (-1, 1, -1, 1, -1, 1), (1, 0, 0), False,
     ["E", "H", "magnitude_E", "magnitude_H", "energy_density"]),
    (3, (2, 2, 2), (0, 2, 0, 1, 0, 1), (0, 0, 0), True, ["E", "H", "energy_density"]),
    (4, (1, 2, 1), (0, 1, 0, 1, 0, 1), (0, 1, 0), True, ["H", "magnitude_E"]),
]


def _states(npts, seed):
    rng = np.random.default_rng(seed)
    out = []
    for step, t in ((0, 0.0), (10, 0.0125), (20, 0.025)):
        st = rng.standard_normal(6 * npts) * 10.0 ** rng.integers(-3, 3, 6 * npts)
        out.append((step, t, st))
    return out


def _product(tmp, n1, nel, bounds, per, cont, fields, states):
    m = nw.Mesh(n1)
    m.createBoxMesh(*nel, *bounds, *[bool(p) for p in per])
    w = nw.FieldWriter()
    w.setFieldSaveOptions(fields)
    w.initialize(os.path.join(tmp, "fields.h5"), m, True, cont)
    for step, t, st in states:
        w.writeStep(step, t, st)
    w.close()
    return m


def _reference(tmp, n1, nel, bounds, per, cont, fields, states):
    sp = os.path.join(tmp, "states.bin")
    with open(sp, "wb") as fh:
        fh.write(struct.pack("<i", len(states)))
        for step, t, st in states:
            fh.write(struct.pack("<id", step, t) + np.ascontiguousarray(st, "<f8").tobytes())
    flags = sum(nw.FieldWriter.FLAGS[f] for f in fields)
    args = [REF_VTK, str(n1)] + [str(v) for v in nel] + [repr(float(b)) for b in bounds] + \
        [str(int(p)) for p in per] + [str(int(cont)), str(flags), sp, os.path.join(tmp, "fields.h5")]
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


@pytest.mark.skipif(not os.path.exists(REF_VTK), reason="reference tree absent: oracle/_ref/ref_vtk_dump not built")
@pytest.mark.parametrize("case", range(len(CASES)))
def test_files_equal_the_reference_writer(tmp_path, case):
    n1, nel, bounds, per, cont, fields = CASES[case]
    npts = nel[0] * nel[1] * nel[2] * n1 ** 3
    states = _states(npts, case)
    a, b = tmp_path / "product", tmp_path / "reference"
    a.mkdir()
    b.mkdir()
    _product(str(a), n1, nel, bounds, per, cont, fields, states)
    _reference(str(b), n1, nel, bounds, per, cont, fields, states)
    names = sorted(os.listdir(b))
    names.remove("states.bin")
    assert sorted(os.listdir(a)) == names
    assert {"fields.pvd", "fields.bin", "fields_step_0.vtu", "fields_step_10.vtu", "fields_step_20.vtu"} <= set(names)
    for n in names:
        assert (a / n).read_bytes() == (b / n).read_bytes(), n


def test_vtu_pvd_bin_structure(tmp_path):
    """Without the reference: the archive holds the exact states, the .vtu holds
    them to the 6 significant digits of the default stream format, the cells
    are the (n1-1)^3 sub-hexahedra of each element, the .pvd lists the steps."""
    n1, nel = 3, (2, 1, 1)
    m = nw.Mesh(n1)
    m.createBoxMesh(*nel, 0, 2, 0, 1, 0, 1)
    npts = m.getTotalPoints()
    states = _states(npts, 7)
    w = nw.FieldWriter()
    w.initialize(str(tmp_path / "snap.h5"), m)
    for step, t, st in states:
        w.writeStep(step, t, st)
    w.close()
    raw = (tmp_path / "snap.bin").read_bytes()
    assert raw[:16] == b"NEKWAVE_BIN_V2\0\0"
    assert struct.unpack_from("<4i", raw, 16) == (n1, 2, npts, 0)
    x, y, z = m.coords()
    xyz = np.frombuffer(raw, "<f8", 3 * npts, 32).reshape(-1, 3)
    assert np.array_equal(xyz, np.stack([x, y, z], 1))
    off = 32 + 24 * npts
    for step, t, st in states:
        assert struct.unpack_from("<2i", raw, off) == (step, 0)
        assert struct.unpack_from("<d", raw, off + 8)[0] == t
        E = np.frombuffer(raw, "<f8", 3 * npts, off + 16).reshape(-1, 3)
        H = np.frombuffer(raw, "<f8", 3 * npts, off + 16 + 24 * npts).reshape(-1, 3)
        assert np.array_equal(E, st[:3 * npts].reshape(3, -1).T) and np.array_equal(H, st[3 * npts:].reshape(3, -1).T)
        assert struct.unpack_from("<i", raw, off + 16 + 48 * npts)[0] == 0
        off += 16 + 48 * npts + 4
    assert off == len(raw)
    pvd = ET.parse(tmp_path / "snap.pvd").getroot()
    ds = pvd.find("Collection").findall("DataSet")
    assert [d.get("file") for d in ds] == ["snap_step_%d.vtu" % s for s, _, _ in states]
    assert [float(d.get("timestep")) for d in ds] == [t for _, t, _ in states]
    vtu = ET.parse(tmp_path / "snap_step_20.vtu").getroot()
    piece = vtu.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == npts and int(piece.get("NumberOfCells")) == 2 * (n1 - 1) ** 3
    arrays = {a.get("Name"): a for a in piece.iter("DataArray")}
    st = states[-1][2]
    e = np.array(arrays["E"].text.split(), float).reshape(-1, 3)
    assert np.allclose(e, st[:3 * npts].reshape(3, -1).T, rtol=5e-6, atol=0)
    conn = np.array(arrays["connectivity"].text.split(), int).reshape(-1, 8)
    assert conn[0].tolist() == [0, 1, 4, 3, 9, 10, 13, 12]  # i + 3 (j + 3 k) corners in VTK order
    assert conn.max() == npts - 1
    assert arrays["types"].text.split() == ["12"] * conn.shape[0]


def test_fields_write_rejects_a_wrong_size(tmp_path):
    m = nw.Mesh(3)
    m.createBoxMesh(1, 1, 1)
    w = nw.FieldWriter()
    w.initialize(str(tmp_path / "f.h5"), m)
    with pytest.raises(nw.NekWaveError):
        w.writeStep(0, 0.0, np.zeros(6 * 27 - 1))
    w.close()
