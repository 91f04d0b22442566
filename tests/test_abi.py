"""The C-ABI library: loads, exports every symbol include/nekwave_amd.h declares,
and its host-side mesh code (no GPU calls) agrees with the oracle."""
import os
import re

import numpy as np
import pytest

import nekwave_amd as nw
import pyoracle as po
from nekwave_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "nekwave_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nw_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 35
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python signatures cover the header
    assert set(syms) <= set(_lib.SIGNATURES), set(syms) - set(_lib.SIGNATURES)
    assert L.nw_abi_version() == 1


@pytest.mark.parametrize("n1", [2, 3, 5, 8, 10])
def test_product_gll_matches_oracle(n1):
    m = nw.Mesh(n1)
    m.createBoxMesh(1, 1, 1)
    z, w, D, _ = po.gll(n1)
    np.testing.assert_array_equal(m.getGllZ(), z)
    np.testing.assert_array_equal(m.getGllW(), w)
    np.testing.assert_array_equal(m.getD(), D)


@pytest.mark.parametrize("args", [
    dict(n1=5, nel=(4, 4, 4), b=(0, 1, 0, 1, 0, 1), per=(0, 0, 0), jit=0.0, seed=0),
    dict(n1=4, nel=(3, 2, 3), b=(-1, 2, 0, 1, 0, 0.5), per=(1, 1, 0), jit=0.1, seed=2),
    dict(n1=8, nel=(2, 2, 2), b=(0, 1, 0, 1, 0, 1), per=(1, 1, 1), jit=0.1, seed=2),
])
def test_product_box_mesh_matches_oracle(args):
    m = nw.Mesh(args["n1"])
    m.createBoxMesh(*args["nel"], *args["b"], *[bool(p) for p in args["per"]], jitter=args["jit"], seed=args["seed"])
    o = po.Mesh.box(args["n1"], args["nel"], args["b"], args["per"], jitter=args["jit"], seed=args["seed"])
    np.testing.assert_array_equal(m.getElementCorners().reshape(-1), o.corners)
    x, y, z = m.coords()
    np.testing.assert_allclose(x, o.x, atol=1e-15)
    np.testing.assert_allclose(y, o.y, atol=1e-15)
    np.testing.assert_allclose(z, o.zc, atol=1e-15)
    assert abs(m.computeMinNodeDistance() - o.dxmin()) < 1e-15
    assert abs(m.computeAutomaticDt(1.0, 0.4) - o.auto_dt(1.0, 0.4)) < 1e-15
    info = m.info()
    assert info.npts == o.npts and info.tfp == o.tfp and info.nelt == o.nelt
    assert info.all_affine == (1 if args["jit"] == 0 else 0)


def test_from_corners_matches_box_connectivity():
    o = po.Mesh.box(4, (3, 3, 2), (0, 1, 0, 1, 0, 1), (0, 0, 0), jitter=0.1, seed=7)
    m = nw.Mesh.fromCorners(4, o.corners, o.f_kind, o.f_nbr_e, o.f_nbr_f)
    assert m.getTotalPoints() == o.npts


def test_error_paths_do_not_abort():
    """Failure handling forced from the host: bad sizes, inverted elements,
    non-conforming faces, unknown kinds come back as status codes."""
    with pytest.raises(nw.NekWaveError) as e:
        nw.Mesh(11).createBoxMesh(1, 1, 1)
    assert e.value.status == _lib.NW_ERR_ARG
    c = po.Mesh.box(3, (1, 1, 1), (0, 1, 0, 1, 0, 1), (0, 0, 0)).corners.reshape(8, 3)
    inv = c[[1, 0, 3, 2, 5, 4, 7, 6]]  # mirrored: negative Jacobian
    k = np.full(6, nw.FACE_PEC); ne = np.full(6, -1); nf = np.full(6, -1)
    with pytest.raises(nw.NekWaveError) as e:
        nw.Mesh.fromCorners(3, inv, k, ne, nf)
    assert e.value.status == _lib.NW_ERR_MESH
    # two unit cubes that do not share a face but claim to be E neighbours
    a = c
    b = c + np.array([5.0, 0, 0])
    k2 = np.full(12, nw.FACE_PEC); ne2 = np.full(12, -1); nf2 = np.full(12, -1)
    k2[1] = nw.FACE_E; ne2[1] = 1; nf2[1] = 3
    with pytest.raises(nw.NekWaveError) as e:
        nw.Mesh.fromCorners(3, np.concatenate([a, b]), k2, ne2, nf2)
    assert e.value.status == _lib.NW_ERR_MESH
    with pytest.raises(nw.NekWaveError):
        nw.Mesh.fromCorners(3, c, np.full(6, 9), ne, nf)
    m = nw.Mesh(3)
    m.createBoxMesh(2, 1, 1)
    with pytest.raises(nw.NekWaveError):
        m.setMaterials(eps=np.array([1.0, -2.0]))


def test_solver_calls_before_initialize_raise():
    s = nw.DgSolver()
    with pytest.raises(nw.NekWaveError) as e:
        s.step(0.1)
    assert e.value.status == _lib.NW_ERR_STATE


def test_cpp_host_interface_compiles_and_links(tmp_path):
    """include/nekwave_amd.hpp (the reference's Mesh/DgSolver/Case signatures)
    builds against the library with the host compiler."""
    from nekwave_amd import build as b

    out = b.build_cpp_kats(str(tmp_path / "kats"))
    assert os.path.exists(out)
