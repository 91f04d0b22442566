"""Pins the oracle's geometry (rows A1-A5) before it is trusted as the checker.

* against the reference's own src/mesh.cpp, compiled in place into
  oracle/_ref/ref_mesh_dump (bit-for-bit comparison of every array
  DgSolver::initialize uploads, dg_solver.cpp:40-241);
* against the reference's unit-test KATs (tests/nekwave_test.cpp:65-129
  MeshQuadratureTest, :353-414 PeriodicBCTest);
* against closed forms (GLL p=4 nodes/weights, polynomial exactness, affine
  metrics of docs/method.md:227-236, dt of a uniform box).
"""
import math

import numpy as np
import pytest

import pyoracle as po
import refdump

ARRAYS = ("z", "w", "D", "Dt", "w3", "corners", "x", "y", "zc", "jac", "rx", "sx", "tx",
          "ry", "sy", "ty", "rz", "sz", "tz", "vM", "vP", "bc", "nx", "ny", "nz", "dA")

needs_ref = pytest.mark.skipif(not refdump.available(), reason="oracle/_ref not built (no reference tree)")


def _compare(m, ref, exact=True):
    for name in ARRAYS:
        a = getattr(m, name)
        b = ref[name]
        assert a.shape == b.shape, name
        if exact:
            assert np.array_equal(a, b), "%s differs: max |d| = %g" % (name, np.max(np.abs(a - b)))
        else:
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-14, err_msg=name)
    assert m.dxmin() == ref["dxmin"]
    assert m.auto_dt(1.0, 0.4) == ref["auto_dt"]


@needs_ref
@pytest.mark.parametrize("n1,nel,bounds,per", [
    (5, (2, 2, 2), (-1, 1, -1, 1, -1, 1), (0, 0, 0)),             # MeshQuadratureTest mesh
    (4, (2, 2, 2), (-2, 2, -1, 1, -0.5, 0.5), (1, 1, 0)),         # PeriodicBCTest mesh
    (4, (2, 2, 2), (-0.5, 0.5, -0.5, 0.5, -0.5, 0.5), (0, 0, 0)),  # MaxwellStabilityTest mesh
    (5, (4, 4, 4), (0, 1, 0, 1, 0, 1), (0, 0, 0)),                 # unit-cube PEC, 4^3, p=4
    (8, (3, 2, 2), (0, 1, 0, 1, 0, 1), (1, 1, 1)),                 # fully periodic, p=7
    (6, (3, 1, 2), (-1, 2, 0, 0.5, 1, 3), (0, 1, 0)),              # anisotropic box
])
def test_box_geometry_matches_reference_bitwise(tmp_path, n1, nel, bounds, per):
    m = po.Mesh.box(n1, nel, bounds, per)
    ref = refdump.box(tmp_path, n1, nel, bounds, per)
    _compare(m, ref)


@needs_ref
@pytest.mark.parametrize("n1", [2, 3, 4, 5, 6, 7, 8, 9, 10])
def test_gll_matches_reference_all_orders(tmp_path, n1):
    m = po.Mesh.box(n1, (1, 1, 1), (-1, 1, -1, 1, -1, 1))
    ref = refdump.box(tmp_path, n1, (1, 1, 1), (-1, 1, -1, 1, -1, 1))
    _compare(m, ref)


@needs_ref
@pytest.mark.parametrize("per", [(0, 0, 0), (1, 1, 1), (1, 0, 0)])
def test_jittered_trilinear_geometry_matches_reference_via_rea(tmp_path, per):
    """Jittered (trilinear) corners fed to the reference through its .rea loader:
    pins computeMetricsFromCorners, Nanson normals and the coordinate match."""
    n1 = 4
    m = po.Mesh.box(n1, (3, 3, 3), (0, 1, 0, 1, 0, 1), per, jitter=0.1, seed=2)
    ref = refdump.rea(tmp_path, n1, m.corners, m.f_kind, m.f_nbr_e, m.f_nbr_f)
    _compare(m, ref)


# ---------------------------------------------------------------- KATs ----

def test_mesh_quadrature_kat():
    """tests/nekwave_test.cpp:65-129, restated on the oracle."""
    N = 5
    m = po.Mesh.box(N, (2, 2, 2), (-1, 1, -1, 1, -1, 1))
    z, w, D = m.z, m.w, m.D
    assert abs(w.sum() - 2.0) <= 1e-12
    for i in range(N):
        assert abs(z[i] + z[N - 1 - i]) <= 1e-12
    for i in range(N):
        assert abs(sum(D[i + j * N] for j in range(N))) <= 1e-11
        assert abs(sum(D[i + j * N] * z[j] for j in range(N)) - 1.0) <= 1e-11
    jac = m.jac.reshape(m.nelt, N, N, N)
    W = np.einsum("i,j,k->kji", w, w, w)
    assert abs((jac * W).sum() - 8.0) <= 1e-10


def test_periodic_bc_kat():
    """tests/nekwave_test.cpp:353-414: 16 PERIODIC and 8 PEC faces, matched
    periodic nodes share transverse coordinates and differ by L."""
    N = 4
    b = (-2.0, 2.0, -1.0, 1.0, -0.5, 0.5)
    m = po.Mesh.box(N, (2, 2, 2), b, (1, 1, 0))
    kind, vM, vP = m.f_kind, m.vM, m.vP
    x, y, z = m.x, m.y, m.zc
    n2 = N * N
    nper = npec = 0
    for fi in range(m.nelt * 6):
        f = fi % 6
        pts = slice(fi * n2, (fi + 1) * n2)
        if kind[fi] == po.FACE_PERIODIC:
            nper += 1
            a, c = vM[pts], vP[pts]
            assert (c >= 0).all()
            if f in (1, 3):
                assert np.allclose(y[a], y[c], atol=1e-12, rtol=0)
                assert np.allclose(z[a], z[c], atol=1e-12, rtol=0)
                assert np.allclose(np.abs(x[a] - x[c]), 4.0, atol=1e-12, rtol=0)
            if f in (0, 2):
                assert np.allclose(x[a], x[c], atol=1e-12, rtol=0)
                assert np.allclose(z[a], z[c], atol=1e-12, rtol=0)
                assert np.allclose(np.abs(y[a] - y[c]), 2.0, atol=1e-12, rtol=0)
        elif kind[fi] == po.FACE_PEC:
            npec += 1
            assert f in (4, 5)
            assert (vP[pts] == -1).all()
    assert nper == 16 and npec == 8


# ---------------------------------------------------------- closed forms ----

def test_gll_p4_closed_form():
    z, w, D, _ = po.gll(5)
    s = math.sqrt(3.0 / 7.0)
    np.testing.assert_allclose(z, [-1, -s, 0, s, 1], atol=2e-16)
    np.testing.assert_allclose(w, [0.1, 49 / 90, 32 / 45, 49 / 90, 0.1], atol=4e-16)


@pytest.mark.parametrize("n1", [2, 3, 4, 5, 6, 7, 8, 9, 10])
def test_gll_exactness(n1):
    """Weights integrate degree <= 2p-1 exactly; D differentiates degree <= p exactly."""
    p = n1 - 1
    z, w, D, Dt = po.gll(n1)
    Dm = D.reshape(n1, n1).T  # D[i + j*n1] = D_ij -> row i, column j
    for deg in range(0, 2 * p):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, z ** deg) - exact) < 5e-14
    for deg in range(0, p + 1):
        d = Dm @ (z ** deg)
        ref = deg * z ** (deg - 1) if deg > 0 else np.zeros(n1)
        np.testing.assert_allclose(d, ref, atol=5e-12 * max(1, deg * deg))
    np.testing.assert_array_equal(Dt.reshape(n1, n1), D.reshape(n1, n1).T)
    # a plausible transposition mistake must fail: D applied the wrong way round
    if n1 > 2:
        bad = Dm.T @ (z ** 1)
        assert np.max(np.abs(bad - 1.0)) > 1e-3


def test_affine_metrics_closed_form():
    """docs/method.md:227-236: J = dx dy dz / 8, rx = 2/dx, sy = 2/dy, tz = 2/dz."""
    m = po.Mesh.box(4, (2, 3, 4), (0, 1, 0, 1.5, -1, 1))
    dx, dy, dz = 0.5, 0.5, 0.5
    np.testing.assert_allclose(m.jac, dx * dy * dz / 8, rtol=1e-14)
    np.testing.assert_allclose(m.rx, 2 / dx, rtol=1e-14)
    np.testing.assert_allclose(m.sy, 2 / dy, rtol=1e-14)
    np.testing.assert_allclose(m.tz, 2 / dz, rtol=1e-14)
    for a in (m.sx, m.tx, m.ry, m.ty, m.rz, m.sz):
        assert np.max(np.abs(a)) < 1e-14


def test_face_invariants_on_jittered_mesh():
    """|n| = 1, n+ = -n- and dA+ = dA- at matched points, sum dA = face area."""
    m = po.Mesh.box(5, (4, 4, 4), (0, 1, 0, 1, 0, 1), (1, 1, 1), jitter=0.1, seed=2)
    nx, ny, nz, dA, vM, vP = m.nx, m.ny, m.nz, m.dA, m.vM, m.vP
    np.testing.assert_allclose(nx ** 2 + ny ** 2 + nz ** 2, 1.0, atol=1e-14)
    # map node -> face point for the partner lookup
    n2 = m.n1 ** 2
    assert (vP >= 0).all()
    # partner face point index: face (e,f) <-> (nbr_e, nbr_f), (p,q) identity
    ne, nf = m.f_nbr_e, m.f_nbr_f
    fidx = np.arange(m.tfp) // n2
    pq = np.arange(m.tfp) % n2
    partner = (ne[fidx] * 6 + nf[fidx]) * n2 + pq
    np.testing.assert_array_equal(vM[partner], vP)
    np.testing.assert_allclose(nx[partner], -nx, atol=1e-13)
    np.testing.assert_allclose(ny[partner], -ny, atol=1e-13)
    np.testing.assert_allclose(nz[partner], -nz, atol=1e-13)
    np.testing.assert_allclose(dA[partner], dA, rtol=1e-12)
    # boundary faces of the unit cube are planar; each x-min face of the
    # 4x4 array sums to total area 1 over the whole x = 0 plane.
    kind = m.f_kind
    tot = 0.0
    for fi in range(m.nelt * 6):
        e, f = divmod(fi, 6)
        if f == 3 and e % 4 == 0:
            tot += dA[fi * n2:(fi + 1) * n2].sum()
    assert abs(tot - 1.0) < 1e-12
    assert kind.size == m.nelt * 6


def test_dt_uniform_box_closed_form():
    """Unit cube 4^3, p=4: dxmin = (1 - sqrt(3/7)) h / 2 with h = 1/4; CFL 0.5."""
    m = po.Mesh.box(5, (4, 4, 4), (0, 1, 0, 1, 0, 1))
    h = 0.25
    dxmin = (1 - math.sqrt(3 / 7)) * h / 2
    assert abs(m.dxmin() - dxmin) < 1e-15
    assert abs(0.5 * m.dxmin() - 0.0215841) < 5e-8
    assert abs(po.lib().or_auto_cfl(0.4) - 0.4 * 1.75 / math.sqrt(3)) < 1e-16


def test_rng_matches_std_mt19937_64():
    """C++ [rand.predef]: the 10000th output of a default-seeded mt19937_64 is
    9981545732273789042."""
    r = po.Rng(5489)
    for _ in range(9999):
        r.next()
    assert r.next() == 9981545732273789042
