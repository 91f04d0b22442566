"""Probe placement and probe configuration on the host (ProbeManager::init,
probe.cpp:33-116; Case::initProbes, case.cpp:381-424; config.hpp:291-322, 473-490).

The placement is pinned by what the geometry fixes: on an affine box mesh a
point at known element coordinates must come back in that element with those
(r, s, t); points on shared faces go to the lowest-index element (strict '<' in
probe.cpp:84); points outside the box are clamped to its surface.
"""
import numpy as np
import pytest

import nekwave_amd as nw
from nekwave_amd.case import Case, Config


def _box(n1=5, nel=(3, 2, 4), bounds=(0.0, 1.5, -1.0, 1.0, 0.0, 2.0)):
    m = nw.Mesh(n1)
    m.createBoxMesh(*nel, *bounds)
    return m


def _elem_index(ex, ey, ez, nel):
    return ex + nel[0] * (ey + nel[1] * ez)


def test_locate_interior_points_known_element_coordinates():
    nel, b = (3, 2, 4), (0.0, 1.5, -1.0, 1.0, 0.0, 2.0)
    m = _box(5, nel, b)
    rng = np.random.default_rng(3)
    hx = (b[1] - b[0]) / nel[0]
    hy = (b[3] - b[2]) / nel[1]
    hz = (b[5] - b[4]) / nel[2]
    cases = []
    for _ in range(40):
        e3 = [int(rng.integers(0, n)) for n in nel]
        rst = rng.uniform(-0.95, 0.95, 3)
        p = (b[0] + (e3[0] + 0.5 * (rst[0] + 1)) * hx, b[2] + (e3[1] + 0.5 * (rst[1] + 1)) * hy,
             b[4] + (e3[2] + 0.5 * (rst[2] + 1)) * hz)
        cases.append((p, _elem_index(*e3, nel), rst))
    # the element ordering of the box mesh must agree with its corners
    corners = m.getElementCorners()
    pts = np.array([c[0] for c in cases])
    elem, rst = m.locateProbes(pts)
    for q, (p, e, r) in enumerate(cases):
        assert elem[q] == e, (p, elem[q], e)
        lo, hi = corners[e].min(axis=0), corners[e].max(axis=0)
        assert np.all(lo <= np.array(p)) and np.all(np.array(p) <= hi)
        np.testing.assert_allclose(rst[q], r, rtol=0, atol=1e-13)


def test_locate_shared_face_goes_to_first_element_and_outside_is_clamped():
    nel, b = (2, 2, 2), (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    m = _box(4, nel, b)
    pts = np.array([
        [0.5, 0.25, 0.25],   # on the x-face between elements 0 and 1
        [0.5, 0.5, 0.5],     # the centre vertex shared by all eight
        [-0.3, 0.25, 0.75],  # outside in -x: clamped into element (0, 0, 1)
        [2.0, 2.0, 2.0],     # outside the far corner
        [0.25, 0.25, 0.25],  # element 0 centre
    ])
    elem, rst = m.locateProbes(pts)
    assert list(elem) == [0, 0, _elem_index(0, 0, 1, nel), 7, 0]
    np.testing.assert_allclose(rst[0], [1.0, 0.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(rst[1], [1.0, 1.0, 1.0], atol=1e-15)
    np.testing.assert_allclose(rst[2], [-1.0, 0.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(rst[3], [1.0, 1.0, 1.0], atol=1e-15)
    np.testing.assert_allclose(rst[4], [0.0, 0.0, 0.0], atol=1e-15)


def test_locate_empty_list():
    m = _box(3, (1, 1, 1), (0, 1, 0, 1, 0, 1))
    elem, rst = m.locateProbes(np.zeros((0, 3)))
    assert elem.shape == (0,) and rst.shape == (0, 3)


def test_config_probe_keys():
    """config.hpp:291-322 and 473-490: probes as [x, y, z] lists or strings;
    relative probes are centre + r * L with L = max - min per axis."""
    c = Config.from_json({"mesh": {"nelx": 2, "nely": 2, "nelz": 2, "xmin": 0.0, "xmax": 2.0,
                                   "ymin": -1.0, "ymax": 1.0, "zmin": 0.0, "zmax": 4.0},
                          "probes": [[0.1, 0.2, 0.3], "0.4, 0.5, 0.6"],
                          "relative_probes": [[0.25, -0.5, 0.125]],
                          "output": {"output_dir": "none"}})
    assert c.probes == [(0.1, 0.2, 0.3), (0.4, 0.5, 0.6)]
    assert c.output_dir == "none"
    case = Case(c)
    case.preprocess()
    case.addProbe(1.0, 0.0, 1.0)
    case.addRelativeProbe(0.0, 0.0, 0.0)
    p = case._probe_list()
    np.testing.assert_allclose(p, [[0.1, 0.2, 0.3], [0.4, 0.5, 0.6], [1.0, 0.0, 1.0],
                                   [1.0 + 0.25 * 2.0, 0.0 - 0.5 * 2.0, 2.0 + 0.125 * 4.0], [1.0, 0.0, 2.0]])


@pytest.mark.parametrize("nel", [(1, 1, 1), (3, 1, 1)])
def test_config_num_probes(nel):
    """case.cpp:397-422: with no explicit probes, num_probes places them at the
    GLL nodes along x (single-element box) or evenly on the centre line."""
    c = Config(nelx=nel[0], nely=nel[1], nelz=nel[2], order=5, num_probes=7,
               xmin=0.0, xmax=2.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0)
    case = Case(c)
    case.preprocess()
    p = case._probe_list()
    if nel == (1, 1, 1):
        z = case.mesh.getGllZ()
        np.testing.assert_allclose(p[:, 0], 1.0 + z, atol=1e-15)  # 5 of 7: min(num_probes, N)
    else:
        np.testing.assert_allclose(p[:, 0], np.linspace(0.1, 1.9, 7), atol=1e-15)
    np.testing.assert_allclose(p[:, 1:], 0.5, atol=0)
