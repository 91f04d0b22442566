"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces every fixture (a regression pin on the oracle
build), and the product's seeded draws (nw_rand_*) equal the oracle's
mt19937_64 stream bit for bit.
GPU: the product's FP64 step, both pipelines, lands on the fixture's
after-step state to 1e-12 relative L2 (dV = J w3).
"""
import glob
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as mg  # noqa: E402
import pyoracle as po  # noqa: E402

FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def test_fixture_set_complete():
    assert FIXTURES == sorted(mg.CASES)


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_fixture(name):
    g = load(name)
    o, eps, mu, c0 = mg.build(name)
    st = mg.seeded_state(o)
    np.testing.assert_array_equal(st, g["state"])
    assert o.auto_dt() * 0.5 == pytest.approx(float(g["dt"]), rel=1e-15)
    n = o.npts
    vol = np.concatenate(po.volume_curl(o, st[3 * n:4 * n], st[4 * n:5 * n], st[5 * n:6 * n], 1.0))
    np.testing.assert_allclose(vol, g["volume_curl_H"], rtol=0, atol=1e-12 * np.abs(g["volume_curl_H"]).max())
    fEN, fHN = po.restrict(o, st)
    fl = po.flux(o, st, fEN, fHN, c0)
    np.testing.assert_allclose(fl, g["flux"], rtol=0, atol=1e-12 * np.abs(g["flux"]).max())
    s = po.Solver(o, c0)
    s.upload(st)
    s.step(float(g["dt"]))
    out = s.download()
    np.testing.assert_allclose(out, g["after_step"], rtol=0, atol=1e-12 * np.abs(g["after_step"]).max())
    assert o.energy(st) == pytest.approx(float(g["energy"]), rel=1e-12)


def test_fixture_step_is_a_small_change():
    """The fixture step is not trivial: the (random, rough) state moved by a bounded amount."""
    for name in FIXTURES:
        g = load(name)
        d = np.abs(g["after_step"] - g["state"]).max()
        assert 1e-6 < d < 10.0, (name, d)


def test_rand_matches_oracle_stream():
    from nekwave_amd import workloads as W

    for seed in (0, 2, 3, 4, 5, 12345678901):
        u = W.uniforms(seed, 257)
        r = po.Rng(seed)
        assert np.array_equal(u, np.array([r.uniform() for _ in range(257)]))
        nn = W.normals(seed, 65)
        r = po.Rng(seed)
        assert np.array_equal(nn, np.array([r.normal() for _ in range(65)]))


def test_rand_closed_forms():
    """mt19937_64 seeded 5489 (the standard default): its 10000th output is
    9981545732273789042 (C++11 [rand.predef]); uniforms lie in [0, 1)."""
    r = po.Rng(5489)
    for _ in range(9999):
        r.next()
    assert r.next() == 9981545732273789042
    from nekwave_amd import workloads as W

    u = W.uniforms(7, 100000)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("pipeline", ["fused", "plain"])
@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_step_matches_fixture(name, pipeline):
    import nekwave_amd as nw

    g = load(name)
    nel = [int(v) for v in g["nel"]]
    m = nw.Mesh(int(g["n1"]))
    m.createBoxMesh(*nel, *[float(v) for v in g["bounds"]], *[bool(v) for v in g["periodic"]],
                    jitter=float(g["jitter"]), seed=int(g["seed"]))
    if g["eps"].size:
        m.setMaterials(g["eps"], g["mu"])
    s = nw.DgSolver(precision="fp64", pipeline=pipeline, use_graph=True)
    s.initialize(m, float(g["c0"]))
    s.uploadState(g["state"])
    s.step(float(g["dt"]))
    out = s.downloadState()
    s.finalize()
    n = m.getTotalPoints()
    dV = m.getJac() * np.tile(m.getW3(), m.getNumElements())
    num = sum(np.sum((out[c * n:(c + 1) * n] - g["after_step"][c * n:(c + 1) * n]) ** 2 * dV) for c in range(6))
    den = sum(np.sum(g["after_step"][c * n:(c + 1) * n] ** 2 * dV) for c in range(6))
    assert np.sqrt(num / den) < 1e-12
