"""On-device probes (ProbeManager, probe.cpp:33-216) through the C ABI.

Two checks pin the device records:
  * exactness: on an affine box mesh a field polynomial of degree <= N-1 per
    coordinate is reproduced exactly by GLL interpolation, so a record of such a
    state equals the polynomial at the (clamped) probe point;
  * parity with ProbeManager::record: every per-step record equals the oracle's
    restatement (w = Lr[i] Ls[j] Lt[k], acc += w * u in node order, FP64, no
    contraction) of the state downloaded after the same step -- bit-exact.
"""
import os

import numpy as np
import pytest

import nekwave_amd as nw
from nekwave_amd.case import Case, Config
import pyoracle as po
from gpu_util import pair, random_state

pytestmark = pytest.mark.gpu


def _host_record(o, pts, state):
    """The oracle's ProbeManager::record (probe.cpp:192-216) of a downloaded state."""
    return o.probe_record(pts, state)


PTS = np.array([[0.13, 0.52, 0.77], [0.5, 0.5, 0.5], [0.999, 0.01, 0.3], [-0.2, 0.4, 1.3], [0.25, 0.25, 0.25]])


@pytest.mark.parametrize("pipeline,precision", [("fused", "fp64"), ("fused", "fp32"), ("plain", "fp64")])
def test_probe_polynomial_exact(pipeline, precision):
    n1 = 5
    m, o = pair(n1, (2, 3, 2))
    x, y, zc = o.x, o.y, o.zc
    deg = n1 - 1
    f = [x ** deg * y + zc, y ** deg - x * zc, zc ** deg + x * y, x * y * zc, (x - y) ** 2, 1.0 + 0 * x]
    st = np.concatenate(f)
    s = nw.DgSolver(precision=precision, pipeline=pipeline)
    s.initialize(m, 1.0)
    s.uploadState(st)
    s.setProbes(m, PTS, capacity=4)
    s.sampleProbes()
    h = s.probeHistory()
    s.finalize()
    assert h.shape == (1, len(PTS), 6)
    cp = np.clip(PTS, 0.0, 1.0)
    X, Y, Z = cp[:, 0], cp[:, 1], cp[:, 2]
    want = np.stack([X ** deg * Y + Z, Y ** deg - X * Z, Z ** deg + X * Y, X * Y * Z, (X - Y) ** 2, 1.0 + 0 * X], 1)
    tol = 1e-12 if precision == "fp64" else 2e-6
    err = np.max(np.abs(h[0] - want))
    print("%s %s max |probe - polynomial| = %.3e" % (pipeline, precision, err))
    assert err <= tol, err


@pytest.mark.parametrize("pipeline,precision,use_graph", [("fused", "fp64", True), ("fused", "fp32", True),
                                                          ("fused", "fp64", False), ("plain", "fp64", True)])
def test_probe_records_match_host_record_each_step(pipeline, precision, use_graph):
    m, o = pair(6, (3, 2, 2), (0, 1.5, 0, 1, 0, 1), jitter=0.08, seed=4)
    st0 = random_state(o, 11)
    dt = 0.5 * o.auto_dt()
    pts = np.array([[0.2, 0.3, 0.4], [1.4, 0.9, 0.1], [0.75, 0.5, 0.5]])
    s = nw.DgSolver(precision=precision, pipeline=pipeline, use_graph=use_graph)
    s.initialize(m, 1.0)
    s.uploadState(st0)
    s.setProbes(m, pts, capacity=8)
    states = [s.downloadState()]
    s.sampleProbes()
    for _ in range(3):
        s.step(dt)
        states.append(s.downloadState())
    s.run(dt, 2)  # several steps in one call record one per step too
    states.append(None)
    states.append(s.downloadState())
    h = s.probeHistory()
    assert h.shape == (6, len(pts), 6)
    for r, stt in enumerate(states):
        if stt is None:
            continue
        want = _host_record(o, pts, stt)
        assert np.array_equal(h[r], want), (r, np.max(np.abs(h[r] - want)))
    # the history starts again after a fetch
    s.step(dt)
    h2 = s.probeHistory()
    assert h2.shape == (1, len(pts), 6)
    np.testing.assert_array_equal(h2[0], _host_record(o, pts, s.downloadState()))
    s.finalize()


def test_probe_history_overflow_is_an_error():
    m, o = pair(4, (2, 2, 2))
    s = nw.DgSolver(precision="fp64")
    s.initialize(m, 1.0)
    s.uploadState(random_state(o, 2))
    s.setProbes(m, PTS[:2], capacity=2)
    s.run(0.5 * o.auto_dt(), 3)
    with pytest.raises(nw.NekWaveError):
        s.probeHistory()
    s.setProbes(m, PTS[:2], capacity=2)  # placing them again starts a new history
    s.step(0.5 * o.auto_dt())
    assert s.probeHistory().shape == (1, 2, 6)
    s.setProbes(m, np.zeros((0, 3)), capacity=1)  # removes the probes
    s.step(0.5 * o.auto_dt())
    assert s.probeHistory().shape == (0, 0, 6)
    s.finalize()


def test_case_writes_probe_history_csv(tmp_path):
    """Case records t = 0 and every step (case.cpp:473-474, 594-596) and writes
    probe_history.csv in ProbeManager's format (probe.cpp:117-133, 192-216)."""
    out_dir = str(tmp_path / "out")
    cfg = Config(order=5, nelx=2, nely=2, nelz=2, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0,
                 num_steps=7, output_frequency=3, output_dir=out_dir, probes=[[0.3, 0.6, 0.2]],
                 relative_probes=[[0.0, 0.0, 0.0]])
    case = Case(cfg)
    case.setInitialConditionHook(lambda x, y, z: (0 * x, 0 * x, np.sin(np.pi * x) * np.sin(np.pi * y), 0 * x,
                                                  0 * x, 0 * x))
    case.preprocess()
    case.simulate()
    assert case.probe_steps == list(range(8))
    assert case.probe_values.shape == (8, 2, 6)
    o = po.Mesh.box(5, (2, 2, 2), (0.0, 1.0, 0.0, 1.0, 0.0, 1.0))
    np.testing.assert_array_equal(case.probe_values[-1], _host_record(o, case.probe_points, case.state))
    lines = open(os.path.join(out_dir, "probe_history.csv")).read().splitlines()
    assert lines[0] == "step,time," + ",".join("%s_%d" % (c, q) for q in (1, 2)
                                                for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    assert len(lines) == 9
    row = lines[-1].split(",")
    assert row[0] == "7" and row[1] == "%.8e" % case.probe_times[-1]
    np.testing.assert_allclose([float(v) for v in row[2:]], case.probe_values[-1].reshape(-1), rtol=1e-8, atol=0)
    import json
    meta = json.load(open(os.path.join(out_dir, "probes.json")))
    assert meta["order"] == 5 and meta["num_probes"] == 2 and meta["elements_x"] == 2
    np.testing.assert_allclose([p["target"] for p in meta["probes"]], case.probe_points, rtol=1e-8)
    case.postprocess()


def test_case_launches_output_intervals_as_one_run():
    """Case runs the steps between outputs as one multi-step launch: same step
    schedule (case.cpp:562-583, including the clamped last dt), same times and
    the same state bit for bit as one synchronised step() per step."""
    cfg = Config(order=4, nelx=2, nely=2, nelz=2, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0,
                 dt=0.01, num_steps=0, final_time=0.075, output_frequency=3, output_dir="none",
                 probes=[[0.4, 0.4, 0.6]])
    case = Case(cfg)
    case.setInitialConditionHook(lambda x, y, z: (0 * x, 0 * x, np.sin(np.pi * x) * np.sin(np.pi * y), 0 * x,
                                                  0 * x, 0 * x))
    case.preprocess()
    st0 = case.state.copy()
    case.simulate()
    s = nw.DgSolver(precision="fp64")
    s.initialize(case.mesh, cfg.c0)
    s.uploadState(st0)
    t, times = 0.0, [0.0]
    while t < 0.075 - 1e-13:
        dt = 0.01 if t + 0.01 <= 0.075 else 0.075 - t
        s.step(dt, t)
        t += dt
        times.append(t)
    assert case.current_step == len(times) - 1 == 8
    assert case.probe_times == times
    assert len(case.history) == 3  # steps 3, 6 and the last
    np.testing.assert_array_equal(case.state, s.downloadState())
    s.finalize()
    case.postprocess()


def test_case_fetches_early_when_the_history_fills(monkeypatch):
    """With a device history shorter than the output interval, Case fetches
    whenever it fills; the records equal those of a run with room to spare."""
    def run(cap):
        monkeypatch.setattr(Case, "PROBE_CAPACITY", cap)
        cfg = Config(order=4, nelx=2, nely=2, nelz=2, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0,
                     num_steps=25, output_frequency=10, output_dir="none", probes=[[0.4, 0.4, 0.6], [0.9, 0.1, 0.5]])
        case = Case(cfg)
        case.setInitialConditionHook(lambda x, y, z: (0 * x, 0 * x, np.sin(np.pi * x) * np.sin(np.pi * y), 0 * x,
                                                      0 * x, 0 * x))
        case.preprocess()
        case.simulate()
        case.postprocess()
        return case
    small, big = run(4), run(4096)
    assert small.probe_steps == big.probe_steps == list(range(26))
    np.testing.assert_array_equal(small.probe_values, big.probe_values)
    np.testing.assert_array_equal(small.state, big.state)


def test_probe_entry_points_reject_bad_calls():
    """Error paths of nw_dg_set_probes / nw_dg_probe_sample: a batch of several
    meshes, a mesh that is not the solver's, a zero capacity and sampling before
    any probe is placed come back as status codes."""
    import ctypes as C

    from nekwave_amd import _lib

    m, o = pair(4, (2, 2, 2))
    m2, _ = pair(4, (3, 2, 2))
    xyz = np.ascontiguousarray(PTS[:1])
    ptr = xyz.ctypes.data_as(C.POINTER(C.c_double))
    b = nw.BatchSolver(precision="fp64")
    b.initialize([m, m2], 1.0)
    assert nw.lib().nw_dg_set_probes(b._h, m.handle, 1, ptr, 4) == _lib.NW_ERR_ARG
    b.finalize()
    s = nw.DgSolver(precision="fp64")
    s.initialize(m, 1.0)
    s.uploadState(random_state(o, 1))
    assert nw.lib().nw_dg_probe_sample(s._handle()) == _lib.NW_ERR_STATE
    assert nw.lib().nw_dg_set_probes(s._handle(), m2.handle, 1, ptr, 4) == _lib.NW_ERR_ARG
    # same n1 and element count, other bounds: not the solver's mesh either
    m3 = nw.Mesh(4)
    m3.createBoxMesh(2, 2, 2, 0.0, 2.0, 0.0, 1.0, 0.0, 1.0)
    assert nw.lib().nw_dg_set_probes(s._handle(), m3.handle, 1, ptr, 4) == _lib.NW_ERR_ARG
    assert "not the one the solver was created from" in nw.lib().nw_last_error().decode()
    assert nw.lib().nw_dg_set_probes(s._handle(), m.handle, 1, ptr, 0) == _lib.NW_ERR_ARG
    s.setProbes(m, xyz, capacity=2)  # and the solver still works afterwards
    s.sampleProbes()
    assert s.probeHistory().shape == (1, 1, 6)
    s.finalize()
