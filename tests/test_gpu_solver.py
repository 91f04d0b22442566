"""Solver parity on the GPU through the C ABI (DgSolver replacement) against the
FP64 oracle on the same seeded inputs.

Tolerances (relative L2 with the J*w3 weight of 3dboxpec_test.cpp:76-119):
  * FP64, plain and fused pipelines: <= 1e-12;
  * FP32 fields (k FP32 or FP64):    <= 2e-5 after the stated steps;
and the reference's KATs on the device: energy non-increasing with upwind
(nekwave_test.cpp:212-293), eigenmode errors (3dboxpec_test.cpp:136).
"""
import math
import os

import numpy as np
import pytest

import analytic as an
import nekwave_amd as nw
import pyoracle as po
from gpu_util import gaussian_state, pair, random_state, rel_l2, run_oracle, run_product

pytestmark = pytest.mark.gpu

CASES = {
    # name: (n1, nel, bounds, periodic, jitter, seed, media, halo, c0)
    "pec_tm110_p4": (5, (4, 4, 4), (0, 1, 0, 1, 0, 1), (0, 0, 0), 0.0, 0, False, False, 1.0),
    "periodic_jitter_p3": (4, (3, 3, 3), (0, 1, 0, 1, 0, 1), (1, 1, 1), 0.1, 2, False, False, 1.0),
    "periodic_jitter_p7": (8, (2, 2, 2), (0, 1, 0, 1, 0, 1), (1, 1, 1), 0.1, 2, False, False, 1.0),
    "media_pec_p5": (6, (3, 2, 2), (0, 1.5, 0, 1, 0, 1), (0, 0, 0), 0.0, 0, True, False, 1.0),
    "media_pec_p7": (8, (3, 2, 2), (0, 1.5, 0, 1, 0, 1), (0, 0, 0), 0.05, 4, True, False, 1.0),
    "halo_media_p4": (5, (4, 2, 2), (0, 1, 0, 1, 0, 1), (1, 0, 0), 0.0, 0, True, True, 1.0),
    "central_jitter_p2": (3, (3, 2, 2), (0, 1, 0, 1, 0, 1), (0, 1, 0), 0.05, 8, False, False, 0.0),
}


def _build(name):
    n1, nel, b, per, jit, seed, media, halo, c0 = CASES[name]
    nelt = nel[0] * nel[1] * nel[2]
    eps = mu = None
    if media:
        rng = np.random.default_rng(5)
        eps = np.exp(0.3 * rng.standard_normal(nelt))
        mu = np.exp(0.3 * rng.standard_normal(nelt))
    m, o = pair(n1, nel, b, per, jit, seed, eps, mu, halo)
    return m, o, c0


@pytest.mark.parametrize("pipeline", ["plain", "fused"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_fp64_matches_oracle(name, pipeline):
    m, o, c0 = _build(name)
    st0 = random_state(o, 3)
    dt = 0.5 * o.auto_dt()
    for nsteps in (1, 7):
        ref = run_oracle(o, st0, dt, nsteps, c0)
        got = run_product(m, st0, dt, nsteps, c0, "fp64", pipeline)
        err = rel_l2(o, got, ref)
        print("%s %s steps=%d relL2=%.3e" % (name, pipeline, nsteps, err))
        assert err <= 1e-12, err


@pytest.mark.parametrize("precision", ["fp32", "fp32_k64"])
@pytest.mark.parametrize("name", ["pec_tm110_p4", "periodic_jitter_p7", "halo_media_p4"])
def test_fp32_matches_oracle(name, precision):
    m, o, c0 = _build(name)
    st0 = gaussian_state(o, width=0.2)
    dt = 0.5 * o.auto_dt()
    ref = run_oracle(o, st0, dt, 20, c0)
    got = run_product(m, st0, dt, 20, c0, precision, "fused")
    err = rel_l2(o, got, ref)
    print("%s %s relL2=%.3e" % (name, precision, err))
    assert err <= 2e-5, err


@pytest.mark.parametrize("name", ["periodic_jitter_p7", "media_pec_p7"])
def test_bf16x3_matches_oracle_within_its_bound(name):
    """Split-BF16 matrix-core contractions (NW_PREC_BF16X3): each operand as
    hi + lo BF16 halves, hi*hi + hi*lo + lo*hi in FP32 -- a product error of
    about 2^-16 relative instead of FP32's 2^-24.  After 20 steps within 2e-4 of
    the FP64 oracle (FP32: 2e-5), and measurably different from the FP32 path
    (the split products are really used); deterministic and graph = stream."""
    m, o, c0 = _build(name)
    st0 = gaussian_state(o, width=0.2)
    dt = 0.5 * o.auto_dt()
    ref = run_oracle(o, st0, dt, 20, c0)
    got = run_product(m, st0, dt, 20, c0, "bf16x3", "fused")
    f32 = run_product(m, st0, dt, 20, c0, "fp32", "fused")
    err, err32 = rel_l2(o, got, ref), rel_l2(o, f32, ref)
    print("%s bf16x3 relL2=%.3e (fp32 %.3e)" % (name, err, err32))
    assert err <= 2e-4, err
    assert not np.array_equal(got, f32)
    again = run_product(m, st0, dt, 20, c0, "bf16x3", "fused", use_graph=False)
    assert np.array_equal(got, again)


def test_bf16x3_needs_n1_8():
    m, o, c0 = _build("pec_tm110_p4")
    s = nw.DgSolver(precision="bf16x3", pipeline="fused")
    with pytest.raises(nw.NekWaveError) as e:
        s.initialize(m, c0)
    assert e.value.status == nw._lib.NW_ERR_UNSUPPORTED


@pytest.mark.parametrize("name", ["pec_tm110_p4", "media_pec_p5", "halo_media_p4", "sheared_n8"])
def test_affine_records_equal_per_node_geometry(name, monkeypatch):
    """On meshes of parallelepipeds the fused kernel reads each element's
    geometry from its record (computed once per element) instead of per node;
    the bits must be those of the per-node computation (NW_GENERAL_GEOMETRY=1)."""
    if name == "sheared_n8":  # n1 = 8 (the matrix-core shapes), parallelepipeds that are not boxes
        corners = []
        for e in range(4):
            for a in range(8):
                i, j, k = a in (1, 2, 5, 6), a in (2, 3, 6, 7), a >= 4
                x, y, z = e + i, 0.5 * j, 0.75 * k
                corners.append((x + 0.3 * y + 0.1 * z, y + 0.2 * z, z))
        corners = np.array(corners).reshape(4, 8, 3)
        kind = np.full((4, 6), nw.FACE_PEC, np.int32)
        ne = np.full((4, 6), -1, np.int32)
        nf = np.full((4, 6), -1, np.int32)
        for e in range(3):
            kind[e, 1] = kind[e + 1, 3] = nw.FACE_E
            ne[e, 1], nf[e, 1], ne[e + 1, 3], nf[e + 1, 3] = e + 1, 3, e, 1
        m = nw.Mesh.fromCorners(8, corners, kind, ne, nf)
        o = po.Mesh.from_corners(8, corners, kind, ne, nf)
        c0 = 1.0
    else:
        m, o, c0 = _build(name)
    assert m.info().all_affine
    st0 = random_state(o, 9)
    dt = 0.5 * o.auto_dt()
    for prec in ("fp64", "fp32", "fp32_k64") + (("bf16x3",) if m.getN() == 8 else ()):
        monkeypatch.delenv("NW_GENERAL_GEOMETRY", raising=False)
        a = run_product(m, st0, dt, 3, c0, prec, "fused")
        monkeypatch.setenv("NW_GENERAL_GEOMETRY", "1")
        b = run_product(m, st0, dt, 3, c0, prec, "fused")
        assert np.array_equal(a, b), prec


def test_graph_and_stream_paths_are_bitwise_equal_and_deterministic():
    m, o, c0 = _build("periodic_jitter_p3")
    st0 = random_state(o, 1)
    dt = o.auto_dt()
    for prec in ("fp64", "fp32"):
        a = run_product(m, st0, dt, 5, c0, prec, "fused", use_graph=True)
        b = run_product(m, st0, dt, 5, c0, prec, "fused", use_graph=False)
        c = run_product(m, st0, dt, 5, c0, prec, "fused", use_graph=True)
        assert np.array_equal(a, b) and np.array_equal(a, c), prec
    a = run_product(m, st0, dt, 3, c0, "fp64", "plain", use_graph=True)
    b = run_product(m, st0, dt, 3, c0, "fp64", "plain", use_graph=False)
    assert np.array_equal(a, b)


def test_graph_replays_with_host_work_between():
    """Replays of the captured step, interleaved with eager work and syncs, each
    equal to the eager result (graph replay hazard check)."""
    m, o, c0 = _build("halo_media_p4")
    st0 = random_state(o, 4)
    dt = o.auto_dt()
    ref = run_product(m, st0, dt, 4, c0, "fp64", "fused", use_graph=False)
    s = nw.DgSolver(precision="fp64", pipeline="fused", use_graph=True)
    s.initialize(m, c0)
    for rep in range(3):
        s.uploadState(st0)
        for _ in range(4):
            s.step(dt)
            s.diagnostics()
        assert np.array_equal(s.downloadState(), ref), rep
    s.finalize()


@pytest.mark.parametrize("name", ["media_pec_p5", "periodic_jitter_p7", "media_pec_p7", "halo_media_p4",
                                  "periodic_jitter_p3"])
def test_tile_options_agree(name):
    """Both tile options give the same bits: other elements per workgroup, and
    at n1 = 8 in FP32 the matrix-core volume (one fused multiply-add per step,
    in the vector path's order)."""
    m, o, c0 = _build(name)
    st0 = random_state(o, 2)
    dt = o.auto_dt()
    for prec in ("fp64", "fp32", "fp32_k64"):
        a = run_product(m, st0, dt, 3, c0, prec, "fused", tile_option=0)
        for opt in (1, 2):
            b = run_product(m, st0, dt, 3, c0, prec, "fused", tile_option=opt)
            assert np.array_equal(a, b), (prec, opt)


@pytest.mark.parametrize("name", ["media_pec_p5", "periodic_jitter_p7", "halo_media_p4"])
def test_segment_choices_give_the_same_bits(name):
    """Every tile option with the LSRK register accesses non-temporal or not
    (nw_dg_set_segment_option) gives the bits of tile option 0 without the hint,
    also when the choice changes between runs of one solver (graphs re-captured),
    and the solver reports the choice it runs."""
    m, o, c0 = _build(name)
    st0 = random_state(o, 3)
    dt = o.auto_dt()
    for prec in ("fp32", "fp64"):
        ref = run_product(m, st0, dt, 4, c0, prec, "fused", tile_option=0)
        for opt in (0, 1, 2):
            for sk in (0, 1):
                s = nw.DgSolver(precision=prec, pipeline="fused", use_graph=True)
                s.initialize(m, c0)
                s.uploadState(st0)
                s.setSegmentOption(0, (opt + 1) % 3, 1 - sk)  # a first choice, replaced after 2 steps
                s.run(dt, 2)
                s.setSegmentOption(0, opt, sk)
                seg = s.segments()[0]
                assert (seg["option"], seg["stream_k"]) == (opt, bool(sk)), seg
                s.run(dt, 2)
                got = s.downloadState()
                s.finalize()
                assert np.array_equal(got, ref), (prec, opt, sk)


def test_tune_between_runs_keeps_the_state():
    """tune() after graph-captured steps: its timing launches write only the
    stage's output buffer and k (the state stays in the input buffer), and the
    graphs are re-captured with the new choices -- the run continues with the
    bits of an untuned run."""
    m, o, c0 = _build("periodic_jitter_p7")
    st0 = random_state(o, 4)
    dt = o.auto_dt()
    ref = run_product(m, st0, dt, 6, c0, "fp32", "fused", tile_option=0)
    s = nw.DgSolver(precision="fp32", pipeline="fused", use_graph=True, tile_option=0)
    s.initialize(m, c0)
    s.uploadState(st0)
    s.run(dt, 3)
    s.tune()
    s.run(dt, 3)
    got = s.downloadState()
    s.finalize()
    assert np.array_equal(got, ref)


def test_set_segment_option_rejects_bad_arguments():
    m, o, c0 = _build("media_pec_p5")
    s = nw.DgSolver(precision="fp32", pipeline="fused")
    s.initialize(m, c0)
    for args in [(1, 0, 0), (-1, 0, 0), (0, 3, 0), (0, -1, 0), (0, 0, 2), (0, 0, -2)]:
        with pytest.raises(nw.NekWaveError):
            s.setSegmentOption(*args)
    s.finalize()
    p = nw.DgSolver(precision="fp64", pipeline="plain")
    p.initialize(m, c0)
    with pytest.raises(nw.NekWaveError):
        p.setSegmentOption(0, 0, 0)
    p.finalize()
    q = nw.DgSolver(precision="fp32", pipeline="fused", groups_per_cu=-1)
    with pytest.raises(nw.NekWaveError):
        q.initialize(m, c0)


def test_segment_option_reports_matrix_volume():
    m, o, c0 = _build("periodic_jitter_p7")
    for opt, prec, want in [(0, "fp32", False), (1, "fp32", True), (1, "fp32_k64", True), (1, "fp64", False)]:
        s = nw.DgSolver(precision=prec, pipeline="fused", tile_option=opt)
        s.initialize(m, c0)
        seg = s.segments()[0]
        s.finalize()
        # a working set far below the 256 MiB Infinity Cache: k stays temporal
        assert seg["stream_k"] is False, seg
        assert (seg["n1"], seg["option"], seg["matrix_volume"]) == (8, opt, want), (prec, seg)


@pytest.mark.parametrize("name", ["media_pec_p5", "periodic_jitter_p7", "halo_media_p4"])
@pytest.mark.parametrize("pipeline,precision", [("plain", "fp64"), ("fused", "fp64"), ("fused", "fp32")])
def test_device_rel_l2_matches_host(name, pipeline, precision):
    """nw_dg_rel_l2 (3dboxpec_test.cpp:76-119 norm, reduced on the device) equals
    the host formula on the downloaded state, and is 0 against the state itself."""
    m, o, c0 = _build(name)
    st0 = random_state(o, 7)
    dt = 0.5 * o.auto_dt()
    s = nw.DgSolver(precision=precision, pipeline=pipeline)
    s.initialize(m, c0)
    s.uploadState(st0)
    s.run(dt, 3)
    got = s.downloadState()
    ref = run_oracle(o, st0, dt, 4, c0)
    host = rel_l2(o, got, ref)
    dev = s.relL2(ref)
    assert s.relL2(got) == 0.0
    s.finalize()
    print("%s %s %s device %.15e host %.15e" % (name, pipeline, precision, dev, host))
    assert host > 0 and abs(dev - host) <= 1e-12 * host, (dev, host)


@pytest.mark.parametrize("pipeline,precision", [("plain", "fp64"), ("fused", "fp64"), ("fused", "fp32")])
def test_device_diagnostics_match_oracle(pipeline, precision):
    m, o, c0 = _build("media_pec_p5")
    st0 = random_state(o, 6)
    s = nw.DgSolver(precision=precision, pipeline=pipeline)
    s.initialize(m, c0)
    s.uploadState(st0)
    d = s.diagnostics()
    tol = 1e-13 if precision == "fp64" else 1e-6
    assert abs(d["energy"] - o.energy(st0)) <= tol * o.energy(st0)
    assert abs(d["max_e"] - o.max_e(st0)) <= tol * o.max_e(st0)
    assert abs(d["max_h"] - o.max_h(st0)) <= tol * o.max_h(st0)
    s.finalize()


@pytest.mark.parametrize("pipeline,precision", [("plain", "fp64"), ("fused", "fp64"), ("fused", "fp32"),
                                                ("fused", "fp32_k64")])
def test_maxwell_stability_kat_on_device(pipeline, precision):
    """nekwave_test.cpp:212-293: n1=4, 2x2x2 on [-0.5,0.5]^3, upwind, Ez=exp(-50 r^2),
    dt=0.005, 5 steps, energy never above E0 + 1e-12 and finite."""
    m = nw.Mesh(4)
    m.createBoxMesh(2, 2, 2, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5)
    x, y, z = m.coords()
    st = np.zeros(6 * x.size)
    st[2 * x.size:3 * x.size] = np.exp(-50.0 * (x * x + y * y + z * z))
    s = nw.DgSolver(precision=precision, pipeline=pipeline)
    s.initialize(m, 1.0)
    s.uploadState(st)
    e0 = s.diagnostics()["energy"]
    assert e0 > 0
    t = 0.0
    for _ in range(5):
        s.step(0.005, t)
        t += 0.005
        e = s.diagnostics()["energy"]
        slack = 1e-12 if precision == "fp64" else 1e-6 * e0
        assert e <= e0 + slack and math.isfinite(e)
    s.finalize()


@pytest.mark.parametrize("pipeline,precision", [("plain", "fp64"), ("fused", "fp64"), ("fused", "fp32")])
def test_tm110_c1_on_device(pipeline, precision):
    """PEC unit cube 4^3, p = 4, CFL 0.5, upwind, TM110, 200 steps."""
    m, o = pair(5, (4, 4, 4))
    dt = 0.5 * o.dxmin()
    st0 = an.pack(an.tm_mode(o.x, o.y, 0.0))
    got = run_product(m, st0, dt, 200, 1.0, precision, pipeline)
    err = an.rel_l2(an.dV(o), got, an.pack(an.tm_mode(o.x, o.y, 200 * dt)))
    assert err < 2e-3, err


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_boxper_mode_on_device(precision):
    """3dboxper_test restated on the device: [0, 2 pi]^3 periodic 3x3x3, order 4,
    central flux, auto dt, 50 steps, relative L2 to the analytic mode <= 0.05;
    and within 2e-5 of the oracle's own error."""
    L = 2 * math.pi
    m, o = pair(5, (3, 3, 3), (0, L, 0, L, 0, L), (1, 1, 1))
    dt = o.auto_dt(1.0, 0.4)
    st0 = an.pack(an.boxper_mode(o.x, o.y, o.zc, 0.0))
    ex = an.pack(an.boxper_mode(o.x, o.y, o.zc, 50 * dt))
    got = run_product(m, st0, dt, 50, 0.0, precision, "fused")
    err = an.rel_l2(an.dV(o), got, ex)
    err_o = an.rel_l2(an.dV(o), run_oracle(o, st0, dt, 50, 0.0), ex)
    print("boxper %s relL2=%.3e (oracle %.3e)" % (precision, err, err_o))
    assert err <= 0.05 and abs(err - err_o) <= 2e-5, (err, err_o)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_p_convergence_on_device(precision):
    """TM110 on a 2^3 PEC cube to T = 0.3: each order cuts the error by more
    than 3 (spectral convergence; oracle errors 8.7e-2, 8.6e-3, 6.2e-4, 3.5e-5,
    1.7e-6 for n1 = 3..7), and each error equals the oracle's to 1e-10 (FP64)
    or 1e-6 (FP32, which stops at n1 = 6, above its rounding floor)."""
    T = 0.3
    errs = []
    tol = 1e-10 if precision == "fp64" else 1e-6
    for n1 in ((3, 4, 5, 6, 7) if precision == "fp64" else (3, 4, 5, 6)):
        m, o = pair(n1, (2, 2, 2))
        nst = int(math.ceil(T / (0.2 * o.dxmin())))
        dt = T / nst
        st0 = an.pack(an.tm_mode(o.x, o.y, 0.0))
        ex = an.pack(an.tm_mode(o.x, o.y, T))
        got = run_product(m, st0, dt, nst, 1.0, precision, "fused")
        e = an.rel_l2(an.dV(o), got, ex)
        e_o = an.rel_l2(an.dV(o), run_oracle(o, st0, dt, nst, 1.0), ex)
        assert abs(e - e_o) <= tol, (n1, e, e_o)
        errs.append(e)
    print("p-convergence %s" % precision, errs)
    for e0, e1 in zip(errs, errs[1:]):
        assert e1 < e0 / 3, errs


def test_case_lifecycle_boxpec_mode():
    """3dboxpec_test.cpp restated: Case with a 3x3x3 [-1,1]^3 box, order 4,
    central flux, auto dt, 50 steps, relative L2 <= 0.05."""
    cfg = nw.Config(order=4, nelx=3, nely=3, nelz=3, c0=0.0, num_steps=50, output_frequency=10)
    c = nw.Case(cfg)
    c.setInitialConditionHook(lambda x, y, z: an.boxpec_mode(x, y, z, 0.0))
    res = {}

    def post(case):
        x, y, z = case.mesh.coords()
        ex = an.pack(an.boxpec_mode(x, y, z, case.current_time))
        jw = case.mesh.getW3()
        o = __import__("pyoracle").Mesh.box(4, (3, 3, 3), (-1, 1, -1, 1, -1, 1))
        res["err"] = an.rel_l2(an.dV(o), case.state, ex)
        assert jw.size == 64

    c.setPostprocessingHook(post)
    c.run()
    assert c.current_step == 50
    assert res["err"] <= 0.05, res
    assert len(c.history) == 5
    energies = [h[4] for h in c.history]
    assert max(energies) - min(energies) < 1e-3 * energies[0]  # central flux: conserved to LSRK error


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cpp_reference_kats_on_device(prec):
    """tests/cpp/reference_kats.cpp through the C++ host header: the reference's
    MaxwellStabilityTest and 3dboxpec Case lifecycle."""
    import subprocess

    from nekwave_amd import build as b

    assert os.path.exists(b.KATS_BIN), "built by __graft_entry__.build()"
    r = subprocess.run([b.KATS_BIN, prec], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ALL PASSED" in r.stdout
