"""Benchmark of the DG-SEM Maxwell LSRK45 step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32] [--configs c3,c4,c5]

Headline workload (the configuration the metric is quoted on, "c2" in
nekwave_amd/workloads.py): periodic [0,1]^3, 16^3 hexahedra with interior
vertices jittered by +-10% h (seed 2), polynomial degree 7 (n1 = 8, 512
nodes/element, 2,097,152 nodes x 6 fields), Gaussian pulse, upwind flux,
LSRK45 at the reference's automatic dt.  A step is one LSRK45 step (5 fused
stage launches) over the whole mesh with the state resident in HBM.

Prints one JSON line: DOF-updates/s of the FP32 fused path (value: the median
of five back-to-back timings of exactly K steps, after an untimed clock-settling
run and W warm-up steps), the roofline of the fused stage kernel (algorithmic
bytes per launch / mean launch time from HIP events on the solver's stream), the
FP64 fused and FP64 reference-decomposed ("plain") rates, FP32 fields with FP64
LSRK registers, the FP32 relative L2 error against the FP64 path after 1000 steps
with the energy drift of both, the PEC-cube TM110 case (relative L2 to the
analytic mode, energy drift), the Case loop with on-device probes against the
download-every-step pattern, the other workloads with their stated errors (c3
PEC 32^3 with media, FP32 and split BF16; c4 256 batched cavities; c5 64^3 share
with the halo stand-in, FP32 and FP64) and the CPU baseline (the oracle's FP64
step on the host's threads, a bounded sample, one thread, and the full TM110
run).

Profiling modes (short runs for rocprofv3): --profile-run (a few steps without
graphs, then a 256 MiB copy as the byte-count calibration) and --profile-graph
(one workload's graph timing alone, with the tile option and k policy given for a
single mesh, tuned for the batch; --profile-workload picks the workload of either,
--profile-groups-per-cu caps the fused launches' workgroups per CU).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DOF-updates/s + % HBM roofline, FP32/batched vs FP64 plain, at stated L2 error"
HBM_PEAK_GBS = 8000.0
SIZES = {"fp64": (8, 8), "fp32": (4, 4), "fp32_k64": (4, 8), "bf16x3": (4, 4)}


def algorithmic_bytes_per_step(npts, nelt, s_u, s_k):
    """B_step = DOF (10 s_u + 9 s_k) + 5 nelt (24*8 + 2*8): per step, stage 1
    reads u, writes u and k; stages 2-5 read and write u and k; each stage reads
    the element's 8 FP64 corners and its eps, mu.  This is the reference's
    traffic, the figure every roofline fraction uses; the fused kernel's fifth
    stage does not store k (nothing reads it), so the kernel itself moves
    DOF (10 s_u + 8 s_k) plus the records."""
    return 6 * npts * (10 * s_u + 9 * s_k) + 5 * nelt * (24 * 8 + 2 * 8)


def make_solver(nw, wl, precision, pipeline="fused"):
    if len(wl.meshes) == 1:
        s = nw.DgSolver(precision=precision, pipeline=pipeline, use_graph=True)
        s.initialize(wl.meshes[0], wl.c0)
        s.uploadState(wl.states[0])
    else:
        s = nw.BatchSolver(precision=precision, use_graph=True)
        s.initialize(wl.meshes, wl.c0)
        for x, st in enumerate(wl.states):
            s.uploadState(x, st)
    return s


SETTLE_S = 0.3


def settle(s, dt, seconds=SETTLE_S):
    """Untimed steps for about `seconds` of device time, queued just before a
    timing, so that the timed region starts with the GPU at its sustained
    clocks: after host-side set-up the clocks ramp for tens of milliseconds, and
    a timing of a few milliseconds that starts right away measures the ramp (five
    back-to-back timings of 20 steps of the 16^3 mesh fell 12 % from first to
    last)."""
    per = s.time(dt, 10) / 10
    s.run(dt, max(1, int(seconds * 1e3 / per)))


def time_path(nw, wl, precision, pipeline, steps, warmup, repeats=5):
    """A clock-settling run (`settle`), `warmup` untimed steps, then `repeats`
    timings of exactly `steps` steps, queued back to back on the solver's stream
    (HIP events around each; no host upload or wait between the warm-up and the
    timings).  Returns the per-repeat milliseconds."""
    s = make_solver(nw, wl, precision, pipeline)
    settle(s, wl.dt)
    t0 = time.perf_counter()
    ms = s.timeRepeats(wl.dt, steps, repeats, warmup=warmup)
    wall = time.perf_counter() - t0
    info = s.info()
    segs = s.segments() if hasattr(s, "segments") else None
    s.finalize()
    return [float(v) for v in ms], wall, info, segs


def rel_l2(a, b, w):
    n = w.size
    num = sum(np.sum((a[c * n:(c + 1) * n] - b[c * n:(c + 1) * n]) ** 2 * w) for c in range(6))
    den = sum(np.sum(b[c * n:(c + 1) * n] ** 2 * w) for c in range(6))
    return math.sqrt(num / den)


def host_threads():
    """The host cores this process may use, as `nproc` counts them: OMP_NUM_THREADS
    when set, else the CPUs of the process's affinity mask."""
    import subprocess

    try:
        return int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.strip().isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_baseline(W, dt, budget_s=12.0):
    """The oracle (FP64 CPU restatement of the reference step) on the host
    cores: a bounded sample of the headline workload on all usable cores and one
    step of it on one thread; and the reference's CPU-runnable case in full (the
    PEC cube 4^3, p = 4, TM110, CFL 0.5, 200 steps) with its error against the
    analytic mode and its energy drift, on one thread and on all cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    threads = host_threads()
    o = po.Mesh.box(8, (16, 16, 16), (0, 1, 0, 1, 0, 1), (1, 1, 1), jitter=0.1, seed=2)
    x, y, z = o.x, o.y, o.zc
    zero = np.zeros_like(x)
    st = np.concatenate([zero, zero, np.exp(-((x - .5) ** 2 + (y - .5) ** 2 + (z - .5) ** 2) / 0.01), zero, zero, zero])
    s = po.Solver(o, 1.0)
    s.upload(st)
    po.set_threads(1)
    t_one = s.run(dt, 1)
    po.set_threads(threads)
    t1 = s.run(dt, 1)  # one step sizes the sample
    n = max(1, min(1000, int(budget_s / max(t1, 1e-3))))
    t = s.run(dt, n)
    del s
    # C1 in full: 200 steps
    c1 = W.c1()
    oc = po.Mesh.box(5, (4, 4, 4), (0, 1, 0, 1, 0, 1))
    st1 = W.pack(W.tm_mode(oc.x, oc.y))
    exact = W.pack(W.tm_mode(oc.x, oc.y, c1.steps * c1.dt))
    w1 = oc.jac * np.tile(oc.w3, oc.nelt)
    c1_out = {"steps": c1.steps, "dt": c1.dt, "nodes": oc.npts}
    for label, nthr in (("1_thread", 1), ("all_cores", threads)):
        po.set_threads(nthr)
        sc = po.Solver(oc, 1.0)
        sc.upload(st1)
        e0 = oc.energy(st1)
        tc = sc.run(c1.dt, c1.steps)
        got = sc.download()
        c1_out[label] = {"dof_updates_per_s": 6 * oc.npts * c1.steps / tc, "seconds": tc, "threads": nthr,
                         "rel_l2_vs_analytic": rel_l2(got, exact, w1), "energy_drift": (oc.energy(got) - e0) / e0}
    po.set_threads(threads)
    return {"value": 6 * o.npts * n / t, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "sample": "%d LSRK45 step(s) of the same 16^3 n1=8 jittered periodic mesh, FP64, %d OpenMP "
                      "threads (%s), %.1f s" % (n, threads, cpu_model(), t),
            "value_1_thread": 6 * o.npts / t_one, "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count(),
            "nproc": threads, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "c1_full_run": c1_out}


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", "traffic_%s.json" % name)
    if os.path.exists(p):
        return json.load(open(p))
    return None


def profile_graph(nw, W, precision, steps, option, stream_k, workload="c2", groups_per_cu=0):
    """The benchmark's timing of one workload alone, for a kernel trace: a single
    mesh with the given tile option and k policy (no autotuner launches), the
    batched cavities with the autotuner's per-segment picks (its timing launches
    come first); a short clock-settling run, 5 warm-up steps, then `steps` steps
    as captured graphs, timed with HIP events like the benchmark; prints the
    event mean per launch (every kernel of a step: with the halo stand-in, pack
    and unpack launches too) for comparison with the trace's mean kernel
    durations.  groups_per_cu caps the fused launches' workgroups per CU."""
    wl = getattr(W, workload)()
    if len(wl.meshes) == 1:
        s = nw.DgSolver(precision=precision, pipeline="fused", use_graph=True, tile_option=option,
                        groups_per_cu=groups_per_cu)
        s.initialize(wl.meshes[0], wl.c0)
        s.uploadState(wl.states[0])
        s.setSegmentOption(0, option, stream_k)
    else:
        s = nw.BatchSolver(precision=precision, use_graph=True, groups_per_cu=groups_per_cu)
        s.initialize(wl.meshes, wl.c0)
        for x, st in enumerate(wl.states):
            s.uploadState(x, st)
        s.tune()
        print("segments:", s.segments(), flush=True)
    settle(s, wl.dt, 0.1)
    ms = s.timeRepeats(wl.dt, steps, 3, warmup=5)
    info = s.info()
    s.finalize()
    per_launch = float(np.median(ms)) / (steps * info["kernels_per_step"])
    print("profile graph %s: %d steps x 3, %.4f ms per step, %d launches per step, event mean per launch %.2f us"
          % (workload, steps, float(np.median(ms)) / steps, info["kernels_per_step"], 1e3 * per_launch), flush=True)


def profile_run(nw, W, precision, steps, workload="c2", option=-1, stream_k=-1):
    """A few steps without graphs, so a kernel trace sees every launch.  A single
    mesh runs tile option `option` with non-temporal LSRK register accesses
    `stream_k` (1 / 0, -1 the default rule) -- the choices the autotuner makes in
    the timed benchmark, given explicitly so that every counter pass profiles the
    same kernel: under a counter pass the tuner's own timings are not
    meaningful; with option -1, and for the batched cavities, the autotuner picks per n1
    segment (its timing launches of every option come first; the batch runs one
    persistent launch per segment and stage)."""
    import ctypes

    wl = getattr(W, workload)()
    if len(wl.meshes) == 1:
        kw = {} if option < 0 else {"tile_option": option}
        s = nw.DgSolver(precision=precision, pipeline="fused", use_graph=False, **kw)
        s.initialize(wl.meshes[0], wl.c0)
        s.uploadState(wl.states[0])
        if option < 0:
            s.tune()
        else:
            s.setSegmentOption(0, option, stream_k)
    else:
        s = nw.BatchSolver(precision=precision, use_graph=False)
        s.initialize(wl.meshes, wl.c0)
        for x, st in enumerate(wl.states):
            s.uploadState(x, st)
        s.tune()
    print("segments:", s.segments(), flush=True)
    s.run(wl.dt, steps)
    s.finalize()
    ms = ctypes.c_double()
    nw._lib.check(nw.lib().nw_bench_copy(256 << 20, 3, ctypes.byref(ms)))
    print("profile run: %d steps (%d fused launches), 4 copies of 256 MiB" % (steps, 5 * steps), flush=True)


def tm110_c1(nw, W):
    """PEC unit cube 4^3, p = 4, CFL 0.5, TM110, 200 steps (the reference's CPU-runnable
    case): relative L2 against the analytic eigenmode and energy drift, FP64 fused and
    FP32 fused."""
    wl = W.c1()
    m = wl.meshes[0]
    x, y, _ = m.coords()
    w = m.getJac() * np.tile(m.getW3(), m.getNumElements())
    exact = W.pack(W.tm_mode(x, y, wl.steps * wl.dt))
    out = {"steps": wl.steps, "dt": wl.dt}
    for prec in ("fp64", "fp32"):
        s = make_solver(nw, wl, prec)
        e0 = s.diagnostics()["energy"]
        s.run(wl.dt, wl.steps)
        e1 = s.diagnostics()["energy"]
        got = s.downloadState()
        s.finalize()
        out[prec] = {"rel_l2_vs_analytic": rel_l2(got, exact, w), "energy_drift": (e1 - e0) / e0}
    return out


def case_loop(nw, wl, precision, steps=200, out_freq=10):
    """Case::simulate's loop on the same mesh, wall clock around the host loop:
    the reference's pattern (the whole state downloaded after every step to
    record probes and energy, case.cpp:586-598) against the on-device pattern
    (probes recorded inside each step's graph, the steps between outputs
    launched together, diagnostics and the probe history fetched every
    `out_freq` steps; the one download of the state at the end is reported
    apart)."""
    m = wl.meshes[0]
    dof = 6 * m.getTotalPoints()
    pts = np.array([[0.3, 0.5, 0.5], [0.5, 0.5, 0.5], [0.7, 0.5, 0.5], [0.5, 0.25, 0.75]])
    host = np.zeros(dof)
    out = {"steps": steps, "output_frequency": out_freq, "probes": len(pts)}
    s = make_solver(nw, wl, precision)
    s.run(wl.dt, 3)
    nw.lib().nw_dev_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(wl.dt)
        s.downloadState(host)
    t1 = time.perf_counter()
    s.finalize()
    out["download_every_step"] = {"dof_updates_per_s": dof * steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / steps}
    s = make_solver(nw, wl, precision)
    s.setProbes(m, pts, capacity=out_freq + 1)
    s.run(wl.dt, 3)
    s.probeHistory()
    nw.lib().nw_dev_sync()
    t0 = time.perf_counter()
    s.sampleProbes()
    nrec = 0
    for step in range(out_freq, steps + out_freq, out_freq):  # as Case: one launch per output interval
        s.run(wl.dt, min(step, steps) - (step - out_freq))
        s.diagnostics()
        nrec += s.probeHistory().shape[0]
    t1 = time.perf_counter()
    s.downloadState(host)  # once per run, at the end
    t2 = time.perf_counter()
    s.finalize()
    assert nrec == steps + 1
    out["on_device"] = {"dof_updates_per_s": dof * steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / steps,
                        "final_download_ms": 1e3 * (t2 - t1)}
    out["speedup"] = out["on_device"]["dof_updates_per_s"] / out["download_every_step"]["dof_updates_per_s"]
    return out


def run_state(nw, wl, precision, steps):
    """The workload advanced `steps` steps; returns the live solver."""
    s = make_solver(nw, wl, precision)
    s.run(wl.dt, steps)
    return s


def other_config(nw, W, name, steps, error_steps=True):
    """Rate of the fused FP32 path (and FP64 at the 64^3 share) on a workload,
    and its stated error: the FP32 state after the workload's full step count
    against the FP64 path's (relative L2, J w_i w_j w_k weighted, reduced on the
    device); for the batched cavities also every instance's FP32 and FP64 error
    against its closed-form TM_mn0 mode."""
    wl = W.BUILDERS[name]()
    out = {"nodes": wl.npts, "elements": wl.nelt, "instances": len(wl.meshes), "description": wl.info["description"],
           "dt": wl.dt}
    precs = {"c5": ["fp32", "fp64"], "c3": ["fp32", "bf16x3"]}.get(name, ["fp32"])
    for prec in precs:
        reps, _, info, segs = time_path(nw, wl, prec, "fused", steps, 2, repeats=3)
        ms = float(np.median(reps))
        s_u, s_k = SIZES[prec]
        b = algorithmic_bytes_per_step(wl.npts, wl.nelt, s_u, s_k)
        out[prec] = {"dof_updates_per_s": 6 * wl.npts * steps / (ms * 1e-3), "ms_per_step": ms / steps,
                     "ms_per_step_repeats": [r / steps for r in reps],
                     "algorithmic_GBps": b / (ms / steps * 1e-3) / 1e9,
                     "hbm_frac": b / (ms / steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "kernels_per_step": info["kernels_per_step"]}
        if segs:
            out[prec]["segments"] = segs
    if error_steps:
        n = wl.steps
        s64 = run_state(nw, wl, "fp64", n)
        ref = [s64.downloadState() if len(wl.meshes) == 1 else s64.downloadState(x) for x in range(len(wl.meshes))]
        s32 = run_state(nw, wl, "fp32", n)
        errs = [s32.relL2(ref[x], x) for x in range(len(wl.meshes))]
        err = {"steps": n, "fp32_vs_fp64_rel_l2": max(errs)}
        if "bf16x3" in precs:
            sb = run_state(nw, wl, "bf16x3", n)
            err["bf16x3_vs_fp64_rel_l2"] = sb.relL2(ref[0], 0)
            sb.finalize()
        if len(wl.meshes) > 1:
            err["fp32_vs_fp64_rel_l2_median"] = float(np.median(errs))
            an32, an64 = [], []
            t = n * wl.dt
            for x, (k, p, eps, mm, nn, ph) in enumerate(wl.info["instances"]):
                xs, ys, _ = wl.meshes[x].coords()
                exact = W.pack(W.tm_mode(xs, ys, t, mm, nn, eps, 1.0, ph))
                an32.append(s32.relL2(exact, x))
                an64.append(s64.relL2(exact, x))
            err["fp32_vs_analytic_rel_l2_max"] = max(an32)
            err["fp32_vs_analytic_rel_l2_median"] = float(np.median(an32))
            err["fp64_vs_analytic_rel_l2_max"] = max(an64)
            k, p_, eps, mm, nn, _ = wl.info["instances"][int(np.argmax(an64))]
            err["worst_instance_vs_analytic"] = {"k": k, "p": p_, "eps": eps, "m": mm, "n": nn,
                                                 "note": "discretization error of the coarsest meshes (k elements of "
                                                         "degree p per side for m, n half-waves), equal in FP32 and FP64"}
            err["fp64_vs_analytic_rel_l2_median"] = float(np.median(an64))
            err["fp32_minus_fp64_analytic_error_max"] = max(abs(a - b) for a, b in zip(an32, an64))
        s32.finalize()
        s64.finalize()
        out["error"] = err
    del wl
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp32_k64", "fp64", "bf16x3"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the full-length FP32-vs-FP64 runs of the other workloads")
    ap.add_argument("--accuracy-steps", type=int, default=1000)
    ap.add_argument("--configs", default="c3,c4,c5", help="other workloads to time (comma list, or 'none')")
    ap.add_argument("--profile-workload", default="c2", choices=["c2", "c3", "c4", "c5"],
                    help="workload of --profile-run and --profile-graph")
    ap.add_argument("--profile-groups-per-cu", type=int, default=0,
                    help="cap on the fused launches' workgroups per CU in --profile-graph (0: as many as fit)")
    ap.add_argument("--profile-option", type=int, default=-1,
                    help="tile option of --profile-run on a single mesh (-1: the autotuner's pick)")
    ap.add_argument("--profile-stream", type=int, default=-1, choices=[-1, 0, 1],
                    help="non-temporal k accesses of --profile-run with --profile-option (-1: default rule)")
    ap.add_argument("--profile-graph", action="store_true",
                    help="the headline timing alone (tile option --profile-option, k policy --profile-stream, "
                         "no autotuner), for a kernel trace of its graph launches")
    ap.add_argument("--profile-run", action="store_true",
                    help="short run for rocprofv3 counter passes: the fused stage on the workload "
                         "plus a 256 MiB device copy as the HBM byte-count calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: nekwave_amd runs on one GPU (the multi-GPU halo path is a design, not code); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import nekwave_amd as nw
    from nekwave_amd import workloads as W

    if args.profile_graph:
        profile_graph(nw, W, args.precision, args.steps, max(args.profile_option, 0), args.profile_stream,
                      workload=args.profile_workload, groups_per_cu=args.profile_groups_per_cu)
        return
    if args.profile_run:
        profile_run(nw, W, args.precision, steps=min(args.steps, 4), workload=args.profile_workload,
                    option=args.profile_option, stream_k=args.profile_stream)
        return

    dev = nw.device_info(0)
    wl = W.c2()
    m = wl.meshes[0]
    npts, nelt, dt = wl.npts, wl.nelt, wl.dt
    dof = 6 * npts

    # stated error: the measured precision vs the FP64 path after 1000 steps,
    # and the energy drift (U_end - U_0) / U_0 of both (upwind: dissipative)
    acc = drift = None
    if args.accuracy_steps > 0 and args.precision != "fp64":
        outs, drift = {}, {}
        for prec in (args.precision, "fp64"):
            s = make_solver(nw, wl, prec)
            e0 = s.diagnostics()["energy"]
            s.run(dt, args.accuracy_steps)
            drift[prec] = (s.diagnostics()["energy"] - e0) / e0
            outs[prec] = s.downloadState()
            s.finalize()
        w = m.getJac() * np.tile(m.getW3(), nelt)  # dV = J w3 (3dboxpec_test.cpp:84)
        acc = rel_l2(outs[args.precision], outs["fp64"], w)

    # comparison paths (fewer steps: they are slower)
    cmp_steps = max(10, args.steps // 5)
    ms64 = float(np.median(time_path(nw, wl, "fp64", "fused", cmp_steps, 5, repeats=3)[0]))
    ms64p = float(np.median(time_path(nw, wl, "fp64", "plain", cmp_steps, 5, repeats=3)[0]))
    v64 = dof * cmp_steps / (ms64 * 1e-3)
    v64p = dof * cmp_steps / (ms64p * 1e-3)
    k64 = None
    if args.precision == "fp32":  # FP32 fields with FP64 LSRK registers (112 B per DOF-update)
        msk = float(np.median(time_path(nw, wl, "fp32_k64", "fused", cmp_steps, 5, repeats=3)[0]))
        b_k64 = algorithmic_bytes_per_step(npts, nelt, 4, 8)
        k64 = {"value": dof * cmp_steps / (msk * 1e-3), "ms_per_step": msk / cmp_steps,
               "hbm_frac": b_k64 * cmp_steps / (msk * 1e-3) / 1e9 / HBM_PEAK_GBS}

    # headline, timed after the runs above (the GPU at its working clocks): W
    # warm-up steps, then 5 back-to-back timings of exactly K steps; the value is
    # their median
    reps, wall, sinfo, segs = time_path(nw, wl, args.precision, "fused", args.steps, args.warmup, repeats=5)
    ms = float(np.median(reps))
    value = dof * args.steps / (ms * 1e-3)
    s_u, s_k = SIZES[args.precision]
    b_step = algorithmic_bytes_per_step(npts, nelt, s_u, s_k)
    launches = sinfo["kernels_per_step"] * args.steps
    per_launch_ms = ms / launches
    achieved = (b_step / 5) / (per_launch_ms * 1e-3) / 1e9
    traffic = load_traffic("fused_stage_%s_c2" % args.precision)

    c1 = tm110_c1(nw, W)
    loop = case_loop(nw, wl, args.precision)

    others = {}
    if args.configs and args.configs != "none":
        for name in args.configs.split(","):
            others[name] = other_config(nw, W, name.strip(), 20, error_steps=not args.no_errors)

    cpu = None if args.no_cpu else cpu_baseline(W, dt)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "timing": "median of 5 back-to-back timings of exactly `steps` steps after an untimed %.1f s "
                  "clock-settling run and `warmup` untimed steps (HIP events on the solver's stream)" % SETTLE_S,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "fp32", "fp32_k64": "fp32+fp64k", "fp64": "fp64", "bf16x3": "fp32+bf16x3"}[args.precision],
        "data": "synthetic",
        "config": {"workload": "periodic [0,1]^3 16^3 hexes, interior vertices jittered +-10% h (seed 2), "
                               "N=7 (n1=8), Gaussian Ez pulse, upwind, LSRK45 auto dt",
                   "nodes": npts, "dof": dof, "elements": nelt, "dt": dt, "pipeline": "fused",
                   "precision": args.precision, "elements_per_workgroup": sinfo["epw"],
                   "tile_option": segs[0]["option"],
                   "volume_contractions": "matrix cores" if segs[0]["matrix_volume"] else "vector ALUs",
                   "lsrk_register_nontemporal": segs[0]["stream_k"],
                   "device": dev["arch"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "kernel": "k_fused_stage<8,*,*,%s>" % args.precision,
                     "algorithmic_bytes_per_launch": b_step / 5, "mean_launch_ms": per_launch_ms},
        "fp64_fused": {"value": v64, "ms_per_step": ms64 / cmp_steps},
        "fp64_plain": {"value": v64p, "ms_per_step": ms64p / cmp_steps},
        "fp32_fields_fp64_registers": k64,
        "speedup_vs_fp64_plain": value / v64p,
        "rel_l2_vs_fp64": acc,
        "rel_l2_steps": args.accuracy_steps,
        "energy_drift": drift,
        "ms_per_step_repeats": [r / args.steps for r in reps],
        "tm110_c1": c1,
        "case_loop": loop,
        "other_configs": others,
        "wall_s_timed": wall,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
