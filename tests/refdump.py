"""Reader/runner for oracle/_ref/ref_mesh_dump (reference mesh.cpp built in place).

Test infrastructure: the dumper is built by oracle/Makefile only where the
reference tree exists; tests that need it skip when it is absent.
"""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DUMPER = os.path.join(ROOT, "oracle", "_ref", "ref_mesh_dump")
PROBE_DUMPER = os.path.join(ROOT, "oracle", "_ref", "ref_probe_dump")


def available():
    return os.path.exists(DUMPER)


def probes_available():
    return os.path.exists(PROBE_DUMPER)


def probes(tmpdir, mesh_args, pts, name="probes"):
    """ProbeManager::init of the reference on a box (mesh_args = ("box", n1, nel,
    bounds)) or a .rea file (("rea", n1, path)): per probe the element offset
    and the n1^3 interpolation weights."""
    pp = os.path.join(str(tmpdir), name + "_pts.bin")
    op = os.path.join(str(tmpdir), name + "_out.bin")
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    with open(pp, "wb") as fh:
        fh.write(np.array([pts.shape[0]], np.int32).tobytes() + pts.tobytes())
    if mesh_args[0] == "box":
        _, n1, nel, bounds = mesh_args
        args = [PROBE_DUMPER, "box", str(n1)] + [str(v) for v in nel] + [repr(float(b)) for b in bounds]
    else:
        _, n1, path = mesh_args
        args = [PROBE_DUMPER, "rea", str(n1), path]
    subprocess.run(args + [pp, op], check=True, stdout=subprocess.DEVNULL)
    raw = open(op, "rb").read()
    n, np_ = np.frombuffer(raw, np.int32, 2)
    offs, wts, off = [], [], 8
    for _ in range(int(n)):
        offs.append(int(np.frombuffer(raw, np.int32, 1, off)[0]))
        wts.append(np.frombuffer(raw, np.float64, int(np_), off + 4).copy())
        off += 4 + 8 * int(np_)
    return np.array(offs), np.array(wts)


def _read(path):
    raw = open(path, "rb").read()
    off = 0

    def take(n, dt):
        nonlocal off
        a = np.frombuffer(raw, dtype=dt, count=n, offset=off).copy()
        off += n * np.dtype(dt).itemsize
        return a

    n1, nelt, npts, tfp = take(4, np.int32)
    out = {"n1": int(n1), "nelt": int(nelt), "npts": int(npts), "tfp": int(tfp)}
    out["z"] = take(n1, np.float64)
    out["w"] = take(n1, np.float64)
    out["D"] = take(n1 * n1, np.float64)
    out["Dt"] = take(n1 * n1, np.float64)
    out["w3"] = take(n1 ** 3, np.float64)
    out["corners"] = take(nelt * 24, np.float64)
    for k in ("x", "y", "zc", "jac", "rx", "sx", "tx", "ry", "sy", "ty", "rz", "sz", "tz"):
        out[k] = take(npts, np.float64)
    for k in ("vM", "vP", "bc"):
        out[k] = take(tfp, np.int32)
    for k in ("nx", "ny", "nz", "dA"):
        out[k] = take(tfp, np.float64)
    out["dxmin"], out["auto_cfl"], out["auto_dt"] = take(3, np.float64)
    assert off == len(raw)
    return out


def box(tmpdir, n1, nel, bounds, periodic=(0, 0, 0)):
    path = os.path.join(str(tmpdir), "box_%d.bin" % n1)
    args = [DUMPER, "box", str(n1)] + [str(v) for v in nel] + [repr(float(b)) for b in bounds] \
        + [str(int(p)) for p in periodic] + [path]
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return _read(path)


BC_NAMES = {0: "E", 1: "PERIODIC", 2: "PEC", 3: "PMC"}


def write_rea(path, corners, kind, nbr_e, nbr_f):
    """NekCEM .rea text in the layout Mesh::loadFromRea parses (mesh.cpp:317-445)."""
    corners = np.asarray(corners, dtype=np.float64).reshape(-1, 8, 3)
    nelt = corners.shape[0]
    lines = ["generated mesh", "**MESH DATA**", "%d 3 %d NEL,NDIM,NELV" % (nelt, nelt)]
    for e in range(nelt):
        lines.append("ELEMENT %d [ 1a] GROUP 0" % (e + 1))
        c = corners[e]
        for lo in (0, 4):
            for d in range(3):
                lines.append(" ".join(repr(float(c[lo + a, d])) for a in range(4)))
    lines.append("***** CURVED SIDE DATA *****")
    lines.append("***** BOUNDARY CONDITIONS *****")
    lines.append("***** FLUID BOUNDARY CONDITIONS *****")
    kind = np.asarray(kind).reshape(-1)
    for fi in range(nelt * 6):
        e, f = divmod(fi, 6)
        ne = int(nbr_e[fi]) + 1 if nbr_e[fi] >= 0 else 0
        nf = int(nbr_f[fi]) + 1 if nbr_f[fi] >= 0 else 0
        lines.append("%s %d %d %d.0 %d.0" % (BC_NAMES[int(kind[fi])], e + 1, f + 1, ne, nf))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def rea(tmpdir, n1, corners, kind, nbr_e, nbr_f, name="mesh"):
    rp = os.path.join(str(tmpdir), name + ".rea")
    bp = os.path.join(str(tmpdir), name + ".bin")
    write_rea(rp, corners, kind, nbr_e, nbr_f)
    subprocess.run([DUMPER, "rea", str(n1), rp, bp], check=True, stdout=subprocess.DEVNULL)
    return _read(bp)


def write_rea_tags(path, corners, tags, nbr_e, nbr_f, nel_line=None):
    """.rea text with raw boundary tags (one per face, e-major) and 1-based
    neighbour columns (0 = none), as NekCEM writes them."""
    corners = np.asarray(corners, dtype=np.float64).reshape(-1, 8, 3)
    nelt = corners.shape[0]
    lines = ["generated mesh", "**MESH DATA**", nel_line or ("%d 3 %d NEL,NDIM,NELV" % (nelt, nelt))]
    for e in range(nelt):
        lines.append("ELEMENT %d [ 1a] GROUP 0" % (e + 1))
        c = corners[e]
        for lo in (0, 4):
            for d in range(3):
                lines.append(" ".join(repr(float(c[lo + a, d])) for a in range(4)))
    lines.append("***** CURVED SIDE DATA *****")
    lines.append("***** BOUNDARY CONDITIONS *****")
    lines.append("***** FLUID BOUNDARY CONDITIONS *****")
    for fi in range(nelt * 6):
        e, f = divmod(fi, 6)
        ne = int(nbr_e[fi]) + 1 if nbr_e[fi] >= 0 else 0
        nf = int(nbr_f[fi]) + 1 if nbr_f[fi] >= 0 else 0
        lines.append(" %-3s %d %d %d.00000 %d.00000 0.00000 0.00000 0.00000" % (tags[fi], e + 1, f + 1, ne, nf))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_re2(path, corners, records, ncurv=0, nel_in_header=True):
    """Binary .re2 in the layout Mesh::loadFromRe2 reads (mesh.cpp:439-636):
    80-byte header, float32 6.54321, per element group + x[8] + y[8] + z[8]
    (float64), ncurv, nbc, then per record e, f, ne, nf, p1, p2, p3 (float64,
    1-based, 0 = none) and an 8-byte tag."""
    corners = np.asarray(corners, dtype=np.float64).reshape(-1, 8, 3)
    nelt = corners.shape[0]
    hdr = ("#v002 %d 3 %d" % (nelt, nelt)) if nel_in_header else "#v002"
    buf = [hdr.ljust(80).encode(), np.array([6.54321], np.float32).tobytes()]
    for e in range(nelt):
        buf.append(np.concatenate([[0.0], corners[e, :, 0], corners[e, :, 1], corners[e, :, 2]]).tobytes())
    buf.append(np.array([float(ncurv), float(len(records))]).tobytes())
    for (e, f, ne, nf, tag) in records:
        buf.append(np.array([e + 1, f + 1, ne + 1 if ne >= 0 else 0, nf + 1 if nf >= 0 else 0, 0, 0, 0],
                            np.float64).tobytes())
        buf.append(tag.encode().ljust(8)[:8])
    with open(path, "wb") as fh:
        fh.write(b"".join(buf))


def dump_file(tmpdir, mode, n1, mesh_path, name="dump"):
    bp = os.path.join(str(tmpdir), name + ".bin")
    r = subprocess.run([DUMPER, mode, str(n1), mesh_path, bp], stdout=subprocess.DEVNULL)
    if r.returncode != 0:
        return None
    return _read(bp)
