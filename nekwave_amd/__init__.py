"""nekwave_amd -- MI355X-native DG-SEM time-domain Maxwell step.

Python mirror of the reference's host interface over the library's C ABI
(include/nekwave_amd.h):

    Mesh       -- reference Mesh (src/mesh.hpp:51-218): createBoxMesh,
                  loadFromRea/loadFromRe2, GLL data, coordinates, dt helpers;
                  extensions for jitter, media and the single-GPU halo stand-in.
    DgSolver   -- reference DgSolver (src/dg_solver.hpp:35-102): initialize,
                  uploadState, downloadState, step, finalize; extensions for
                  precision, pipeline, graphs, timed runs and device diagnostics.
    Case       -- reference Case lifecycle (src/case.hpp:48-153):
                  preprocess / simulate / postprocess over a Config subset.

All numerics run in the HIP library; nothing here computes the step.
"""
import ctypes as C
import os
import sys

import numpy as np

from . import _lib
from ._lib import FACE_E, FACE_PEC, FACE_PERIODIC, FACE_PMC, NekWaveError, check, lib

__all__ = ["Mesh", "DgSolver", "BatchSolver", "Case", "Config", "DeviceBuffer", "FieldWriter", "NekWaveError", "device_info",
           "FACE_E", "FACE_PERIODIC", "FACE_PEC", "FACE_PMC"]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def device_info(device=0):
    n = C.c_int(0)
    name = C.create_string_buffer(64)
    cus = C.c_int(0)
    mem = C.c_size_t(0)
    check(lib().nw_device_info(device, C.byref(n), name, 64, C.byref(cus), C.byref(mem)))
    return {"count": n.value, "arch": name.value.decode(), "cus": cus.value, "mem_bytes": mem.value}


class Mesh:
    """Reference Mesh: construct with the number of GLL points per direction N
    (the reference's `order`), then build with createBoxMesh (or fromCorners)."""

    def __init__(self, N=3, numElements=1):
        self._N = int(N)
        self._h = None
        self._nel0 = int(numElements)

    # -- construction -------------------------------------------------------
    def setN(self, N):
        if self._h is not None:
            raise NekWaveError(_lib.NW_ERR_STATE, "setN after the mesh is built")
        self._N = int(N)

    def createBoxMesh(self, nelx, nely, nelz, xmin=-1.0, xmax=1.0, ymin=-1.0, ymax=1.0, zmin=-1.0, zmax=1.0,
                      periodicX=False, periodicY=False, periodicZ=False, jitter=0.0, seed=0):
        self._free()
        b = np.array([xmin, xmax, ymin, ymax, zmin, zmax], dtype=np.float64)
        h = C.c_void_p()
        check(lib().nw_mesh_box(self._N, int(nelx), int(nely), int(nelz), _dp(b), int(bool(periodicX)),
                                int(bool(periodicY)), int(bool(periodicZ)), float(jitter), int(seed), C.byref(h)))
        self._h = h
        self.box = (int(nelx), int(nely), int(nelz))
        return True

    @classmethod
    def fromCorners(cls, N, corners, kind, nbr_elem, nbr_face):
        m = cls(N)
        corners = np.ascontiguousarray(corners, dtype=np.float64).reshape(-1)
        kind = np.ascontiguousarray(kind, dtype=np.int32).reshape(-1)
        ne = np.ascontiguousarray(nbr_elem, dtype=np.int32).reshape(-1)
        nf = np.ascontiguousarray(nbr_face, dtype=np.int32).reshape(-1)
        h = C.c_void_p()
        check(lib().nw_mesh_from_corners(m._N, corners.size // 24, _dp(corners), _ip(kind), _ip(ne), _ip(nf),
                                         C.byref(h)))
        m._h = h
        return m

    def load(self, filename, periodic_p=False):
        """Read a NekCEM mesh file (.re2 when the name contains ".re2", else .rea);
        raises NekWaveError on failure."""
        self._free()
        h = C.c_void_p()
        check(lib().nw_mesh_load(os.fsencode(filename), self._N,
                                 _lib.NW_MESH_REA_PERIODIC if periodic_p else 0, C.byref(h)))
        self._h = h
        return True

    def loadFromRea(self, filename, periodic_p=False):
        """Mesh::loadFromRea (mesh.cpp:317-436): True on success; on failure the
        message goes to stderr and False is returned, as in the reference."""
        try:
            return self.load(filename, periodic_p)
        except NekWaveError as e:
            print("Error: %s" % e, file=sys.stderr)
            return False

    def loadFromRe2(self, filename):
        """Mesh::loadFromRe2 (mesh.cpp:439-636); True/False like loadFromRea."""
        if ".re2" not in filename:
            print("Error: %s is not a .re2 file name" % filename, file=sys.stderr)
            return False
        return self.loadFromRea(filename)

    def getFaces(self):
        """The face list (kind NW_FACE_*, neighbour element, neighbour face), nelt x 6 each."""
        n = 6 * self.getNumElements()
        k = np.zeros(n, np.int32); ne = np.zeros(n, np.int32); nf = np.zeros(n, np.int32)
        check(lib().nw_mesh_faces(self.handle, _ip(k), _ip(ne), _ip(nf)))
        return k.reshape(-1, 6), ne.reshape(-1, 6), nf.reshape(-1, 6)

    def setMaterials(self, eps=None, mu=None):
        e = None if eps is None else np.ascontiguousarray(eps, dtype=np.float64)
        u = None if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        for a in (e, u):
            if a is not None and a.size != self.getNumElements():
                raise NekWaveError(_lib.NW_ERR_ARG, "materials need one value per element")
        check(lib().nw_mesh_set_materials(self.handle, _dp(e) if e is not None else None,
                                          _dp(u) if u is not None else None))

    def selfHaloX(self):
        n = C.c_int(0)
        check(lib().nw_mesh_self_halo_x(self.handle, C.byref(n)))
        return n.value

    # -- queries --------------------------------------------------------------
    @property
    def handle(self):
        if self._h is None:
            raise NekWaveError(_lib.NW_ERR_STATE, "mesh not built")
        return self._h

    def info(self):
        o = _lib.MeshInfo()
        check(lib().nw_mesh_get_info(self.handle, C.byref(o)))
        return o

    def getN(self):
        return self._N

    def getNumPointsPerElement(self):
        return self._N ** 3

    def getNumElements(self):
        return self.info().nelt if self._h is not None else self._nel0

    def getTotalPoints(self):
        return self.info().npts

    def getNumHaloPoints(self):
        return self.info().nhalo

    def isPartitioned(self):
        return self.getNumHaloPoints() > 0

    def _gll(self):
        n = self._N
        z = np.zeros(n); w = np.zeros(n); D = np.zeros(n * n); w3 = np.zeros(n ** 3)
        check(lib().nw_mesh_gll(self.handle, _dp(z), _dp(w), _dp(D), _dp(w3)))
        return z, w, D, w3

    def getGllZ(self):
        return self._gll()[0]

    def getGllW(self):
        return self._gll()[1]

    def getD(self):
        return self._gll()[2]

    def getW3(self):
        return self._gll()[3]

    def coords(self):
        n = self.getTotalPoints()
        x = np.zeros(n); y = np.zeros(n); z = np.zeros(n)
        check(lib().nw_mesh_coords(self.handle, _dp(x), _dp(y), _dp(z)))
        return x, y, z

    def getCoordX(self):
        return self.coords()[0]

    def getCoordY(self):
        return self.coords()[1]

    def getCoordZ(self):
        return self.coords()[2]

    def getJac(self):
        j = np.zeros(self.getTotalPoints())
        check(lib().nw_mesh_jacobian(self.handle, _dp(j)))
        return j

    def getElementCorners(self):
        c = np.zeros(self.getNumElements() * 24)
        check(lib().nw_mesh_corners(self.handle, _dp(c)))
        return c.reshape(-1, 8, 3)

    def locateProbes(self, coords):
        """ProbeManager::init's placement (probe.cpp:57-96): per probe the
        element index and the clamped reference coordinates (r, s, t)."""
        xyz = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        elem = np.zeros(n, dtype=np.int32)
        rst = np.zeros((n, 3))
        check(lib().nw_mesh_locate_probes(self.handle, n, _dp(xyz), _ip(elem), _dp(rst)))
        return elem, rst

    def computeMinNodeDistance(self):
        v = C.c_double()
        check(lib().nw_mesh_min_node_distance(self.handle, C.byref(v)))
        return v.value

    @staticmethod
    def computeAutomaticCFL(safetyFactor=0.4):
        return (safetyFactor * 1.75) / np.sqrt(3.0)

    def computeAutomaticDt(self, waveSpeed=1.0, safetyFactor=0.4):
        v = C.c_double()
        check(lib().nw_mesh_auto_dt(self.handle, float(waveSpeed), float(safetyFactor), C.byref(v)))
        return v.value

    def _free(self):
        if self._h is not None:
            lib().nw_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass


class DgSolver:
    """Reference DgSolver.  Extensions: precision ('fp64', 'fp32', 'fp32_k64'),
    pipeline ('fused' or 'plain' = the reference's 7-kernel stage), graphs, the
    fused tile option (-1: tuned) and groups_per_cu, a cap on the fused launch's
    workgroups per CU (0: as many as fit)."""

    def __init__(self, precision="fp64", pipeline="fused", use_graph=True, device=0, tile_option=-1,
                 groups_per_cu=0):
        self.opts = _lib.Opts()
        lib().nw_opts_default(C.byref(self.opts))
        self.opts.precision = _lib.PREC[precision]
        self.opts.pipeline = _lib.PIPE[pipeline]
        self.opts.use_graph = int(bool(use_graph))
        self.opts.device = int(device)
        self.opts.tile_option = int(tile_option)
        self.opts.groups_per_cu = int(groups_per_cu)
        self._h = None
        self.npts = 0

    def initialize(self, mesh, c0):
        self.finalize()
        h = C.c_void_p()
        check(lib().nw_dg_create(mesh.handle, float(c0), C.byref(self.opts), C.byref(h)))
        self._h = h
        self.npts = mesh.getTotalPoints()
        return True

    def uploadState(self, hostState, size=None):
        a = np.ascontiguousarray(hostState, dtype=np.float64).reshape(-1)
        n = a.size if size is None else int(size)
        check(lib().nw_dg_upload(self._handle(), _dp(a), n))

    def downloadState(self, hostState=None, size=None):
        n = 6 * self.npts if size is None else int(size)
        out = np.zeros(n) if hostState is None else hostState
        if out.dtype != np.float64 or not out.flags.c_contiguous:
            raise NekWaveError(_lib.NW_ERR_ARG, "downloadState needs a contiguous float64 array")
        check(lib().nw_dg_download(self._handle(), _dp(out), n))
        return out

    def step(self, dt, time=0.0):
        check(lib().nw_dg_step(self._handle(), float(dt), float(time)))

    def run(self, dt, nsteps):
        check(lib().nw_dg_run(self._handle(), float(dt), int(nsteps)))

    def time(self, dt, nsteps):
        ms = C.c_double()
        check(lib().nw_dg_time(self._handle(), float(dt), int(nsteps), C.byref(ms)))
        return ms.value

    def timeRepeats(self, dt, nsteps, repeats, warmup=0):
        """`warmup` untimed steps, then `repeats` back-to-back timings of exactly
        `nsteps` steps (ms each, HIP events on the solver's stream)."""
        ms = np.zeros(int(repeats))
        check(lib().nw_dg_time_repeats(self._handle(), float(dt), int(warmup), int(nsteps), int(repeats), _dp(ms)))
        return ms

    def diagnostics(self):
        o = _lib.Diag()
        check(lib().nw_dg_diag(self._handle(), C.byref(o)))
        return {"energy": o.energy, "max_e": o.max_e, "max_h": o.max_h}

    def relL2(self, reference, instance=0):
        """Relative L2 distance of the current state to `reference` (uploadState
        layout), weighted by J w_i w_j w_k as 3dboxpec_test.cpp:76-119, reduced
        on the device in FP64 (nw_dg_rel_l2)."""
        a = np.ascontiguousarray(reference, dtype=np.float64).reshape(-1)
        out = C.c_double()
        check(lib().nw_dg_rel_l2(self._handle(), int(instance), _dp(a), a.size, C.byref(out)))
        return out.value

    def setProbes(self, mesh, coords, capacity=1024):
        """On-device ProbeManager (probe.cpp:33-216): place probes at `coords`
        ((n, 3) array) on `mesh`, the mesh this solver was created from; every
        later step appends one record to a device history of `capacity` records."""
        xyz = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
        check(lib().nw_dg_set_probes(self._handle(), mesh.handle, xyz.shape[0], _dp(xyz), int(capacity)))
        self._nprobes = xyz.shape[0]
        self._probe_cap = int(capacity)

    def sampleProbes(self):
        """Append one record of the current state (nw_dg_probe_sample)."""
        check(lib().nw_dg_probe_sample(self._handle()))

    def probeHistory(self):
        """Records taken since the last call, oldest first: array (records,
        probes, 6) with components Ex Ey Ez Hx Hy Hz."""
        n = getattr(self, "_nprobes", 0)
        cap = max(getattr(self, "_probe_cap", 1), 1)
        out = np.zeros(cap * max(n, 1) * 6)
        nrec = C.c_int()
        check(lib().nw_dg_probe_history(self._handle(), _dp(out), cap, C.byref(nrec)))
        return out[: nrec.value * n * 6].reshape(nrec.value, n, 6).copy()

    def info(self):
        o = _lib.DgInfo()
        check(lib().nw_dg_get_info(self._handle(), C.byref(o)))
        return {f: getattr(o, f) for f, _ in o._fields_}

    def finalize(self):
        if self._h is not None:
            lib().nw_dg_destroy(self._h)
            self._h = None

    def _handle(self):
        if self._h is None:
            raise NekWaveError(_lib.NW_ERR_STATE, "solver not initialized")
        return self._h

    def tune(self):
        """Time every distinct tile option per segment (with and without
        non-temporal k accesses where the working set exceeds the Infinity
        Cache) and keep the fastest (nw_dg_tune)."""
        check(lib().nw_dg_tune(self._handle()))

    def setSegmentOption(self, seg, option, stream_k=-1):
        """Set segment `seg`'s tile option and non-temporal k accesses (1, 0, or
        -1 for the default rule), e.g. as an earlier process tuned them
        (`segments()`); autotuning then keeps them until `tune()`."""
        check(lib().nw_dg_set_segment_option(self._handle(), int(seg), int(option), int(stream_k)))

    def setWalk(self, mode):
        """Walk direction of the fused stages (nw_dg_set_walk): 1 every other
        stage backwards (a batch then also takes its segments in reverse
        order), 0 all forwards, -1 the default rule (alternate where the largest
        segment's working set exceeds the Infinity Cache by at most 8x)."""
        check(lib().nw_dg_set_walk(self._handle(), int(mode)))

    def alternatingWalk(self):
        """Whether every other fused stage walks backwards (nw_dg_walk)."""
        v = C.c_int()
        check(lib().nw_dg_walk(self._handle(), C.byref(v)))
        return bool(v.value)

    def segments(self):
        """Per n1 segment: n1, elements, elements per workgroup, tile option,
        whether it runs the volume contractions on the matrix cores and whether
        its LSRK register accesses are non-temporal."""
        out = []
        for s in range(self.info()["segments"]):
            n1, nelt, epw = C.c_int(), C.c_int(), C.c_int()
            check(lib().nw_dg_segment_info(self._handle(), s, C.byref(n1), C.byref(nelt), C.byref(epw)))
            opt, mx, sk = C.c_int(), C.c_int(), C.c_int()
            check(lib().nw_dg_segment_option(self._handle(), s, C.byref(opt), C.byref(mx), C.byref(sk)))
            out.append({"n1": n1.value, "nelt": nelt.value, "epw": epw.value, "option": opt.value,
                        "matrix_volume": bool(mx.value), "stream_k": bool(sk.value)})
        return out

    def __del__(self):
        try:
            self.finalize()
        except Exception:
            pass


class DeviceBuffer:
    """A device allocation owned by the library's allocator (for the kernel-level ABI)."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        check(lib().nw_dev_alloc(self.nbytes, C.byref(self.ptr)))

    @classmethod
    def of(cls, arr):
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(lib().nw_dev_h2d(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def download(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        check(lib().nw_dev_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def offset(self, nbytes):
        return C.c_void_p(self.ptr.value + int(nbytes))

    def free(self):
        if self.ptr:
            lib().nw_dev_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FieldWriter:
    """Hdf5Writer (hdf5_writer.hpp:22-120) as the reference runs it without the
    HDF5 library: initialize() writes <dir>/<stem>.bin's header, writeStep()
    appends the step to the archive and writes <stem>_step_<n>.vtu and
    <stem>.pvd.  Save options: "E", "H", "magnitude_E", "magnitude_H",
    "energy_density" (the derived curl/div fields are not offered)."""

    FLAGS = {"E": 1, "H": 2, "magnitude_E": 4, "magnitude_H": 8, "energy_density": 16}

    def __init__(self):
        self._h = None
        self._flags = self.FLAGS["E"] | self.FLAGS["H"]

    def setFieldSaveOptions(self, fields):
        self._flags = 0
        for f in fields:
            self._flags |= self.FLAGS[f]

    def initialize(self, h5_path, mesh, enable_xdmf=True, export_continuous=False):
        self.close()
        h = C.c_void_p()
        check(lib().nw_fields_open(os.fsencode(h5_path), mesh.handle, int(bool(export_continuous)), self._flags,
                                   C.byref(h)))
        self._h = h
        return True

    def writeStep(self, step, time, state):
        if self._h is None:
            raise NekWaveError(_lib.NW_ERR_STATE, "writeStep before initialize")
        st = np.ascontiguousarray(state, dtype=np.float64)
        check(lib().nw_fields_write(self._h, int(step), float(time), _dp(st), st.size))
        return True

    def close(self):
        if self._h is not None:
            lib().nw_fields_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSolver(DgSolver):
    """Many independent meshes (any mix of n1, common dt) in one solver: each
    stage launches one persistent fused kernel per n1 segment, one after
    another.  Instances are addressed by their index in the list given to
    initialize()."""

    def __init__(self, precision="fp32", use_graph=True, device=0, tile_option=-1, groups_per_cu=0):
        super().__init__(precision=precision, pipeline="fused", use_graph=use_graph, device=device,
                         tile_option=tile_option, groups_per_cu=groups_per_cu)
        self.sizes = []

    def initialize(self, meshes, c0):
        self.finalize()
        arr = (C.c_void_p * len(meshes))(*[m.handle.value for m in meshes])
        h = C.c_void_p()
        check(lib().nw_batch_create(len(meshes), arr, float(c0), C.byref(self.opts), C.byref(h)))
        self._h = h
        self.sizes = [m.getTotalPoints() for m in meshes]
        self.npts = self.sizes[0]
        return True

    def uploadState(self, instance, hostState):
        a = np.ascontiguousarray(hostState, dtype=np.float64).reshape(-1)
        check(lib().nw_dg_upload_instance(self._handle(), int(instance), _dp(a), a.size))

    def downloadState(self, instance):
        out = np.zeros(6 * self.sizes[instance])
        check(lib().nw_dg_download_instance(self._handle(), int(instance), _dp(out), out.size))
        return out

    def diagnostics(self, instance=0):
        o = _lib.Diag()
        check(lib().nw_dg_diag_instance(self._handle(), int(instance), C.byref(o)))
        return {"energy": o.energy, "max_e": o.max_e, "max_h": o.max_h}


from .case import Case, Config  # noqa: E402
