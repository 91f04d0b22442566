"""ctypes view of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

FACE_E, FACE_PERIODIC, FACE_PEC, FACE_PMC, FACE_HALO = 0, 1, 2, 3, 4
MAXN1 = 12


class _Mesh(C.Structure):
    _fields_ = [
        ("n1", C.c_int), ("np", C.c_int), ("nelt", C.c_int), ("npts", C.c_int),
        ("tfp", C.c_int), ("nhalo", C.c_int),
        ("z", C.c_double * MAXN1), ("w", C.c_double * MAXN1),
        ("D", C.c_double * (MAXN1 * MAXN1)), ("Dt", C.c_double * (MAXN1 * MAXN1)),
        ("w3", C.POINTER(C.c_double)), ("corners", C.POINTER(C.c_double)),
        ("f_kind", C.POINTER(C.c_int)), ("f_nbr_e", C.POINTER(C.c_int)), ("f_nbr_f", C.POINTER(C.c_int)),
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)), ("zc", C.POINTER(C.c_double)),
        ("jac", C.POINTER(C.c_double)),
        ("rx", C.POINTER(C.c_double)), ("sx", C.POINTER(C.c_double)), ("tx", C.POINTER(C.c_double)),
        ("ry", C.POINTER(C.c_double)), ("sy", C.POINTER(C.c_double)), ("ty", C.POINTER(C.c_double)),
        ("rz", C.POINTER(C.c_double)), ("sz", C.POINTER(C.c_double)), ("tz", C.POINTER(C.c_double)),
        ("vM", C.POINTER(C.c_int)), ("vP", C.POINTER(C.c_int)), ("bc", C.POINTER(C.c_int)),
        ("nx", C.POINTER(C.c_double)), ("ny", C.POINTER(C.c_double)), ("nz", C.POINTER(C.c_double)),
        ("dA", C.POINTER(C.c_double)),
        ("eps", C.POINTER(C.c_double)), ("mu", C.POINTER(C.c_double)),
        ("send_vol", C.POINTER(C.c_int)), ("ghost_elem", C.POINTER(C.c_int)),
    ]


class _Solver(C.Structure):
    _fields_ = [
        ("m", C.POINTER(_Mesh)), ("c0", C.c_double), ("stride", C.c_int),
        ("state", C.POINTER(C.c_double)), ("k", C.POINTER(C.c_double)), ("rhs", C.POINTER(C.c_double)),
        ("fEN", C.POINTER(C.c_double)), ("fHN", C.POINTER(C.c_double)), ("flux", C.POINTER(C.c_double)),
        ("sendbuf", C.POINTER(C.c_double)), ("recvbuf", C.POINTER(C.c_double)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(_LIB)
        P = C.POINTER
        d, i, u64 = C.c_double, C.c_int, C.c_uint64
        dp, ip = P(d), P(i)
        sig = {
            "or_gll": (i, [i, dp, dp, dp, dp]),
            "or_mesh_box": (P(_Mesh), [i, i, i, i, dp, i, i, i, d, u64]),
            "or_mesh_from_corners": (P(_Mesh), [i, i, dp, ip, ip, ip]),
            "or_mesh_free": (None, [P(_Mesh)]),
            "or_mesh_set_materials": (i, [P(_Mesh), dp, dp]),
            "or_mesh_self_halo_x": (i, [P(_Mesh), i]),
            "or_dxmin": (d, [P(_Mesh)]),
            "or_auto_cfl": (d, [d]),
            "or_auto_dt": (d, [P(_Mesh), d, d]),
            "or_volume_curl": (None, [P(_Mesh), dp, dp, dp, dp, dp, dp, d]),
            "or_restrict": (None, [P(_Mesh), dp, i, dp, dp]),
            "or_flux": (None, [P(_Mesh), dp, i, dp, dp, d, dp]),
            "or_add_flux": (None, [P(_Mesh), dp, dp]),
            "or_inv_mass": (None, [P(_Mesh), dp]),
            "or_lsrk_update": (None, [P(_Mesh), dp, i, dp, dp, d, d, d]),
            "or_pack_halo": (None, [P(_Mesh), dp, i, dp]),
            "or_unpack_halo": (None, [P(_Mesh), dp, dp, i]),
            "or_rk_coeffs": (None, [dp, dp]),
            "or_solver_new": (P(_Solver), [P(_Mesh), d]),
            "or_solver_free": (None, [P(_Solver)]),
            "or_solver_upload": (None, [P(_Solver), dp]),
            "or_solver_download": (None, [P(_Solver), dp]),
            "or_solver_stage": (None, [P(_Solver), i, d]),
            "or_solver_step": (None, [P(_Solver), d]),
            "or_solver_run": (d, [P(_Solver), d, i]),
            "or_energy": (d, [P(_Mesh), dp, i]),
            "or_max_e": (d, [P(_Mesh), dp, i]),
            "or_max_h": (d, [P(_Mesh), dp, i]),
            "or_lagrange_1d": (None, [i, dp, d, dp]),
            "or_locate_probes": (None, [P(_Mesh), i, dp, ip, dp]),
            "or_probe_record": (None, [P(_Mesh), i, ip, dp, dp, i, dp]),
            "or_set_threads": (None, [i]),
            "or_rng_seed": (None, [C.c_void_p, u64]),
            "or_rng_next": (u64, [C.c_void_p]),
            "or_rng_uniform": (d, [C.c_void_p]),
            "or_rng_normal": (d, [C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _arr(ptr, n, dtype):
    if not ptr or n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Rng:
    """mt19937_64 with the project transforms (uniform=(x>>11)*2^-53, Box-Muller)."""

    def __init__(self, seed):
        self._buf = C.create_string_buffer(312 * 8 + 16)
        lib().or_rng_seed(self._buf, seed)

    def next(self):
        return lib().or_rng_next(self._buf)

    def uniform(self):
        return lib().or_rng_uniform(self._buf)

    def normal(self):
        return lib().or_rng_normal(self._buf)


def gll(n1):
    z = np.zeros(n1); w = np.zeros(n1); D = np.zeros(n1 * n1); Dt = np.zeros(n1 * n1)
    if lib().or_gll(n1, dptr(z), dptr(w), dptr(D), dptr(Dt)) != 0:
        raise ValueError("bad n1")
    return z, w, D, Dt


class Mesh:
    """Oracle mesh (reference Mesh, mesh.hpp:51-218) with numpy copies of its arrays."""

    def __init__(self, ptr):
        if not ptr:
            raise ValueError("oracle mesh construction failed")
        self._p = ptr

    @classmethod
    def box(cls, n1, nel, bounds=(-1, 1, -1, 1, -1, 1), periodic=(False, False, False),
            jitter=0.0, seed=0):
        b = np.asarray(bounds, dtype=np.float64)
        nelx, nely, nelz = nel
        m = cls(lib().or_mesh_box(n1, nelx, nely, nelz, dptr(b), int(periodic[0]),
                                  int(periodic[1]), int(periodic[2]), float(jitter), seed))
        m.nel = tuple(nel)
        return m

    @classmethod
    def from_corners(cls, n1, corners, kind, nbr_e, nbr_f):
        corners = np.ascontiguousarray(corners, dtype=np.float64).reshape(-1)
        kind = np.ascontiguousarray(kind, dtype=np.int32).reshape(-1)
        nbr_e = np.ascontiguousarray(nbr_e, dtype=np.int32).reshape(-1)
        nbr_f = np.ascontiguousarray(nbr_f, dtype=np.int32).reshape(-1)
        nelt = corners.size // 24
        return cls(lib().or_mesh_from_corners(n1, nelt, dptr(corners), iptr(kind), iptr(nbr_e), iptr(nbr_f)))

    def __del__(self):
        try:
            if self._p:
                lib().or_mesh_free(self._p)
                self._p = None
        except Exception:
            pass

    @property
    def s(self):
        return self._p.contents

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        s = self._p.contents
        scalars = ("n1", "np", "nelt", "npts", "tfp", "nhalo")
        if name in scalars:
            return getattr(s, name)
        if name in ("z", "w"):
            return np.array(getattr(s, name)[: s.n1])
        if name in ("D", "Dt"):
            return np.array(getattr(s, name)[: s.n1 * s.n1])
        if name == "w3":
            return _arr(s.w3, s.np, np.float64)
        if name == "corners":
            return _arr(s.corners, s.nelt * 24, np.float64)
        if name in ("f_kind", "f_nbr_e", "f_nbr_f"):
            return _arr(getattr(s, name), s.nelt * 6, np.int32)
        if name in ("x", "y", "zc", "jac", "rx", "sx", "tx", "ry", "sy", "ty", "rz", "sz", "tz"):
            return _arr(getattr(s, name), s.npts, np.float64)
        if name in ("vM", "vP", "bc"):
            return _arr(getattr(s, name), s.tfp, np.int32)
        if name in ("nx", "ny", "nz", "dA"):
            return _arr(getattr(s, name), s.tfp, np.float64)
        if name in ("eps", "mu"):
            return _arr(getattr(s, name), s.nelt, np.float64)
        if name in ("send_vol", "ghost_elem"):
            return _arr(getattr(s, name), s.nhalo, np.int32)
        raise AttributeError(name)

    def set_materials(self, eps=None, mu=None):
        e = None if eps is None else np.ascontiguousarray(eps, dtype=np.float64)
        u = None if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        lib().or_mesh_set_materials(self._p, dptr(e) if e is not None else None,
                                    dptr(u) if u is not None else None)
        self._keep = (e, u)

    def self_halo_x(self):
        return lib().or_mesh_self_halo_x(self._p, self.nel[0])

    def dxmin(self):
        return lib().or_dxmin(self._p)

    def auto_dt(self, speed=1.0, safety=0.4):
        return lib().or_auto_dt(self._p, speed, safety)

    def energy(self, state, stride=None):
        state = np.ascontiguousarray(state, dtype=np.float64)
        return lib().or_energy(self._p, dptr(state), stride or self.npts)

    def locate_probes(self, xyz):
        """ProbeManager::init's placement (probe.cpp:64-91): element, (r, s, t)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        elem = np.zeros(n, dtype=np.int32)
        rst = np.zeros((n, 3))
        lib().or_locate_probes(self._p, n, dptr(xyz), iptr(elem), dptr(rst))
        return elem, rst

    def probe_record(self, xyz, state, stride=None):
        """ProbeManager::record (probe.cpp:192-216) of `state`: (probes, 6)."""
        elem, rst = self.locate_probes(xyz)
        state = np.ascontiguousarray(state, dtype=np.float64)
        out = np.zeros((elem.size, 6))
        lib().or_probe_record(self._p, elem.size, iptr(elem), dptr(rst), dptr(state), stride or self.npts, dptr(out))
        return out

    def max_e(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        return lib().or_max_e(self._p, dptr(state), self.npts)

    def max_h(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        return lib().or_max_h(self._p, dptr(state), self.npts)


def rk_coeffs():
    a = np.zeros(5); b = np.zeros(5)
    lib().or_rk_coeffs(dptr(a), dptr(b))
    return a, b


class Solver:
    """Oracle DgSolver (dg_solver.hpp:35-102): upload/download/step."""

    def __init__(self, mesh, c0=1.0):
        self.mesh = mesh
        self._p = lib().or_solver_new(mesh._p, c0)

    def __del__(self):
        try:
            if self._p:
                lib().or_solver_free(self._p)
                self._p = None
        except Exception:
            pass

    def upload(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        assert state.size == 6 * self.mesh.npts
        lib().or_solver_upload(self._p, dptr(state))

    def download(self):
        out = np.zeros(6 * self.mesh.npts)
        lib().or_solver_download(self._p, dptr(out))
        return out

    def stage(self, s, dt):
        lib().or_solver_stage(self._p, s, dt)

    def step(self, dt, nsteps=1):
        for _ in range(nsteps):
            lib().or_solver_step(self._p, dt)

    def run(self, dt, nsteps):
        return lib().or_solver_run(self._p, dt, nsteps)

    def k(self):
        return _arr(self._p.contents.k, 6 * self.mesh.npts, np.float64)


def set_threads(n):
    lib().or_set_threads(int(n))


# ---- kernel-level wrappers (numpy in, numpy out) ----
def volume_curl(mesh, u1, u2, u3, sign):
    outs = [np.zeros(mesh.npts) for _ in range(3)]
    ins = [np.ascontiguousarray(u, dtype=np.float64) for u in (u1, u2, u3)]
    lib().or_volume_curl(mesh._p, *[dptr(a) for a in ins], *[dptr(a) for a in outs], float(sign))
    return outs


def restrict(mesh, state, stride=None):
    state = np.ascontiguousarray(state, dtype=np.float64)
    fEN = np.zeros(3 * mesh.tfp); fHN = np.zeros(3 * mesh.tfp)
    lib().or_restrict(mesh._p, dptr(state), stride or mesh.npts, dptr(fEN), dptr(fHN))
    return fEN, fHN


def flux(mesh, state, fEN, fHN, c0, stride=None):
    state = np.ascontiguousarray(state, dtype=np.float64)
    out = np.zeros(6 * mesh.tfp)
    lib().or_flux(mesh._p, dptr(state), stride or mesh.npts, dptr(np.ascontiguousarray(fEN)),
                  dptr(np.ascontiguousarray(fHN)), float(c0), dptr(out))
    return out


def add_flux(mesh, fl, rhs):
    rhs = np.array(rhs, dtype=np.float64, copy=True)
    lib().or_add_flux(mesh._p, dptr(np.ascontiguousarray(fl)), dptr(rhs))
    return rhs


def inv_mass(mesh, rhs):
    rhs = np.array(rhs, dtype=np.float64, copy=True)
    lib().or_inv_mass(mesh._p, dptr(rhs))
    return rhs


def lsrk_update(mesh, state, k, rhs, a, b, dt, stride=None):
    state = np.array(state, dtype=np.float64, copy=True)
    k = np.array(k, dtype=np.float64, copy=True)
    lib().or_lsrk_update(mesh._p, dptr(state), stride or mesh.npts, dptr(k),
                         dptr(np.ascontiguousarray(rhs)), a, b, dt)
    return state, k
