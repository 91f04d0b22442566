"""Loader and ctypes signatures of libnekwave_amd.so (the C ABI in include/nekwave_amd.h).

The shared library is built in-tree by nekwave_amd.build; there is no fallback:
if it is missing, every entry point raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnekwave_amd.so")

NW_OK, NW_ERR_ARG, NW_ERR_HIP, NW_ERR_ALLOC, NW_ERR_STATE, NW_ERR_MESH, NW_ERR_UNSUPPORTED = range(7)
STATUS_NAMES = {0: "NW_OK", 1: "NW_ERR_ARG", 2: "NW_ERR_HIP", 3: "NW_ERR_ALLOC", 4: "NW_ERR_STATE",
                5: "NW_ERR_MESH", 6: "NW_ERR_UNSUPPORTED"}
PREC = {"fp64": 0, "fp32": 1, "fp32_k64": 2, "bf16x3": 3}
PIPE = {"plain": 0, "fused": 1}
FACE_E, FACE_PERIODIC, FACE_PEC, FACE_PMC = 0, 1, 2, 3
NW_MESH_REA_PERIODIC = 1


class NekWaveError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


class MeshInfo(C.Structure):
    _fields_ = [("n1", C.c_int), ("np", C.c_int), ("nelt", C.c_int), ("nhalo", C.c_int),
                ("npts", C.c_int64), ("tfp", C.c_int64), ("has_media", C.c_int), ("all_affine", C.c_int)]


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("pipeline", C.c_int), ("use_graph", C.c_int),
                ("device", C.c_int), ("tile_option", C.c_int), ("groups_per_cu", C.c_int)]


class Diag(C.Structure):
    _fields_ = [("energy", C.c_double), ("max_e", C.c_double), ("max_h", C.c_double)]


class DgInfo(C.Structure):
    _fields_ = [("n1", C.c_int), ("nelt", C.c_int), ("precision", C.c_int), ("pipeline", C.c_int),
                ("use_graph", C.c_int), ("epw", C.c_int), ("npts", C.c_int64), ("stride", C.c_int64),
                ("device_bytes", C.c_int64), ("kernels_per_step", C.c_int), ("instances", C.c_int),
                ("segments", C.c_int)]


_lib = None

vp, i, d, i64, sz = C.c_void_p, C.c_int, C.c_double, C.c_int64, C.c_size_t
dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)

SIGNATURES = {
    "nw_last_error": (C.c_char_p, []),
    "nw_abi_version": (i, []),
    "nw_device_info": (i, [i, ip, C.c_char_p, sz, ip, C.POINTER(sz)]),
    "nw_mesh_box": (i, [i, i, i, i, dp, i, i, i, d, C.c_uint64, C.POINTER(vp)]),
    "nw_mesh_from_corners": (i, [i, i, dp, ip, ip, ip, C.POINTER(vp)]),
    "nw_mesh_load": (i, [C.c_char_p, i, i, C.POINTER(vp)]),
    "nw_mesh_faces": (i, [vp, ip, ip, ip]),
    "nw_mesh_set_materials": (i, [vp, dp, dp]),
    "nw_mesh_self_halo_x": (i, [vp, ip]),
    "nw_mesh_get_info": (i, [vp, C.POINTER(MeshInfo)]),
    "nw_mesh_gll": (i, [vp, dp, dp, dp, dp]),
    "nw_mesh_coords": (i, [vp, dp, dp, dp]),
    "nw_mesh_corners": (i, [vp, dp]),
    "nw_mesh_jacobian": (i, [vp, dp]),
    "nw_mesh_min_node_distance": (i, [vp, dp]),
    "nw_mesh_auto_dt": (i, [vp, d, d, dp]),
    "nw_mesh_destroy": (None, [vp]),
    "nw_opts_default": (None, [C.POINTER(Opts)]),
    "nw_dg_create": (i, [vp, d, C.POINTER(Opts), C.POINTER(vp)]),
    "nw_dg_upload": (i, [vp, dp, sz]),
    "nw_dg_download": (i, [vp, dp, sz]),
    "nw_dg_step": (i, [vp, d, d]),
    "nw_dg_run": (i, [vp, d, i]),
    "nw_dg_time": (i, [vp, d, i, dp]),
    "nw_dg_time_repeats": (i, [vp, d, i, i, i, dp]),
    "nw_dg_sync": (i, [vp]),
    "nw_dg_diag": (i, [vp, C.POINTER(Diag)]),
    "nw_dg_rel_l2": (i, [vp, i, dp, C.c_size_t, dp]),
    "nw_dg_set_probes": (i, [vp, vp, i, dp, i]),
    "nw_dg_probe_sample": (i, [vp]),
    "nw_dg_probe_history": (i, [vp, dp, i, ip]),
    "nw_mesh_locate_probes": (i, [vp, i, dp, ip, dp]),
    "nw_dg_get_info": (i, [vp, C.POINTER(DgInfo)]),
    "nw_dg_destroy": (None, [vp]),
    "nw_batch_create": (i, [i, C.POINTER(vp), d, C.POINTER(Opts), C.POINTER(vp)]),
    "nw_dg_upload_instance": (i, [vp, i, dp, sz]),
    "nw_dg_download_instance": (i, [vp, i, dp, sz]),
    "nw_dg_diag_instance": (i, [vp, i, C.POINTER(Diag)]),
    "nw_dg_tune": (i, [vp]),
    "nw_dg_segment_info": (i, [vp, i, ip, ip, ip]),
    "nw_dg_segment_option": (i, [vp, i, ip, ip, ip]),
    "nw_dg_set_segment_option": (i, [vp, i, i, i]),
    "nw_dg_set_walk": (i, [vp, i]),
    "nw_dg_walk": (i, [vp, ip]),
    "nw_launch_volume_curl": (i, [vp] * 18 + [d, i, i, vp]),
    "nw_launch_restrict_faces": (i, [vp, vp, vp, vp, i64, i64, vp]),
    "nw_launch_compute_flux": (i, [vp] * 9 + [d, i64, i64, vp, vp, vp, vp, i, i64, vp]),
    "nw_launch_add_flux": (i, [vp, vp, vp, vp, i64, i64, vp]),
    "nw_fields_open": (i, [C.c_char_p, vp, i, i, vp]),
    "nw_fields_write": (i, [vp, i, C.c_double, vp, C.c_size_t]),
    "nw_fields_close": (None, [vp]),
    "nw_launch_diagnostics": (i, [vp, i64, vp, vp, vp, vp, i, i, vp, vp]),
    "nw_launch_diagnostics_noalloc": (i, [vp, i64, vp, vp, vp, vp, i, i, vp, vp]),
    "nw_launch_inv_mass": (i, [vp, vp, vp, i, i, i64, vp, vp, vp]),
    "nw_launch_lsrk45_update": (i, [vp, vp, vp, d, d, d, i64, i64, vp]),
    "nw_launch_pack_halo": (i, [vp, vp, vp, vp, vp, i, i64, vp]),
    "nw_launch_unpack_halo": (i, [vp, vp, vp, vp, i, i64, i64, vp]),
    "nw_rand_uniform": (i, [C.c_uint64, sz, dp]),
    "nw_rand_normal": (i, [C.c_uint64, sz, dp]),
    "nw_dev_alloc": (i, [sz, C.POINTER(vp)]),
    "nw_dev_free": (i, [vp]),
    "nw_dev_h2d": (i, [vp, vp, sz]),
    "nw_dev_d2h": (i, [vp, vp, sz]),
    "nw_dev_sync": (i, []),
    "nw_bench_copy": (i, [sz, i, dp]),
}


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("nekwave_amd: %s is not built; run `python -m nekwave_amd.build`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != NW_OK:
        raise NekWaveError(status, lib().nw_last_error().decode())
    return status


def last_error():
    return lib().nw_last_error().decode()
