"""Build the nekwave_amd HIP library for gfx950 (in-tree, no JIT cache).

    python -m nekwave_amd.build        # or nekwave_amd.build.build()

Every translation unit is compiled with hipcc --offload-arch=gfx950 into
_build/ and linked into nekwave_amd/libnekwave_amd.so.  Packed FP32 VALU ops
are disabled for all device code.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build", "nekwave_amd")
LIB = os.path.join(HERE, "libnekwave_amd.so")
ARCH = os.environ.get("NW_OFFLOAD_ARCH", "gfx950")

DEVICE_SOURCES = ["nw_fused.hip", "nw_plain.hip"]
HOST_SOURCES = ["nw_mesh.cpp", "nw_meshio.cpp", "nw_fields.cpp", "nw_solver.cpp", "nw_api.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-Wall", "-Wno-unused-result",
          "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-ffp-contract=on"]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def _compile(src, obj, device):
    cmd = [hipcc()] + COMMON + (DEVICE if device else ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]) + \
        ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    units = [(s, True) for s in DEVICE_SOURCES] + [(s, False) for s in HOST_SOURCES]
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, os.path.join(BUILD, os.path.splitext(s)[0] + ".o"), d) for s, d in units]
        for f in futs:
            objs.append(f.result())
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_cpp_kats()
    if verbose:
        print("built", LIB)
    return LIB


KATS_SRC = os.path.join(ROOT, "tests", "cpp", "reference_kats.cpp")
KATS_BIN = os.path.join(HERE, "nw_reference_kats")


def build_cpp_kats(out=KATS_BIN):
    """The reference's C++ test cases against the C++ host header, linked to the library."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), KATS_SRC,
           "-L" + HERE, "-lnekwave_amd", "-Wl,-rpath,$ORIGIN", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("C++ host interface build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
