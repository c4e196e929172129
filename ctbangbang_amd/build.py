"""In-tree build of libctbb_hip.so and the ctbangbang driver for gfx950.

Host C++ files are compiled with g++, HIP files with hipcc
(--offload-arch=gfx950, packed-FP32 instructions disabled), and everything is
linked by hipcc into ctbangbang_amd/libctbb_hip.so.  The filter tables of the
supported scanners are generated from their closed form into
ctbangbang_amd/resources/filters.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libctbb_hip.so")
EXE = os.path.join(PKG, "ctbangbang")

HOST_SRCS = ["setup.cpp", "filters.cpp", "geometry.cpp", "synth_host.cpp", "dicom.cpp", "cpu_path.cpp"]
HIP_SRCS = ["bp.hip", "rebin.hip", "filter.hip", "adaptive.hip", "finalize.hip", "synth.hip", "plan.hip",
            "ctx.hip", "compat.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-pthread", "-Wall", "-Wno-unused-function", f"-I{INC}", f"-I{CSRC}"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-Wno-unused-result", "-Wno-unused-value",
             f"-I{INC}", f"-I{CSRC}"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


def _obj(src):
    return os.path.join(BUILD, os.path.splitext(src)[0] + ".o")


def _compile(src):
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        return _run([HIPCC] + HIP_FLAGS + ["-c", path, "-o", _obj(src)])
    return _run([CXX] + HOST_FLAGS + ["-c", path, "-o", _obj(src)])


def write_filters(verbose=False):
    """Generate resources/filters from the ramp closed form (see filters.cpp)."""
    import ctypes
    import math
    lib = ctypes.CDLL(LIB)
    lib.ctbb_ramp_filter.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.POINTER(ctypes.c_float)]
    out = os.path.join(PKG, "resources", "filters")
    os.makedirs(out, exist_ok=True)
    # (n_channels, r_f, fan angle increment): Definition AS spacing as the
    # reference's tables use it (0.067864 deg, src/setup.cu:332); the
    # 672-channel tables use the Sensation 64 spacing (0.07758621 deg,
    # src/setup.cu:356); 920 channels use the Force scanner file values.
    scanners = [(736, 595.0, 0.067864 * math.pi / 180.0),
                (672, 570.0, 0.07758621 * math.pi / 180.0),
                (920, 595.0, 0.00094729)]
    kernels = {"sharp": (1.0, 1.0), "ramp": (1.0, 1.0), "medium": (0.65, 0.5), "smooth": (0.4, 0.5),
               "exp": (0.4, 0.5)}
    for nch, rf, dbeta in scanners:
        n = 2 * nch
        ds = rf * math.sin(dbeta / 2.0)
        for name, (c, a) in kernels.items():
            taps = (ctypes.c_float * (2 * n))()
            if lib.ctbb_ramp_filter(n, ds, c, a, taps) != 0:
                raise RuntimeError("ctbb_ramp_filter failed")
            with open(os.path.join(out, f"f_{nch}_{name}.txt"), "wb") as f:
                f.write(bytes(taps))
    if verbose:
        print("filters written to", out)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = HOST_SRCS + HIP_SRCS
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for src, log in zip(srcs, ex.map(_compile, srcs)):
            if verbose and log.strip():
                print(f"[{src}]", log.strip()[:2000])
    objs = [_obj(s) for s in srcs]
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", LIB + ".tmp"] + objs)
    os.replace(LIB + ".tmp", LIB)
    _run([HIPCC] + HIP_FLAGS + [os.path.join(CSRC, "main.cpp"), "-x", "none", "-L" + PKG, "-lctbb_hip",
                                "-Wl,-rpath,$ORIGIN", "-o", EXE])
    write_filters(verbose)
    if verbose:
        print("built", LIB)
    return LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    for p in (LIB, EXE):
        if os.path.exists(p):
            os.remove(p)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
