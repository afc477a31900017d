"""Builds libnbx.so, the HIP library behind the C ABI in include/nbx.h, and
the nblib-compatible C++ layer over it.

The library is compiled in-tree for gfx950 with hipcc: every .hip file under
gmxapi_amd/csrc is compiled (device and host code) and linked into
gmxapi_amd/lib/libnbx.so.  Packed FP32 instructions are disabled
(-packed-fp32-ops); float atomics map to the hardware instructions
(-munsafe-fp-atomics).  gmxapi_amd/nblib/nblib.cpp is compiled with g++ into
gmxapi_amd/lib/libnblib_amd.so (linked to libnbx.so, found through
$ORIGIN), and tests/cpp/test_nblib.cpp into gmxapi_amd/lib/test_nblib.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import tempfile
import concurrent.futures

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libnbx.so"
ARCH = os.environ.get("NBX_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: install ROCm or set HIPCC")


def library_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def nblib_path() -> str:
    return os.path.join(LIBDIR, "libnblib_amd.so")


def nblib_test_path() -> str:
    return os.path.join(LIBDIR, "test_nblib")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (res.stderr or "").splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed ({res.returncode}):\n{errs}")
    if verbose and errs.strip():
        print(errs)


def build_nblib(verbose: bool = False) -> str:
    """Compile the C++ nblib layer and its test program (needs libnbx.so)."""
    root = os.path.dirname(HERE)
    inc = ["-I", os.path.join(root, "include"), "-I", os.path.join(HERE, "nblib")]
    rpath = "-Wl,-rpath,$ORIGIN"
    lib = nblib_path()
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall"] + inc
         + [os.path.join(HERE, "nblib", "nblib.cpp"), "-L", LIBDIR, "-lnbx", rpath, "-o", lib + ".tmp"], verbose)
    os.replace(lib + ".tmp", lib)
    exe = nblib_test_path()
    _run(["g++", "-O2", "-std=c++17", "-Wall"] + inc
         + [os.path.join(root, "tests", "cpp", "test_nblib.cpp"), "-L", LIBDIR, "-lnblib_amd", "-lnbx", rpath,
            "-o", exe + ".tmp"], verbose)
    os.replace(exe + ".tmp", exe)
    return lib


def build(verbose: bool = False) -> str:
    """Compile the library; returns its path.  Each source file is compiled
    on its own (in parallel, one process per file up to the CPU count), then
    the objects are linked."""
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    out = library_path()
    tmp = out + ".tmp"
    flags = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
             "-I", os.path.join(os.path.dirname(HERE), "include")]

    def run(cmd):
        if verbose:
            print(" ".join(cmd))
        res = subprocess.run(cmd, capture_output=True, text=True)
        errs = "\n".join(l for l in (res.stderr or "").splitlines() if "packed-fp32-ops" not in l)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed ({res.returncode}):\n{errs}")
        if verbose and errs.strip():
            print(errs)

    with tempfile.TemporaryDirectory(prefix="nbx_build_") as objdir:
        objs = [os.path.join(objdir, os.path.basename(src) + ".o") for src in srcs]
        workers = max(1, min(len(srcs), os.cpu_count() or 1, 16))
        with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
            list(pool.map(lambda so: run([hipcc()] + flags + ["-c", so[0], "-o", so[1]]), zip(srcs, objs)))
        run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    build_nblib(verbose)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
