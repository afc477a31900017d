"""Builds libnbx.so, the HIP library behind the C ABI in include/nbx.h.

The library is compiled in-tree for gfx950 with hipcc: every .hip file under
gmxapi_amd/csrc is compiled (device and host code) and linked into
gmxapi_amd/lib/libnbx.so.  Packed FP32 instructions are disabled
(-packed-fp32-ops); float atomics map to the hardware instructions
(-munsafe-fp-atomics).
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess

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


def build(verbose: bool = False) -> str:
    """Compile the library; returns its path."""
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    out = library_path()
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
           "-I", os.path.join(os.path.dirname(HERE), "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (res.stderr or "").splitlines()
                     if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{errs}")
    if verbose and errs.strip():
        print(errs)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
