"""The nblib-compatible C++ layer (gmxapi_amd/nblib) through its own test
program, tests/cpp/test_nblib.cpp, which mirrors the reference's
api/nblib/tests/nbkernelsystem.cpp.

Host-only checks run on the CPU; the force tests run on the GPU with the
argon and SPC-methanol systems of the reference's testsystems.cpp (inputs
and reference forces from tests/golden)."""
import os
import subprocess

import pytest

from gmxapi_amd import build
from tests.helpers import load_golden

# particle masses of testsystems.cpp:53-67 (ParticleLibrary)
MASSES = {"Ow": 15.99940, "H": 1.008, "OMet": 15.999, "CMet": 15.035, "Ar": 39.94800}


def _binary():
    exe = build.nblib_test_path()
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/test_nblib is not built: run __graft_entry__.build()")
    return exe


def _run(args):
    res = subprocess.run([_binary()] + args, capture_output=True, text=True, timeout=120)
    print(res.stdout)
    print(res.stderr)
    return res


def test_host_side_validation():
    res = _run(["host"])
    assert res.returncode == 0, res.stdout
    assert "0 failure(s)" in res.stdout


def _write_system(path, name, d):
    n = len(d["x"])
    names = d["type_names"]
    with open(path, "w") as f:
        f.write(f"{name}\n{n} {len(names)}\n{d['box']!r} {d['rc']!r}\n")
        for t, nm in enumerate(names):
            f.write(f"{nm} {MASSES[nm]!r} {d['c6'][t]!r} {d['c12'][t]!r}\n")
        for a in range(n):
            x = d["x"][a]
            f.write(f"{d['types'][a]} {d['charges'][a]!r} {x[0]!r} {x[1]!r} {x[2]!r}\n")
        for a in range(n):
            ex = d["excl_idx"][d["excl_off"][a]:d["excl_off"][a + 1]]
            f.write(" ".join(str(v) for v in [len(ex)] + list(ex)) + "\n")
        for fr in d["forces"]:
            f.write(f"{fr[0]!r} {fr[1]!r} {fr[2]!r}\n")


@pytest.mark.gpu
def test_nbkernelsystem_on_device(tmp_path):
    files = []
    for name, fixture in (("Argon", "nblib_argon.json"), ("SpcMethanol", "nblib_spc_methanol.json")):
        p = tmp_path / f"{name}.txt"
        _write_system(p, name, load_golden(fixture))
        files.append(str(p))
    print("[nblib-cpp] device tests", flush=True)
    res = _run(["device"] + files)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failure(s)" in res.stdout
