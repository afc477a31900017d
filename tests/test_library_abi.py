"""CPU checks of the C ABI library: it loads and exports every symbol the
public header declares (no compute calls without a GPU)."""
import ctypes
import os

from gmxapi_amd import build, nbx


def test_library_is_built_in_tree():
    assert os.path.exists(build.library_path()), "run __graft_entry__.build() first"


def test_library_exports_every_header_symbol():
    lib = nbx.library()
    syms = nbx.header_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(lib, name), name


def test_default_options_match_nblib():
    lib = nbx.library()
    o = nbx._Options()
    lib.nbx_default_options(ctypes.byref(o))
    # NBKernelOptions defaults (api/nblib/kerneloptions.h:86-112)
    assert o.coulomb_type == nbx.COULOMB["pme"]
    assert o.pairlist_cutoff == 1.0
    assert o.ewald_rtol == 1e-5
    assert o.ewald_potential_shift == 0


def test_null_arguments_are_input_errors():
    lib = nbx.library()
    assert lib.nbx_create(None, None) == nbx.NBX_ERROR_INPUT
    assert b"NULL" in lib.nbx_last_error()
    assert lib.nbx_compute(None, None, None, 1, 0) == nbx.NBX_ERROR_INPUT
