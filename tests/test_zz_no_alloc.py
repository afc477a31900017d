"""nbx_list_pairs_noalloc (the allocation-free variant of nbx_list_pairs)
against nbx_list_pairs, through the C ABI: the same pair set bit for bit, in
the list's order and the same on every call, on two calls in a row (the
second with a smaller system, whose list is shorter), and its input checks.
The keys are written to a device array the test allocates; the library
allocates nothing for the call."""
import ctypes as C

import numpy as np
import pytest

from gmxapi_amd import nbx, systems as S

pytestmark = pytest.mark.gpu


def _noalloc(ctx, r, out_ptr, cap):
    lib = nbx.library()
    fn = lib.nbx_list_pairs_noalloc
    fn.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_int64]
    fn.restype = C.c_int64
    return int(fn(ctx.handle, float(r), out_ptr, int(cap)))


def _pairs_noalloc(ctx, r):
    import torch

    n = _noalloc(ctx, r, None, 0)
    assert n >= 0, nbx.library().nbx_last_error()
    out = torch.full((max(n, 1),), -1, dtype=torch.int64, device="cuda")
    m = _noalloc(ctx, r, out.data_ptr(), n)
    assert m == n
    torch.cuda.synchronize()
    return out[:n].cpu().numpy()


@pytest.mark.parametrize("first,second", [
    (("water500", dict(n=500, L=2.6, rc=0.9, rlist=1.0)), ("water216", dict(n=216, L=None, rc=1.0, rlist=1.0))),
])
def test_noalloc_list_pairs_match_list_pairs(first, second):
    for name, p in (first, second):
        s = (S.water_box(p["n"], p["L"], seed=9, rc=p["rc"], rlist=p["rlist"]) if p["L"]
             else S.small_water_box(p["n"], seed=11))
        print(f"[noalloc] {name}: {s.natoms} atoms", flush=True)
        ctx = nbx.context_for(s)
        for r in (s.rlist, s.rcoulomb * 0.8):
            ref = ctx.list_pairs(r)
            got = _pairs_noalloc(ctx, r)
            again = _pairs_noalloc(ctx, r)
            print(f"  r {r}: {len(got)} pairs (nbx_list_pairs {len(ref)})", flush=True)
            np.testing.assert_array_equal(got, again)  # list order, every call
            np.testing.assert_array_equal(np.sort(got), ref)
        # a short output array receives the first cap keys of the same order
        cap = len(got) // 3
        import torch

        out = torch.full((len(got),), -1, dtype=torch.int64, device="cuda")
        assert _noalloc(ctx, r, out.data_ptr(), cap) == len(got)
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        np.testing.assert_array_equal(o[:cap], got[:cap])
        assert np.all(o[cap:] == -1)


def test_noalloc_input_errors():
    s = S.small_water_box(216, seed=11)
    ctx = nbx.context_for(s)
    host = np.zeros(16, np.int64)
    assert _noalloc(ctx, 1.0, host.ctypes.data, 16) == -nbx.NBX_ERROR_INPUT
    assert _noalloc(ctx, 1.0, None, 16) == -nbx.NBX_ERROR_INPUT
    assert _noalloc(ctx, -1.0, None, 0) == -nbx.NBX_ERROR_INPUT
    fresh = nbx.Context()
    assert _noalloc(fresh, 1.0, None, 0) == -nbx.NBX_ERROR_STATE
