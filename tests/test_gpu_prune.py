"""Dynamic pruning on the GPU (nbx_prune_pairlist) against the oracle's
restatement of the reference's prune kernel (kernel_ref_prune.cpp:39-123,
pinned in test_oracle_prune.py): the pruned list is identical entry for
entry, forces and energies computed with it match the FP64 oracle and the
unpruned list, and every prune starts from the list as searched."""
import numpy as np
import pytest

from gmxapi_amd import nbx, systems as S
from oracle import oracle as O
from tests.helpers import oracle_for, rms

pytestmark = pytest.mark.gpu

R_INNER = 1.0


@pytest.fixture(scope="module")
def case():
    s = S.water_box(400, 2.4, seed=21, rc=0.9, rlist=1.15)
    rng = np.random.default_rng(4)
    x1 = s.x + rng.normal(scale=0.01, size=s.x.shape)  # moved, not wrapped
    return s, x1


def _gpu_pruned(s, x1, r):
    ctx = nbx.context_for(s)
    ctx.convert_coordinates(x1)
    ctx.prune_pairlist(r)
    return ctx


def _entries(ie, cj, ex):
    # (ci, shift, cj, exclusion mask bits 4 i + j)
    return [(int(a), int(sh), int(cj[k]), int(ex[k]) & 0xFFFF) for a, sh, b, e in ie for k in range(b, e)]


def test_pruned_list_matches_the_oracle(case):
    # The reference's plain-C search and prune compare filler atoms too (all
    # fillers sit at one far-away point, so two partly filled clusters look
    # in range under the central shift); the search here skips fillers, so
    # its list lacks those entries (same atom pair set).  The prune itself
    # follows the reference, fillers included: the GPU's pruned entries are
    # the oracle's pruned entries among the GPU's own list as searched.
    print("[prune] 1200-atom water, rlist 1.15 -> 1.0", flush=True)
    s, x1 = case
    full = nbx.context_for(s)
    fie, fje = full.pairlist()
    outer_seq = _entries(fie, fje[:, 0], fje[:, 1])
    outer = set(outer_seq)
    ctx = _gpu_pruned(s, x1, R_INNER)
    ie, je = ctx.pairlist()
    assert np.all(ie[:, 3] > ie[:, 2])  # emptied i-entries are dropped
    got = _entries(ie, je[:, 0], je[:, 1])
    assert len(set(got)) == len(got) and 0 < len(got) < len(outer)
    it = iter(outer_seq)
    assert all(any(e == b for b in it) for e in got)  # the searched order is kept
    o = oracle_for(s, "f32")
    o.prune(R_INNER, x=x1)
    ci, cj, ex = o.pairlist()
    want = {e for e in _entries(ci, cj, ex) if e in outer}
    assert set(got) == want
    np.testing.assert_array_equal(ctx.list_pairs(R_INNER), o.list_pairs(R_INNER))
    np.testing.assert_array_equal(ctx.list_pairs(R_INNER), O.brute_pairs(x1, s.box, R_INNER, "f32"))


def test_forces_with_the_pruned_list(case):
    s, x1 = case
    ctx = _gpu_pruned(s, x1, R_INNER)
    f = ctx.compute(x1, energy=True)
    e = ctx.energies()
    full = nbx.context_for(s)
    f_full = full.compute(x1, energy=True)
    e_full = full.energies()
    r64 = oracle_for(s, "f64").compute(x1)
    err = rms(f - r64["f"]) / rms(r64["f"])
    print(f"  force rms err vs f64 {err:.2e}; vs unpruned {rms(f - f_full) / rms(f_full):.2e}; "
          f"E {e} unpruned {e_full}", flush=True)
    assert err <= 5e-4
    assert rms(f - f_full) <= 1e-6 * rms(f_full)
    for a, b in zip(e, e_full):
        assert abs(a - b) <= 1e-6 * abs(b)


def test_prune_starts_from_the_searched_list(case):
    s, x1 = case
    ctx = _gpu_pruned(s, x1, 0.95)
    ctx.prune_pairlist(R_INNER)
    direct = _gpu_pruned(s, x1, R_INNER)
    for a, b in zip(ctx.pairlist(), direct.pairlist()):
        np.testing.assert_array_equal(a, b)
    # a new search (here on the original coordinates) replaces the outer list
    ctx.put_on_grid(s.box, s.x)
    ctx.construct_pairlist()
    for a, b in zip(ctx.pairlist(), nbx.context_for(s).pairlist()):
        np.testing.assert_array_equal(a, b)


def test_prune_radius_errors(case):
    s, _ = case
    ctx = nbx.context_for(s)
    for r in (0.85, 1.2):
        with pytest.raises(nbx.NbxInputError):
            ctx.prune_pairlist(r)
