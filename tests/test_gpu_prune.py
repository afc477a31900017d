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


# -- pruning inside the kernel's work layout (nbx_prune_rolling) ----------


def _layout_entries(wg, nc):
    """Per chunk: (ci, shift, [(cj, mask) of its non-padding work entries in
    order]), with the tile and masked-tile counts."""
    out = []
    for ci, flags, ws, tiles in wg["chunks"]:
        ents = [(int(wg["work"][k, 0]), int(wg["work"][k, 1]) & 0xFFFF) for k in range(ws, ws + 16 * tiles)]
        real = [e for e in ents if e[0] != nc]
        # padding only at the end
        assert all(e[0] == nc for e in ents[len(real):])
        out.append((int(ci), int(flags) & 0xFF, real, int(tiles), int(flags) >> 9))
    return out


def test_rolling_prune_matches_the_oracle(case):
    print("[rolling prune] 1200-atom water, rlist 1.15 -> 1.0, one part", flush=True)
    s, x1 = case
    nc = nbx.context_for(s).info()["num_clusters"]
    outer_ctx = nbx.context_for(s)
    outer = _layout_entries(outer_ctx.work_groups(), nc)
    ctx = nbx.context_for(s)
    ctx.convert_coordinates(x1)
    ctx.prune_rolling(R_INNER)
    ctx.synchronize()
    pruned = _layout_entries(ctx.work_groups(), nc)
    o = oracle_for(s, "f32")
    o.prune(R_INNER, x=x1)
    ci, cj, ex = o.pairlist()
    want = set(_entries(ci, cj, ex))
    assert len(pruned) == len(outer)
    n_out = n_in = 0
    for (c, sh, ents, tiles, special), (c0, sh0, ents0, _t0, _s0) in zip(pruned, outer):
        assert (c, sh) == (c0, sh0)
        # the kept entries: the oracle's pruned entries among the layout's,
        # in the layout's order, masked ones first
        keep0 = [e for e in ents0 if (c, sh, e[0], e[1]) in want]
        assert ents == keep0
        assert tiles == max(1, (len(ents) + 15) // 16)
        nmask = sum(1 for e in ents if e[1] != 0xFFFF)
        assert special == (nmask + 15) // 16
        assert all(e[1] != 0xFFFF for e in ents[:nmask])
        n_out += len(ents0)
        n_in += len(ents)
    print(f"  {n_out} -> {n_in} entries", flush=True)
    assert 0 < n_in < n_out


def test_rolling_prune_forces(case):
    s, x1 = case
    ctx = nbx.context_for(s)
    ctx.convert_coordinates(x1)
    ctx.prune_rolling(R_INNER)
    f = ctx.compute(x1, energy=True, virial=True)
    e = ctx.energies()
    full = nbx.context_for(s)
    f_full = full.compute(x1, energy=True, virial=True)
    e_full = full.energies()
    r64 = oracle_for(s, "f64").compute(x1)
    err = rms(f - r64["f"]) / rms(r64["f"])
    print(f"  force rms err vs f64 {err:.2e}; vs unpruned {rms(f - f_full) / rms(f_full):.2e}", flush=True)
    assert err <= 5e-4
    assert rms(f - f_full) <= 1e-6 * rms(f_full)
    for a, b in zip(e, e_full):
        assert abs(a - b) <= 1e-6 * abs(b)
    np.testing.assert_allclose(ctx.virial(), full.virial(), rtol=0, atol=1e-6 * np.abs(full.virial()).max())


def test_rolling_parts_compose(case):
    # pruning the parts one after another gives the one-part prune; a part
    # left alone keeps the layout as built (or its last prune)
    s, x1 = case
    one = nbx.context_for(s)
    one.convert_coordinates(x1)
    one.prune_rolling(R_INNER)
    parts = nbx.context_for(s)
    parts.convert_coordinates(x1)
    parts.prune_rolling(R_INNER, 0, 3)
    parts.synchronize()
    half = parts.work_groups()
    outer = nbx.context_for(s).work_groups()
    wg1 = one.work_groups()
    changed = np.any(half["chunks"] != outer["chunks"], axis=1)
    assert changed.any() and not changed.all()
    parts.prune_rolling(R_INNER, 1, 3)
    parts.prune_rolling(R_INNER, 2, 3)
    parts.synchronize()
    ws = parts.work_groups()
    np.testing.assert_array_equal(ws["chunks"], wg1["chunks"])
    nc = parts.info()["num_clusters"]
    assert _layout_entries(ws, nc) == _layout_entries(wg1, nc)
    # pruning again from the layout as built, at a larger radius, keeps more
    parts.prune_rolling(1.1)
    parts.synchronize()
    assert parts.work_groups()["chunks"][:, 3].sum() >= wg1["chunks"][:, 3].sum()
    # a new search rebuilds the layout and the outer copy
    parts.put_on_grid(s.box, s.x)
    parts.construct_pairlist()
    np.testing.assert_array_equal(parts.work_groups()["chunks"], outer["chunks"])


def test_rolling_prune_errors(case):
    s, _ = case
    ctx = nbx.context_for(s)
    for args in ((0.85, 0, 1), (1.2, 0, 1), (1.0, 1, 1), (1.0, -1, 2), (1.0, 0, 0)):
        with pytest.raises(nbx.NbxInputError):
            ctx.prune_rolling(*args)
