"""Pins the oracle's grid and pair search.

The set of atom pairs the cluster list holds within rlist must equal the
brute-force O(N^2) set computed with the same float rounding and the same
orientation rule (shifts at or below the central index only,
pairlist.cpp:3339-3342).  Grid and list invariants follow grid.cpp and
pairlist.cpp.
"""
import numpy as np
import pytest

from gmxapi_amd import systems as S
from oracle import oracle as O
from tests.helpers import oracle_for


def _systems():
    yield S.small_water_box(216, seed=11)
    yield S.water_box(300, 2.1, seed=5)
    yield S.water_box(500, 2.6, seed=9, rc=0.9, rlist=1.0)
    yield S.hetero_chains(natoms=1200, box_length=2.3, ntypes=5, seed=3, rc=0.9, rlist=1.0)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_list_pair_set_equals_brute_force(prec):
    for s in _systems():
        o = oracle_for(s, prec)
        lp = o.list_pairs(s.rlist)
        bp = O.brute_pairs(s.x, s.box, s.rlist, prec)
        assert len(np.unique(lp)) == len(lp), "a pair appears twice in the list"
        np.testing.assert_array_equal(lp, bp)


def test_brute_force_pairs_against_numpy():
    # an independent min-image count in float64: only pairs within a few ulp
    # of rlist may differ from the float32 canonical set
    s = S.small_water_box(216, seed=4)
    x = s.x
    d = x[:, None, :] - x[None, :, :]
    d -= s.box * np.round(d / s.box)
    r2 = np.sum(d * d, axis=-1)
    iu = np.triu_indices(len(x), 1)
    n64 = int(np.count_nonzero(r2[iu] < s.rlist ** 2))
    n32 = len(O.brute_pairs(x, s.box, s.rlist, "f32"))
    near = int(np.count_nonzero(np.abs(r2[iu] - s.rlist ** 2) < 1e-5))
    assert abs(n64 - n32) <= near


def test_grid_invariants():
    s = S.water_box(500, 2.6, seed=9)
    o = oracle_for(s, "f32")
    g = o.grid()
    ai = g["atomIndices"]
    real = ai[ai >= 0]
    np.testing.assert_array_equal(np.sort(real), np.arange(s.natoms))
    np.testing.assert_array_equal(ai[g["cells"]], np.arange(s.natoms))
    # Grid::setDimensions: t_len = cbrt(4 / density), ncx = int(L / t_len)
    dens = np.float32(s.natoms) / np.float32(np.float32(s.box[0]) * np.float32(np.float32(s.box[1]) * np.float32(s.box[2])))
    tlen = np.cbrt(np.float32(4) / dens)
    assert g["ncx"] == int(np.float32(s.box[0]) / tlen)
    cxy = g["cxy_ind"]
    x32 = s.x.astype(np.float32)
    for col in range(g["ncx"] * g["ncy"]):
        a = ai[cxy[col] * 4:cxy[col + 1] * 4]
        a = a[a >= 0]
        keys = list(zip(x32[a, 2], a))
        assert keys == sorted(keys), "column not sorted on (z, index)"
        assert cxy[col + 1] - cxy[col] == (len(a) + 3) // 4


def test_list_invariants():
    s = S.water_box(300, 2.1, seed=5)
    o = oracle_for(s, "f32")
    ci, cj, ex = o.pairlist()
    assert np.all(ci[:, 1] <= 22)
    g = o.grid()
    cells = g["cells"]
    for e in range(len(ci)):
        c, sh, a, b = ci[e]
        assert b > a
        js = cj[a:b]
        if sh == 22:
            assert np.all(js >= c)
            assert js[0] == c  # the diagonal cluster is the first entry
            # diagonal mask NBNXN_INTERACTION_MASK_DIAG = 0x08ce: only j > i
            assert int(ex[a]) & 0xFFFF & ~0x08CE == 0
        # masked entries come first (sort_cj_excl)
        masked = ex[a:b] != 0xFFFFFFFF
        if masked.any():
            first_unmasked = np.argmin(masked) if not masked.all() else len(masked)
            assert not masked[first_unmasked:].any()
    # every cleared bit of a non-diagonal entry is a topology exclusion
    ai = g["atomIndices"]
    excl = set()
    for a in range(s.natoms):
        for b in s.excl_idx[s.excl_off[a]:s.excl_off[a + 1]]:
            excl.add((a, int(b)))
    for e in range(len(ci)):
        c, sh, a0, b0 = ci[e]
        for k in range(a0, b0):
            for i in range(4):
                for j in range(4):
                    bit = (int(ex[k]) >> (4 * i + j)) & 1
                    if bit:
                        continue
                    if sh == 22 and cj[k] == c and j <= i:
                        continue
                    ia, ja = ai[c * 4 + i], ai[cj[k] * 4 + j]
                    assert (ia, ja) in excl


def test_ensemble_generator():
    # the batched-ensemble workload: counts log-uniform in [1000, 16000],
    # boxes at 33.4 molecules/nm^3, all larger than twice the cut-off
    from gmxapi_amd import systems as S
    reps = S.ensemble(64)
    n = np.array([s.natoms // 3 for s in reps])
    assert len(reps) == 64 and n.min() >= 1000 and n.max() <= 16000
    for s in reps[:4]:
        assert abs(s.natoms / 3 / s.box[0] ** 3 - 33.4) < 1e-6 * 33.4
        assert s.box[0] > 2 * s.rlist


def test_pair_at_exactly_rlist_is_outside():
    # the float32 test d^2 < rlist^2 with every product and sum rounded puts
    # this pair exactly on rlist^2, outside the list; fusing either square
    # into the sum (a multiply-add) would round it to just below and list it
    from fractions import Fraction as Fr

    s = S.boundary_pair_system()
    dx, dy = np.float32(S.BOUNDARY_DX), np.float32(S.BOUNDARY_DY)
    assert float(np.float32(s.x[1, 0]) - np.float32(s.x[0, 0])) == float(dx)
    assert float(np.float32(s.x[1, 1]) - np.float32(s.x[0, 1])) == float(dy)
    r2 = np.float32(np.float32(s.rlist) * np.float32(s.rlist))
    assert np.float32(np.float32(dx * dx) + np.float32(dy * dy)) == r2

    def fused(u, v):  # round(u*u + fl(v*v)) to float32, exactly
        e = Fr(float(u)) ** 2 + Fr(float(np.float32(v * v)))
        f = np.float32(float(e))
        return min((f, np.nextafter(f, np.float32(2)), np.nextafter(f, np.float32(0))),
                   key=lambda c: abs(Fr(float(c)) - e))

    assert fused(dx, dy) < r2 and fused(dy, dx) < r2
    key = (0 << 32) | 1
    for prec in ("f32",):
        lp = oracle_for(s, prec).list_pairs(s.rlist)
        bp = O.brute_pairs(s.x, s.box, s.rlist, prec)
        np.testing.assert_array_equal(lp, bp)
        assert key not in set(lp.tolist())
        # just inside: listed
        assert key in set(oracle_for(s, prec).list_pairs(np.nextafter(np.float32(1.25), np.float32(2))).tolist())
