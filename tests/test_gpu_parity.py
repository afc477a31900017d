"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars (the path's contract):
  * grid: atom order bit-identical to the oracle's restatement of the
    reference grid;
  * pair search: the set of atom pairs within rlist bit-exact with the
    oracle's list and with brute-force enumeration;
  * forces within 5e-4 of the RMS force, energies within 1e-6 relative of
    the FP64 oracle (the reference kernel in double precision), and close
    to the FP32 oracle (the reference plain-C kernel in single precision).
"""
import numpy as np
import pytest

from gmxapi_amd import nbx, systems as S
from oracle import oracle as O
from tests.helpers import fp32_energy_noise, load_golden, oracle_for, rms

pytestmark = pytest.mark.gpu

FORCE_TOL_RMS = 5e-4
ENERGY_TOL_REL = 1e-6


def _systems():
    return [
        ("water216", S.small_water_box(216, seed=11)),
        ("water500_rc0.9", S.water_box(500, 2.6, seed=9, rc=0.9, rlist=1.0)),
        ("chains1200", S.hetero_chains(1200, 2.3, ntypes=5, seed=3, rc=0.9, rlist=1.0)),
    ]


@pytest.fixture(scope="module")
def water3000():
    return S.config(0)


@pytest.mark.parametrize("name,s", _systems())
def test_grid_is_bit_identical(name, s):
    print(f"[grid] {name} natoms={s.natoms}", flush=True)
    ctx = nbx.context_for(s)
    o = oracle_for(s, "f32")
    g = o.grid()
    ai, cells = ctx.grid()
    info = ctx.info()
    assert info["num_columns_x"] == g["ncx"] and info["num_columns_y"] == g["ncy"]
    np.testing.assert_array_equal(ai, g["atomIndices"])
    np.testing.assert_array_equal(cells, g["cells"])


@pytest.mark.parametrize("name,s", _systems())
def test_pair_set_is_bit_exact(name, s):
    print(f"[pairs] {name}", flush=True)
    ctx = nbx.context_for(s)
    gpu = ctx.list_pairs(s.rlist)
    assert len(np.unique(gpu)) == len(gpu)
    np.testing.assert_array_equal(gpu, oracle_for(s, "f32").list_pairs(s.rlist))
    np.testing.assert_array_equal(gpu, O.brute_pairs(s.x, s.box, s.rlist, "f32"))


def test_pair_at_exactly_rlist():
    # a pair whose float32 squared distance rounds to exactly rlist^2 (outside)
    # and to just below it if a product is fused into the sum: the search, the
    # pair expansion and the prune must all round as the reference's test does
    s = S.boundary_pair_system()
    print(f"[pairs] boundary pair, {s.natoms} atoms", flush=True)
    ctx = nbx.context_for(s)
    ref = oracle_for(s, "f32").list_pairs(s.rlist)
    gpu = ctx.list_pairs(s.rlist)
    assert (0 << 32 | 1) not in set(gpu.tolist())
    np.testing.assert_array_equal(gpu, ref)
    r_in = float(np.nextafter(np.float32(1.25), np.float32(2)))
    assert (0 << 32 | 1) in set(ctx.list_pairs(r_in).tolist())
    ctx.prune_pairlist(1.25)
    np.testing.assert_array_equal(ctx.list_pairs(s.rlist), ref)


def test_rebuilt_lists_match_the_oracle():
    # a list rebuilt on a context fills each cluster's region sized from the
    # previous search (no count pass) and falls back to counting when a
    # cluster outgrows it: dilute box first (small regions), then the dense
    # box (regions too small: counted), then moved atoms (regions fit)
    import dataclasses

    _, s = _systems()[1]
    ctx = nbx.context_for(s)
    dilute = dataclasses.replace(s, box=s.box * 1.6, x=s.x * 1.6)
    rng = np.random.default_rng(5)
    moved = dataclasses.replace(s, x=np.mod(s.x + 0.02 * rng.standard_normal(s.x.shape), s.box))
    for name, t in (("dilute", dilute), ("dense", s), ("moved", moved)):
        print(f"[rebuild] {name}", flush=True)
        ctx.put_on_grid(t.box, t.x)
        ctx.construct_pairlist()
        np.testing.assert_array_equal(ctx.list_pairs(t.rlist), oracle_for(t, "f32").list_pairs(t.rlist))
        f = ctx.compute(t.x)
        r64 = oracle_for(t, "f64").compute(t.x)
        assert rms(f - r64["f"]) / rms(r64["f"]) <= FORCE_TOL_RMS


def test_group_union_paths_agree(water3000, monkeypatch):
    # the group pass builds each group's union from a bitmap over its
    # j-cluster range, or by sorting the keys when the range is too wide;
    # both must give the same work layout (union order, slots, chunks)
    print("[groups] bitmap and sort unions, 9000 atoms", flush=True)
    s = water3000
    layouts = []
    for sort in ("0", "1"):
        monkeypatch.setenv("NBX_UNION_SORT", sort)
        ctx = nbx.context_for(s)
        layouts.append((ctx.work_groups(), ctx.compute(s.x)))
    (a, fa), (b, fb) = layouts
    for key in ("group_size", "num_groups", "num_chunks", "window_slots"):
        assert a[key] == b[key], key
    for key in ("groups", "chunks", "union", "work"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    np.testing.assert_allclose(fa, fb, rtol=1e-5, atol=1e-3)


def test_pair_set_config0(water3000):
    print("[pairs] config0 9000 atoms", flush=True)
    s = water3000
    ctx = nbx.context_for(s)
    np.testing.assert_array_equal(ctx.list_pairs(s.rlist), oracle_for(s, "f32").list_pairs(s.rlist))


@pytest.mark.parametrize("name,s", _systems())
def test_list_contract(name, s):
    ctx = nbx.context_for(s)
    ie, je = ctx.pairlist()
    ai, cells = ctx.grid()
    assert np.all(ie[:, 1] <= 22)
    excl = set()
    for a in range(s.natoms):
        for b in s.excl_idx[s.excl_off[a]:s.excl_off[a + 1]]:
            excl.add((a, int(b)))
    central = ie[ie[:, 1] == 22]
    # every non-empty cluster has exactly one central entry holding its diagonal
    assert len(np.unique(central[:, 0])) == len(central) == info_clusters(ctx)
    for c, sh, a0, b0 in ie:
        js = je[a0:b0, 0].astype(np.int64)
        if sh == 22:
            assert np.all(js >= c) and c in js
        for k in range(a0, b0):
            m = int(je[k, 1])
            cj = int(je[k, 0])
            for i in range(4):
                for j in range(4):
                    if (m >> (4 * i + j)) & 1:
                        continue
                    if sh == 22 and cj == c and j <= i:
                        continue
                    assert (int(ai[c * 4 + i]), int(ai[cj * 4 + j])) in excl


def info_clusters(ctx):
    return ctx.info()["num_clusters"]


_NOISE = {}


def _fp32_noise(s, coulomb="ewald", ewald_exact=False, copies=6, vdw=("potential-shift", 0.0), lj_pme=None,
                rvdw=None):
    """The reference FP32 kernel's RMS energy rounding noise on this box
    (tests.helpers.fp32_energy_noise: the FP32 oracle's deviation from FP64
    over rigid translations).  On boxes of a few hundred molecules that noise
    is above 1e-6 of the total (float coordinate differences and shift
    vectors carry ~1e-7 relative rounding that r^-12 and the Coulomb sum
    amplify)."""
    key = (id(s), coulomb, ewald_exact, vdw, lj_pme, rvdw)
    if key not in _NOISE:
        _NOISE[key] = fp32_energy_noise(s, copies, coulomb=coulomb, ewald_exact=ewald_exact, vdw=vdw,
                                        lj_pme=lj_pme, rvdw=rvdw)
    return _NOISE[key]


def _energy_bar(s, key, scale, coulomb="ewald", ewald_exact=False, vdw=("potential-shift", 0.0), lj_pme=None,
                rvdw=None):
    """Energy tolerance: 1e-6 relative, or on small boxes three times the
    FP32 reference kernel's own RMS rounding noise there, whichever is larger
    (the 9,000- and 288,000-atom boxes are held to the plain 1e-6)."""
    return max(ENERGY_TOL_REL * scale,
               3.0 * _fp32_noise(s, coulomb, ewald_exact, vdw=vdw, lj_pme=lj_pme, rvdw=rvdw)[key])


def _compare(s, ctx, coulomb="ewald", widen_energy=True, vdw=("potential-shift", 0.0), lj_pme=None, rvdw=None,
             ewald_exact=False):
    f = ctx.compute(s.x, energy=True, virial=True)
    vlj, vc = ctx.energies()
    r64 = oracle_for(s, "f64", coulomb=coulomb, vdw=vdw, lj_pme=lj_pme, rvdw=rvdw, ewald_exact=ewald_exact).compute(s.x)
    r32 = oracle_for(s, "f32", coulomb=coulomb, vdw=vdw, lj_pme=lj_pme, rvdw=rvdw,
                     ewald_exact=ewald_exact).compute(s.x)
    ferr64 = rms(f - r64["f"]) / rms(r64["f"])
    ferr32 = rms(f - r32["f"]) / rms(r64["f"])
    e64 = r64["Vvdw"] + r64["Vc"]
    print(f"  force rms err vs f64 {ferr64:.2e} vs f32 {ferr32:.2e}; "
          f"Vlj {vlj:.6f} ({r64['Vvdw']:.6f}, f32 {r32['Vvdw']:.6f}) "
          f"Vc {vc:.6f} ({r64['Vc']:.6f}, f32 {r32['Vc']:.6f})", flush=True)
    assert ferr64 <= FORCE_TOL_RMS
    assert ferr32 <= FORCE_TOL_RMS
    for got, key in ((vlj, "Vvdw"), (vc, "Vc")):
        scale = max(abs(r64[key]), abs(e64))
        bar = (_energy_bar(s, key, scale, coulomb, ewald_exact, vdw=vdw, lj_pme=lj_pme, rvdw=rvdw) if widen_energy
               else ENERGY_TOL_REL * scale)
        assert abs(got - r64[key]) <= bar, (key, got, r64[key], bar)
    vir = ctx.virial()
    np.testing.assert_allclose(vir, r64["virial"], rtol=0, atol=1e-4 * np.abs(r64["virial"]).max())
    fsh = ctx.shift_forces()
    np.testing.assert_allclose(fsh, r64["fshift"], rtol=0, atol=1e-4 * max(1.0, np.abs(r64["fshift"]).max()))
    return f


@pytest.mark.parametrize("name,s", _systems())
def test_forces_energies_virial(name, s):
    print(f"[forces] {name}", flush=True)
    ctx = nbx.context_for(s)
    _compare(s, ctx)


def test_forces_config0(water3000):
    print("[forces] config0 9000 atoms", flush=True)
    ctx = nbx.context_for(water3000)
    _compare(water3000, ctx, widen_energy=False)


def test_forces_energies_water288k():
    # the full-size water box: forces against the FP64 oracle, energies to
    # 1e-6 of each component on its own
    print("[forces] water 288k atoms", flush=True)
    s = S.config(1)
    ctx = nbx.context_for(s)
    f = ctx.compute(s.x, energy=True, virial=True)
    vlj, vc = ctx.energies()
    r64 = oracle_for(s, "f64").compute(s.x)
    ferr = rms(f - r64["f"]) / rms(r64["f"])
    print(f"  force rms err {ferr:.2e}; Vlj {vlj:.6f} ({r64['Vvdw']:.6f}) Vc {vc:.6f} ({r64['Vc']:.6f})", flush=True)
    assert ferr <= FORCE_TOL_RMS
    assert abs(vlj - r64["Vvdw"]) <= ENERGY_TOL_REL * abs(r64["Vvdw"])
    assert abs(vc - r64["Vc"]) <= ENERGY_TOL_REL * abs(r64["Vc"])
    np.testing.assert_allclose(ctx.virial(), r64["virial"], rtol=0, atol=1e-5 * np.abs(r64["virial"]).max())


@pytest.mark.parametrize("name,s", _systems()[1:])
def test_analytical_ewald_against_exact_oracle(name, s):
    # the analytical correction against the FP64 oracle with exact erfc
    # (what the reference's SIMD kernels approximate, simd_math.h:1609-1720)
    print(f"[analytical] {name}", flush=True)
    ctx = nbx.context_for(s, ewald_correction="analytical")
    f = ctx.compute(s.x, energy=True, virial=True)
    vlj, vc = ctx.energies()
    r64 = oracle_for(s, "f64", ewald_exact=True).compute(s.x)
    ferr = rms(f - r64["f"]) / rms(r64["f"])
    e64 = r64["Vvdw"] + r64["Vc"]
    print(f"  force rms err {ferr:.2e}; Vlj {vlj:.6f} ({r64['Vvdw']:.6f}) Vc {vc:.6f} ({r64['Vc']:.6f})", flush=True)
    assert ferr <= FORCE_TOL_RMS
    for got, key in ((vlj, "Vvdw"), (vc, "Vc")):
        scale = max(abs(r64[key]), abs(e64))
        assert abs(got - r64[key]) <= _energy_bar(s, key, scale, ewald_exact=True), (key, got, r64[key])


def test_analytical_ewald_config0(water3000):
    print("[analytical] config0 9000 atoms", flush=True)
    s = water3000
    ctx = nbx.context_for(s, ewald_correction="analytical")
    f = ctx.compute(s.x, energy=True)
    vlj, vc = ctx.energies()
    r64 = oracle_for(s, "f64", ewald_exact=True).compute(s.x)
    ferr = rms(f - r64["f"]) / rms(r64["f"])
    print(f"  force rms err {ferr:.2e}; Vc {vc:.6f} ({r64['Vc']:.6f})", flush=True)
    assert ferr <= FORCE_TOL_RMS
    assert abs(vlj - r64["Vvdw"]) <= ENERGY_TOL_REL * abs(r64["Vvdw"] + r64["Vc"])
    assert abs(vc - r64["Vc"]) <= ENERGY_TOL_REL * abs(r64["Vc"])


def test_energies_are_bitwise_reproducible():
    # the static schedule fixes which wavefront sums which tiles, and the
    # per-wavefront group sums are added in a fixed order
    s = S.config(0)
    ctx = nbx.context_for(s)
    ctx.compute(s.x, energy=True)
    e1 = ctx.energies()
    ctx.compute(s.x, energy=True)
    assert ctx.energies() == e1


@pytest.mark.parametrize("coulomb", ["cutoff", "reaction-field"])
def test_reaction_field_flavours(coulomb):
    s = S.water_box(300, 2.1, seed=5)
    ctx = nbx.context_for(s, coulomb=coulomb)
    _compare(s, ctx, coulomb="rf")


@pytest.mark.parametrize("modifier", ["force-switch", "potential-switch"])
@pytest.mark.parametrize("name,s", _systems()[1:])
def test_lj_modifiers(name, s, modifier):
    # the plain-C kernel's LJ_FORCE_SWITCH / LJ_POT_SWITCH flavours
    # (kernel_ref_inner.h:155-205) against the FP64 and FP32 oracle; the
    # water box has types without LJ (the skip test), the chains 5 types and
    # bonded exclusions
    print(f"[lj modifier] {name} {modifier}", flush=True)
    vdw = (modifier, 0.6)
    ctx = nbx.context_for(s, vdw_modifier=modifier, rvdw_switch=vdw[1])
    _compare(s, ctx, vdw=vdw)


@pytest.mark.parametrize("rule", ["geometric", "lorentz-berthelot"])
@pytest.mark.parametrize("name,s", _systems()[1:])
def test_lj_pme(name, s, rule):
    # LJ-PME's real-space kernel (kernel_ref_inner.h:207-248) with the grid's
    # C6 by either combination rule, self terms included, against the FP64
    # and FP32 oracle
    print(f"[lj-pme] {name} {rule}", flush=True)
    ctx = nbx.context_for(s, vdw_type="pme", ljpme_comb_rule=rule)
    _compare(s, ctx, lj_pme=rule)


@pytest.mark.parametrize("opts,oracle_opts", [
    ({"vdw_type": "pme"}, {"lj_pme": "geometric"}),
    ({"vdw_modifier": "force-switch", "rvdw_switch": 0.8}, {"vdw": ("force-switch", 0.8)}),
    ({"vdw_cutoff": 0.8}, {"rvdw": 0.8}),
], ids=["lj-pme", "force-switch", "twin-cutoff"])
def test_lj_flavours_water288k(opts, oracle_opts):
    # the full-size water box with another LJ flavour: forces against the
    # FP64 oracle, energies to 1e-6 of each component on its own
    print(f"[lj flavour] water 288k atoms {opts}", flush=True)
    s = S.config(1)
    ctx = nbx.context_for(s, **opts)
    f = ctx.compute(s.x, energy=True, virial=True)
    vlj, vc = ctx.energies()
    r64 = oracle_for(s, "f64", **oracle_opts).compute(s.x)
    ferr = rms(f - r64["f"]) / rms(r64["f"])
    print(f"  force rms err {ferr:.2e}; Vlj {vlj:.6f} ({r64['Vvdw']:.6f}) Vc {vc:.6f} ({r64['Vc']:.6f})", flush=True)
    assert ferr <= FORCE_TOL_RMS
    assert abs(vlj - r64["Vvdw"]) <= ENERGY_TOL_REL * abs(r64["Vvdw"])
    assert abs(vc - r64["Vc"]) <= ENERGY_TOL_REL * abs(r64["Vc"])


def test_lorentz_berthelot_parameters():
    # the kernel reads the full C6/C12 matrices, so Lorentz-Berthelot
    # parameters (nbnxm's ljcrLB, atomdata.cpp:540-590) need only the host's
    # matrix builder; checked against the FP64 oracle on the same matrices
    _, s = _systems()[2]
    sig = (s.c12 / s.c6) ** (1.0 / 6.0)
    eps = s.c6 ** 2 / (4.0 * s.c12)
    m6, m12 = nbx.lorentz_berthelot_c6c12_matrices(sig, eps)
    ctx = nbx.Context(rlist=s.rlist, rcoulomb=s.rcoulomb)
    ctx.set_nonbonded_parameters(m6, m12)
    ctx.set_topology(s.types, s.charges, s.excl_off, s.excl_idx)
    ctx.put_on_grid(s.box, s.x)
    ctx.construct_pairlist()
    f = ctx.compute(s.x, energy=True)
    vlj, vc = ctx.energies()
    o = O.Oracle("f64", "ewald", rlist=s.rlist, rcoulomb=s.rcoulomb)
    nbfp = np.stack([6.0 * m6, 12.0 * m12], axis=-1).reshape(-1)
    o.setup(s.box, s.x, s.types, s.charges, nbfp, s.ntypes, s.excl_off, s.excl_idx)
    r64 = o.compute(s.x)
    geo = oracle_for(s, "f64").compute(s.x)
    ferr = rms(f - r64["f"]) / rms(r64["f"])
    print(f"  LB force rms err {ferr:.2e}; Vlj {vlj:.6f} ({r64['Vvdw']:.6f}, geometric {geo['Vvdw']:.6f})",
          flush=True)
    assert ferr <= FORCE_TOL_RMS
    # energies: 1e-6, or three times the FP32 kernel's rounding noise on
    # this box (estimated with the geometric matrices: same system, same
    # magnitudes), as for the other small boxes
    e64 = r64["Vvdw"] + r64["Vc"]
    for got, key in ((vlj, "Vvdw"), (vc, "Vc")):
        scale = max(abs(r64[key]), abs(e64))
        assert abs(got - r64[key]) <= _energy_bar(s, key, scale), (key, got, r64[key])
    # the rule matters on this system
    assert abs(geo["Vvdw"] - r64["Vvdw"]) > 1e-3 * abs(r64["Vvdw"])


def test_lj_modifier_option_errors():
    s = _systems()[0][1]
    with pytest.raises(nbx.NbxInputError):
        nbx.context_for(s, vdw_modifier="force-switch", rvdw_switch=s.rcoulomb)
    with pytest.raises(nbx.NbxInputError):
        nbx.context_for(s, vdw_modifier="potential-switch", rvdw_switch=-0.1)
    with pytest.raises(nbx.NbxInputError):
        nbx.context_for(s, vdw_modifier="exact-cutoff")
    # the adjoint covers the LJ cut-off with the potential shift only
    ctx = nbx.context_for(s, vdw_modifier="force-switch", rvdw_switch=0.5)
    with pytest.raises(nbx.NbxInputError):
        ctx.compute_adjoint(s.x, np.zeros_like(s.x))
    ctx = nbx.context_for(s, vdw_type="pme")
    with pytest.raises(nbx.NbxInputError):
        ctx.compute_adjoint(s.x, np.zeros_like(s.x))
    # LJ-PME takes the potential shift
    with pytest.raises(nbx.NbxInputError):
        nbx.context_for(s, vdw_type="pme", vdw_modifier="force-switch", rvdw_switch=0.5)


@pytest.mark.parametrize("name", ["nblib_argon.json", "nblib_spc_methanol.json"])
def test_nblib_golden_forces(name):
    # api/nblib/tests/nbkernelsystem.cpp:69-85,187-203 refdata
    d = load_golden(name)
    ctx = nbx.Context(coulomb="cutoff", rlist=d["rc"])
    m6, m12 = nbx.nblib_c6c12_matrices(d["c6"], d["c12"])
    ctx.set_nonbonded_parameters(m6, m12)
    ctx.set_topology(d["types"], d["charges"], d["excl_off"], d["excl_idx"])
    ctx.put_on_grid([d["box"]] * 3, d["x"])
    ctx.construct_pairlist()
    f = ctx.compute(d["x"])
    ref = np.array(d["forces"])
    assert np.abs(f - ref).max() <= 5e-5 * np.abs(ref).max()


def test_repeated_compute_is_reproducible():
    # nbkernelsystem.cpp:150-167 compares repeated computes with equality:
    # the reference's CPU kernel is bitwise reproducible for a fixed thread
    # count.  Here the j-force windows are staged and summed per cluster in
    # a fixed order and the energies in fixed slots, so repeated calls are
    # bit-identical too (no window spills at this size).
    s = S.small_water_box(216, seed=11)
    ctx = nbx.context_for(s)
    a = ctx.compute(s.x, energy=True, virial=True)
    ea = ctx.energies()
    va = ctx.virial()
    b = ctx.compute(s.x, energy=True, virial=True)
    eb = ctx.energies()
    vb = ctx.virial()
    np.testing.assert_array_equal(a, b)
    assert tuple(ea) == tuple(eb)
    np.testing.assert_array_equal(va, vb)


def test_shift_forces_and_virial_are_reproducible(water3000):
    # each wavefront sums its own shift forces and the workgroup adds the
    # wavefronts' arrays in order: shift forces and the virial's shift term
    # are bit-identical from call to call (the 4.48 nm box puts most
    # clusters within rlist of a face, so many wavefronts carry shifts)
    # The j-forces of clusters that do not fit a group's LDS window (slot
    # 0xffff in the work list; long cut-offs relative to the box) are added
    # with float atomics: then forces and the virial's x (x) f term agree to
    # rounding only, while the shift forces (from the i-forces) and the
    # energies stay bit-identical.
    # NBX_COMPUTE_REPRODUCIBLE sums those in fixed point instead: then
    # everything is bit-identical, and equal to the default to rounding.
    s = water3000
    ctx = nbx.context_for(s)
    spill = bool(np.any((ctx.work_groups()["work"][:, 1] >> 16) == 0xFFFF))
    print(f"[reproducible] config0, window spill: {spill}", flush=True)
    fixed = []
    for _ in range(3):
        f = ctx.compute(s.x, energy=True, shift_forces=True, virial=True, reproducible=True)
        fixed.append((f, ctx.shift_forces(), ctx.virial(), tuple(ctx.energies())))
    for got in fixed[1:]:
        for k in range(3):
            np.testing.assert_array_equal(got[k], fixed[0][k])
        assert got[3] == fixed[0][3]
    ref = None
    for _ in range(5):
        f = ctx.compute(s.x, energy=True, shift_forces=True, virial=True)
        got = (f, ctx.shift_forces(), ctx.virial(), tuple(ctx.energies()))
        if ref is None:
            ref = got
            assert np.abs(got[1]).max() > 0
            continue
        np.testing.assert_array_equal(got[1], ref[1])
        assert got[3] == ref[3]
        if spill:
            np.testing.assert_allclose(got[0], ref[0], rtol=0, atol=1e-6 * np.abs(ref[0]).max())
            np.testing.assert_allclose(got[2], ref[2], rtol=0, atol=1e-6 * np.abs(ref[2]).max())
        else:
            np.testing.assert_array_equal(got[0], ref[0])
            np.testing.assert_array_equal(got[2], ref[2])
    np.testing.assert_allclose(fixed[0][0], ref[0], rtol=0, atol=1e-5 * np.abs(ref[0]).max())
    np.testing.assert_allclose(fixed[0][2], ref[2], rtol=0, atol=1e-6 * np.abs(ref[2]).max())
    assert fixed[0][3] == ref[3]


def test_water288k_reproducible_mode():
    # the 288,000-atom water box overflows some windows (about 0.2 % of the
    # j-entries): with NBX_COMPUTE_REPRODUCIBLE its forces, virial and shift
    # forces are bit-identical from call to call and agree with the FP64
    # oracle as the default path does
    s = S.config(1)
    ctx = nbx.context_for(s)
    spill = int(np.sum((ctx.work_groups()["work"][:, 1] >> 16) == 0xFFFF))
    print(f"[reproducible] water 288k, overflow entries: {spill}", flush=True)
    runs = []
    for _ in range(3):
        f = ctx.compute(s.x, energy=True, shift_forces=True, virial=True, reproducible=True)
        runs.append((f, ctx.virial(), ctx.shift_forces()))
    for got in runs[1:]:
        for k in range(3):
            np.testing.assert_array_equal(got[k], runs[0][k])
    r64 = oracle_for(s, "f64").compute(s.x)
    assert rms(runs[0][0] - r64["f"]) / rms(r64["f"]) <= FORCE_TOL_RMS


def test_update_changes_forces():
    # nbkernelsystem.cpp:169-184: forces change after the coordinates move
    s = S.small_water_box(216, seed=11)
    ctx = nbx.context_for(s)
    a = ctx.compute(s.x)
    x2 = s.x + 0.01 * np.random.default_rng(0).standard_normal(s.x.shape)
    x2 = np.mod(x2, s.box)
    b = ctx.compute(x2)
    assert np.all(np.abs(a - b).sum(axis=1) > 0)


def test_input_errors():
    s = S.small_water_box(216, seed=11)
    ctx = nbx.Context(rlist=1.0)
    with pytest.raises(nbx.NbxStateError):
        ctx.construct_pairlist()
    m6, m12 = nbx.nblib_c6c12_matrices(s.c6, s.c12)
    ctx.set_nonbonded_parameters(m6, m12)
    bad = s.types.copy()
    bad[0] = 7
    with pytest.raises(nbx.NbxInputError):
        ctx.set_topology(bad, s.charges, s.excl_off, s.excl_idx)
    ctx.set_topology(s.types, s.charges, s.excl_off, s.excl_idx)
    x = s.x.copy()
    x[5, 0] = -1.5
    with pytest.raises(nbx.NbxInputError):
        ctx.put_on_grid(s.box, x)
    with pytest.raises(nbx.NbxInputError):
        ctx.put_on_grid([1.5, 1.5, 1.5], s.x)
    with pytest.raises(nbx.NbxStateError):
        ctx.compute(s.x)
    with pytest.raises(nbx.NbxInputError):
        nbx.Context(rlist=1.0, rcoulomb=1.2)


@pytest.mark.parametrize("name,s", _systems())
def test_work_list_covers_the_pair_list(name, s):
    # The kernel's work list holds every (ci, shift, cj, mask) entry of the
    # pair list exactly once, in a chunk of ci's group, with a window slot
    # that maps back to cj; padding entries are the far-away dummy cluster in
    # the i-cluster's own slot; masked entries lie in the chunk's leading
    # "special" tiles; one self chunk per cluster.
    print(f"[work] {name}", flush=True)
    ctx = nbx.context_for(s)
    wg = ctx.work_groups()
    ie, je = ctx.pairlist()
    nc = ctx.info()["num_clusters"]
    G = wg["group_size"]
    assert G >= 1 and wg["num_groups"] == (nc + G - 1) // G
    want = sorted((int(c), int(sh), int(je[k, 0]), int(je[k, 1]) & 0xFFFF)
                  for c, sh, a, b in ie for k in range(a, b))
    got, self_chunks = [], []
    for g, (ch0, ch1, uo, us) in enumerate(wg["groups"]):
        union = wg["union"][uo:uo + us]
        assert np.all(np.diff(union) > 0)
        assert not np.any((union >= g * G) & (union < (g + 1) * G))
        for c in range(ch0, ch1):
            ci, flags, ws, tiles = (int(v) for v in wg["chunks"][c])
            sh, is_self, special = flags & 0xFF, (flags >> 8) & 1, flags >> 9
            assert g * G <= ci < (g + 1) * G and tiles >= 1 and special <= tiles
            if is_self:
                assert sh == 22
                self_chunks.append(ci)
            for k in range(ws, ws + 16 * tiles):
                cj, w = int(wg["work"][k, 0]), int(wg["work"][k, 1])
                slot, mask = w >> 16, w & 0xFFFF
                if cj == nc:
                    # padding: the far-away dummy cluster in the spare slot
                    assert slot == G + len(union) and mask == 0xFFFF
                    continue
                if slot == 0xFFFF:
                    # a union cluster beyond the window's capacity
                    assert not (g * G <= cj < (g + 1) * G)
                else:
                    assert (g * G + slot if slot < G else union[slot - G]) == cj
                if mask != 0xFFFF:
                    assert (k - ws) // 16 < special
                got.append((ci, sh, cj, mask))
    assert sorted(got) == want
    assert sorted(self_chunks) == list(range(nc))
    # groups, chunks and work entries are laid out back to back in group
    # order: the kernel's static schedule walks them as one stream
    gr, ch = wg["groups"], wg["chunks"]
    assert gr[0, 0] == 0 and np.all(gr[1:, 0] == gr[:-1, 1]) and gr[-1, 1] == len(ch)
    assert ch[0, 2] == 0 and np.all(ch[1:, 2] == ch[:-1, 2] + 16 * ch[:-1, 3])
    assert ch[-1, 2] + 16 * ch[-1, 3] == len(wg["work"])
    for c, sh, a, b in ie:
        masked = (je[a:b, 1] & 0xFFFF) != 0xFFFF
        assert not np.any(np.diff(masked.astype(int)) > 0), "masked entries must lead"


def test_add_forces_accumulates_on_device():
    # NBX_COMPUTE_ADD_FORCES: f += nonbonded forces (atomdata_add_nbat_f_to_f
    # into a buffer holding other force sources); forces are bitwise
    # reproducible, so the sum is exact to the one float addition
    import torch

    print("[forces] add to an existing device force buffer", flush=True)
    s = S.small_water_box(216, seed=11)
    ctx = nbx.context_for(s)
    dev = torch.device("cuda:0")
    x = torch.tensor(s.x, dtype=torch.float32, device=dev).contiguous()
    f_nb = torch.empty_like(x)
    ctx.compute_device(x.data_ptr(), f_nb.data_ptr(), nbx.COMPUTE_FORCES)
    f0 = torch.randn_like(x)
    f = f0.clone()
    ctx.compute_device(x.data_ptr(), f.data_ptr(), nbx.COMPUTE_FORCES | nbx.COMPUTE_ADD_FORCES)
    torch.cuda.synchronize()
    assert torch.equal(f, f0 + f_nb)
    # and the added forces are the path's forces: against the FP64 oracle
    r64 = oracle_for(s, "f64").compute(s.x)
    added = (f - f0).cpu().numpy().astype(np.float64)
    ferr = rms(added - r64["f"]) / rms(r64["f"])
    print(f"  added forces rms err vs FP64 oracle {ferr:.2e}", flush=True)
    assert ferr <= FORCE_TOL_RMS
    # host arrays too
    h0 = np.random.default_rng(3).normal(size=s.x.shape).astype(np.float32)
    h = h0.copy()
    xs = np.ascontiguousarray(s.x, np.float32)
    ctx._check(ctx._lib.nbx_compute(ctx._h, nbx._ptr(xs), nbx._ptr(h),
                                    nbx.COMPUTE_FORCES | nbx.COMPUTE_ADD_FORCES, 0), "nbx_compute")
    np.testing.assert_array_equal(h, h0 + f_nb.cpu().numpy())


@pytest.mark.parametrize("vdw", [("potential-shift", 0.0), ("force-switch", 0.5), ("potential-switch", 0.5)],
                         ids=["shift", "force-switch", "potential-switch"])
@pytest.mark.parametrize("correction", ["table", "analytical"])
@pytest.mark.parametrize("name,s", _systems()[1:])
def test_twin_cutoff(name, s, correction, vdw):
    # rvdw < rcoulomb: the TAB_TWIN / EWALD_TWIN kernels with LJ masked
    # beyond rvdw (kernel_ref_inner.h:254-271), shift and switches at rvdw
    print(f"[twin cut-off] {name} rvdw 0.7 {correction} {vdw[0]}", flush=True)
    ctx = nbx.context_for(s, ewald_correction=correction, vdw_cutoff=0.7, vdw_modifier=vdw[0],
                          rvdw_switch=vdw[1])
    _compare(s, ctx, vdw=vdw, rvdw=0.7, ewald_exact=correction == "analytical")


def test_twin_cutoff_option_errors():
    s = _systems()[0][1]
    for opts in ({"coulomb": "reaction-field", "vdw_cutoff": 0.8}, {"vdw_type": "pme", "vdw_cutoff": 0.8},
                 {"vdw_cutoff": 1.2}, {"vdw_cutoff": 0.8, "vdw_modifier": "force-switch", "rvdw_switch": 0.8}):
        with pytest.raises(nbx.NbxInputError):
            nbx.context_for(s, **opts)
    ctx = nbx.context_for(s, vdw_cutoff=0.8)
    with pytest.raises(nbx.NbxInputError):
        ctx.compute_adjoint(s.x, np.zeros_like(s.x))


@pytest.mark.parametrize("opts,oracle_opts", [
    ({}, {}),
    ({"coulomb": "reaction-field"}, {"coulomb": "rf"}),
    ({"vdw_modifier": "force-switch", "rvdw_switch": 0.6}, {"vdw": ("force-switch", 0.6)}),
    ({"vdw_modifier": "potential-switch", "rvdw_switch": 0.6}, {"vdw": ("potential-switch", 0.6)}),
    ({"vdw_type": "pme"}, {"lj_pme": "geometric"}),
    ({"vdw_cutoff": 0.8, "vdw_modifier": "force-switch", "rvdw_switch": 0.6},
     {"rvdw": 0.8, "vdw": ("force-switch", 0.6)}),
], ids=["ewald", "rf", "force-switch", "potential-switch", "lj-pme", "twin-force-switch"])
def test_interaction_constants_match_oracle(opts, oracle_opts):
    # interaction_const_t as the library computes it (float) against the
    # oracle's FP32 restatement: beta, table scale and size, epsfac, RF
    # constants, potential shifts, switch and LJ-PME constants, and the
    # Ewald correction table the kernel reads
    s = S.small_water_box(216, seed=11)
    ctx = nbx.context_for(s, **opts)
    o = oracle_for(s, "f32", **oracle_opts)
    got, ref = ctx.interaction_constants(), o.constants()
    for k in ref:
        assert got[k] == pytest.approx(ref[k], rel=2e-7, abs=1e-30), (k, got[k], ref[k])
    got, ref = ctx.vdw_constants(), o.vdw_constants()
    for k in ref:
        assert got[k] == pytest.approx(ref[k], rel=2e-6, abs=1e-30), (k, got[k], ref[k])
    tab = ctx.ewald_table()
    if oracle_opts.get("coulomb", "ewald") == "ewald":
        F, V, fdv0 = o.tables()
        fdv0 = fdv0.reshape(-1, 4)
        assert tab.shape == fdv0.shape
        np.testing.assert_array_equal(tab[:, :3], fdv0[:, :3])
        # the fourth entry holds V + sh_ewald (sh_ewald = 0 on nblib's path)
        np.testing.assert_array_equal(tab[:, 3], fdv0[:, 2] + np.float32(ref_sh(o)))
    else:
        assert tab.shape[0] == 0


def ref_sh(o):
    return o.constants()["sh_ewald"]


def _sparse(name, box, x, q):
    n = len(x)
    return S.System(name, np.array(box, float), np.array(x, float), np.zeros(n, np.int32), np.array(q, float),
                    np.array([0.0026]), np.array([2.6e-6]), np.arange(n + 1, dtype=np.int32),
                    np.arange(n, dtype=np.int32), 1.0, 1.0)


def _corner_lattice():
    g = np.arange(3) * 0.35 + 0.2
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return _sparse("corner27", [6.0, 6.0, 6.0], x, np.tile([0.4, -0.4], 14)[:27])


@pytest.mark.parametrize("s", [
    _sparse("two-far", [5.0, 5.0, 5.0], [[0.5, 0.5, 0.5], [3.0, 3.0, 3.0]], [0.5, -0.5]),
    _sparse("two-near", [5.0, 5.0, 5.0], [[0.5, 0.5, 0.5], [0.9, 0.5, 0.5]], [0.5, -0.5]),
    _sparse("two-across-the-boundary", [5.0, 5.0, 5.0], [[0.1, 2.5, 2.5], [4.8, 2.5, 2.5]], [0.5, -0.5]),
    _corner_lattice(),
], ids=lambda s: s.name)
def test_sparse_systems(s):
    # nearly empty boxes: single-atom clusters, columns without atoms, a
    # pair interacting only through the periodic boundary, a lattice filling
    # one corner; forces and energies against the FP64 oracle.  The energy
    # bar adds what float coordinates allow: a distance formed from
    # coordinates up to L is off by up to about L 2^-23, which changes a
    # pair's energy by |F| times that (a single pair across a 5 nm boundary,
    # d = 0.1 + 5 - 4.8 in float, is 1e-6 off: 3e-5 of its LJ energy)
    print(f"[sparse] {s.name}: {s.natoms} atoms", flush=True)
    ctx = nbx.context_for(s)
    f = ctx.compute(s.x, energy=True, virial=True)
    vlj, vc = ctx.energies()
    r64 = oracle_for(s, "f64").compute(s.x)
    np.testing.assert_allclose(f, r64["f"], rtol=0, atol=1e-4 * max(1.0, np.abs(r64["f"]).max()))
    coord = 0.5 * np.abs(r64["f"]).sum() * np.max(s.box) * 2.0 ** -23
    assert abs(vlj - r64["Vvdw"]) <= 1e-6 * abs(r64["Vvdw"]) + coord
    assert abs(vc - r64["Vc"]) <= 1e-6 * abs(r64["Vc"]) + coord
    np.testing.assert_allclose(ctx.list_pairs(s.rlist), oracle_for(s, "f32").list_pairs(s.rlist))


def test_water288k_invariances():
    # size-independent properties at full size: the forces sum to zero
    # (every pair force appears twice, opposite), and translating every atom
    # (wrapped into the box, so grid, clusters and list all change) leaves
    # each atom's force and the energies unchanged up to float rounding: the
    # translated coordinates round differently in float (measured 8e-6 of the
    # RMS force), while one O-O pair missing from the list would move the RMS
    # difference by about 1.4e-4 of it
    s = S.config(1)
    ctx = nbx.context_for(s)
    f = ctx.compute(s.x, energy=True).astype(np.float64)
    e = ctx.energies()
    scale = rms(f)
    print(f"[invariance] 288k: |sum f| {np.abs(f.sum(0)).max():.3e}, rms f {scale:.3e}", flush=True)
    assert np.abs(f.sum(0)).max() <= 1e-6 * scale * np.sqrt(s.natoms)
    t = np.array([0.37, 1.1, 2.3])
    moved = np.mod(s.x + t, s.box)
    ctx2 = nbx.context_for(s)
    ctx2.put_on_grid(s.box, moved)
    ctx2.construct_pairlist()
    f2 = ctx2.compute(moved, energy=True).astype(np.float64)
    e2 = ctx2.energies()
    assert rms(f2 - f) <= 3e-5 * scale
    assert abs(e2[0] - e[0]) <= 1e-6 * abs(e[0])
    assert abs(e2[1] - e[1]) <= 1e-6 * abs(e[1])
