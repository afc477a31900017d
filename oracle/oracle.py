"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle is a plain-C restatement of the reference's nbnxm path (grid,
4x4 pair search, exclusion masks, the plain-C reference kernel and the force
reduction), built in float ("f32", the reference's mixed-precision build) and
in double ("f64", the reference's double build).  Only tests/, the smoke
check and bench.py's CPU-baseline leg use it, and only as the checker.

Each method cites the reference function it stands for; the C source cites
file:line for every step.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

_LIBS: dict[str, C.CDLL] = {}


def build() -> None:
    """Compile both precisions with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(precision: str) -> C.CDLL:
    """f32 / f64: the checker; bench_f32: the CPU-baseline build of the float
    oracle (-O3, AVX2/FMA, OpenMP), which only bench.py's cpu_baseline uses."""
    if precision not in ("f32", "f64", "bench_f32"):
        raise ValueError(f"precision must be f32, f64 or bench_f32, not {precision!r}")
    if precision in _LIBS:
        return _LIBS[precision]
    path = os.path.join(_BUILD, f"liboracle_{precision}.so")
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    real = C.c_double if precision == "f64" else C.c_float
    if precision == "bench_f32":
        lib.orc_compute_threads.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        lib.orc_compute_simd.restype = C.c_int
        lib.orc_compute_simd.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        lib.orc_baseline_simd_available.restype = C.c_int
        lib.orc_baseline_simd_available.argtypes = [C.c_void_p]
    vp = C.c_void_p
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_int]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_set_vdw_cutoff.restype = C.c_int
    lib.orc_set_vdw_cutoff.argtypes = [vp, C.c_double]
    lib.orc_set_vdw_modifier.restype = C.c_int
    lib.orc_set_vdw_modifier.argtypes = [vp, C.c_int, C.c_double]
    lib.orc_set_lj_pme.restype = C.c_int
    lib.orc_set_lj_pme.argtypes = [vp, C.c_int, C.c_double]
    lib.orc_lj_pme_beta.restype = real
    lib.orc_lj_pme_beta.argtypes = [vp]
    lib.orc_ewald_coeff_lj.restype = real
    lib.orc_ewald_coeff_lj.argtypes = [real, real]
    lib.orc_ewald_coeff.restype = real
    lib.orc_ewald_coeff.argtypes = [real, real]
    lib.orc_ewald_table_scale.restype = real
    lib.orc_ewald_table_scale.argtypes = [real, real]
    lib.orc_ewald_table_size.restype = C.c_int
    lib.orc_ewald_table_size.argtypes = [real, real]
    lib.orc_ewald_tables.argtypes = [C.c_int, C.c_double, real, vp, vp, vp]
    lib.orc_ewald_v_exact.restype = C.c_double
    lib.orc_ewald_v_exact.argtypes = [C.c_double, C.c_double]
    lib.orc_ewald_f_exact.restype = C.c_double
    lib.orc_ewald_f_exact.argtypes = [C.c_double, C.c_double]
    lib.orc_one_4pi_eps0.restype = C.c_double
    lib.orc_get_constants.argtypes = [vp, vp]
    lib.orc_get_tables.argtypes = [vp, vp, vp, vp]
    lib.orc_get_vdw_constants.argtypes = [vp, vp]
    lib.orc_set_nbfp.argtypes = [vp, C.c_int, vp]
    lib.orc_set_exclusions.argtypes = [vp, C.c_int, vp, vp]
    lib.orc_put_on_grid.restype = C.c_int
    lib.orc_put_on_grid.argtypes = [vp, vp, C.c_int, vp]
    lib.orc_set_atom_props.argtypes = [vp, vp, vp]
    lib.orc_convert_coordinates.argtypes = [vp, vp]
    lib.orc_get_grid_dims.argtypes = [vp, vp]
    lib.orc_get_grid.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.orc_get_shiftvec.argtypes = [vp, vp]
    lib.orc_construct_pairlist.restype = C.c_int
    lib.orc_construct_pairlist.argtypes = [vp]
    lib.orc_prune_pairlist.restype = C.c_int
    lib.orc_prune_pairlist.argtypes = [vp, C.c_double]
    lib.orc_num_ci.restype = C.c_int
    lib.orc_num_ci.argtypes = [vp]
    lib.orc_num_cj.restype = C.c_int
    lib.orc_num_cj.argtypes = [vp]
    lib.orc_get_list.argtypes = [vp, vp, vp, vp]
    lib.orc_list_pairs.restype = C.c_int64
    lib.orc_list_pairs.argtypes = [vp, C.c_double, vp, C.c_int64]
    lib.orc_brute_pairs.restype = C.c_int64
    lib.orc_brute_pairs.argtypes = [C.c_int, vp, vp, C.c_double, vp, C.c_int64]
    lib.orc_kernel.argtypes = [vp, C.c_int]
    lib.orc_compute.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp]
    lib.orc_brute_forces.argtypes = [C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp, C.c_int, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, vp, vp, vp]
    _LIBS[precision] = lib
    return lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def one_4pi_eps0() -> float:
    """ONE_4PI_EPS0 (math/units.h:79) in kJ mol^-1 nm e^-2."""
    return float(_lib("f64").orc_one_4pi_eps0())


def ewald_coeff(rc: float, rtol: float, precision: str = "f32") -> float:
    """calc_ewaldcoeff_q (ewald/ewald_utils.cpp:46-74)."""
    return float(_lib(precision).orc_ewald_coeff(rc, rtol))


def ewald_tables(rc: float, beta: float, precision: str = "f32"):
    """init_ewald_f_table + generateEwaldCorrectionTables
    (mdlib/forcerec.cpp:726-764, tables/forcetable.cpp:144-387).
    Returns (scale, F, V, FDV0)."""
    lib = _lib(precision)
    dt = np.float32 if precision == "f32" else np.float64
    scale = lib.orc_ewald_table_scale(beta, rc)
    n = lib.orc_ewald_table_size(rc, scale)
    F = np.zeros(n, dt)
    V = np.zeros(n, dt)
    FDV0 = np.zeros(4 * n, dt)
    lib.orc_ewald_tables(n, float(scale), beta, _p(F), _p(V), _p(FDV0))
    return float(scale), F, V, FDV0


def ewald_v_exact(beta: float, r: float) -> float:
    return float(_lib("f64").orc_ewald_v_exact(beta, r))


def ewald_f_exact(beta: float, r: float) -> float:
    return float(_lib("f64").orc_ewald_f_exact(beta, r))


def nblib_nbfp(c6_type, c12_type, precision: str) -> np.ndarray:
    """nbfp as nblib builds it: geometric C6/C12 combination in real
    (api/nblib/interactions.cpp:83-93) times 6 and 12 in real
    (gmxsetup.cpp:143-166).  Returns ntype*ntype*2 reals."""
    dt = np.float32 if precision == "f32" else np.float64
    c6 = np.asarray(c6_type, dt)
    c12 = np.asarray(c12_type, dt)
    t = len(c6)
    out = np.zeros((t, t, 2), dt)
    for i in range(t):
        for j in range(t):
            out[i, j, 0] = np.sqrt(dt(c6[i] * c6[j])) * dt(6.0)
            out[i, j, 1] = np.sqrt(dt(c12[i] * c12[j])) * dt(12.0)
    return out.reshape(-1)


def brute_pairs(x, box, rlist: float, precision: str = "f32") -> np.ndarray:
    """All atom pairs with canonical squared distance < rlist^2 by O(N^2)
    enumeration, as sorted int64 keys (min << 32 | max)."""
    lib = _lib(precision)
    dt = np.float32 if precision == "f32" else np.float64
    xr = np.ascontiguousarray(x, dt).reshape(-1)
    b = np.ascontiguousarray(box, dt)
    r2 = float(dt(rlist) * dt(rlist))
    n = lib.orc_brute_pairs(len(xr) // 3, _p(xr), _p(b), r2, 0, 0)
    out = np.zeros(max(n, 1), np.int64)
    lib.orc_brute_pairs(len(xr) // 3, _p(xr), _p(b), r2, _p(out), n)
    return np.sort(out[:n])


def brute_forces(x, box, types, q, nbfp, ntype, excl_off, excl_idx, coulomb: str, rc: float, beta: float,
                 epsfac: float, k_rf: float = 0.0, c_rf: float = 0.0, sh_ewald: float = 0.0):
    """Direct FP64 pair sum with minimum image and exact Ewald closed form.
    Returns (f[N,3], (Vvdw, Vc), virial[3,3])."""
    lib = _lib("f64")
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    n = len(x) // 3
    b = np.ascontiguousarray(box, np.float64)
    t = np.ascontiguousarray(types, np.int32)
    qq = np.ascontiguousarray(q, np.float64)
    fp = np.ascontiguousarray(nbfp, np.float64)
    eo = np.ascontiguousarray(excl_off, np.int32)
    ei = np.ascontiguousarray(excl_idx, np.int32)
    f = np.zeros(3 * n)
    e = np.zeros(2)
    vir = np.zeros(9)
    lib.orc_brute_forces(n, _p(x), _p(b), _p(t), _p(qq), ntype, _p(fp), _p(eo), _p(ei),
                         1 if coulomb == "ewald" else 0, rc, beta, epsfac, k_rf, c_rf, sh_ewald,
                         _p(f), _p(e), _p(vir))
    return f.reshape(n, 3), (float(e[0]), float(e[1])), vir.reshape(3, 3)


# interaction_const_t vdw_modifier values the oracle implements
VDW_MODIFIER = {"potential-shift": 0, "force-switch": 1, "potential-switch": 2}
LJPME_COMB_RULE = {"geometric": 0, "lorentz-berthelot": 1}
VDW_CONSTANT_KEYS = ["rvdw", "rvdw_switch", "fsw_c2d", "fsw_c3d", "fsw_c2r", "fsw_c3r", "sw_c3", "sw_c4", "sw_c5",
                     "lje_coeff2", "lje_coeff6_6", "sh_lj_ewald"]


class Oracle:
    """One nblib-style force calculator on the CPU.

    Mirrors GmxSetupDirector::setupGmxForceCalculator (api/nblib/gmxsetup.cpp:306-325)
    followed by GmxForceCalculator::compute (gmxcalculator.cpp:70-83).
    """

    def __init__(self, precision="f64", coulomb="ewald", ewald_exact=False, rlist=1.0, rcoulomb=None,
                 epsilon_r=1.0, epsilon_rf=1.0, ewald_rtol=1e-5, ewald_potential_shift=False,
                 vdw_modifier="potential-shift", rvdw_switch=0.0, lj_pme=None, ewald_rtol_lj=1e-3, rvdw=None):
        self.precision = precision
        self.lib = _lib(precision)
        self.dt = np.float64 if precision == "f64" else np.float32
        self.coulomb = coulomb
        rc = rlist if rcoulomb is None else rcoulomb
        self.h = self.lib.orc_create(1 if coulomb == "ewald" else 0, 1 if ewald_exact else 0, rlist, rc,
                                     epsilon_r, epsilon_rf, ewald_rtol, 1 if ewald_potential_shift else 0)
        self.natoms = 0
        # rvdw: None (= rcoulomb) or the twin cut-off's LJ radius
        if rvdw is not None and self.lib.orc_set_vdw_cutoff(self.h, float(rvdw)) != 0:
            raise ValueError(f"bad LJ cut-off {rvdw} (needs 0 < rvdw <= rcoulomb, Ewald, no LJ-PME)")
        if self.lib.orc_set_vdw_modifier(self.h, VDW_MODIFIER[vdw_modifier], rvdw_switch) != 0:
            raise ValueError(f"bad LJ modifier {vdw_modifier!r} / rvdw_switch {rvdw_switch}")
        # lj_pme: None (LJ cut-off) or the grid's combination rule
        if lj_pme is not None:
            if self.lib.orc_set_lj_pme(self.h, LJPME_COMB_RULE[lj_pme], ewald_rtol_lj) != 0:
                raise ValueError(f"bad LJ-PME set-up {lj_pme!r} / rtol {ewald_rtol_lj}")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.orc_destroy(h)
            self.h = None

    def lj_pme_beta(self) -> float:
        """ewaldcoeff_lj (0 without LJ-PME)."""
        return float(self.lib.orc_lj_pme_beta(self.h))

    def constants(self) -> dict:
        out = np.zeros(9)
        self.lib.orc_get_constants(self.h, _p(out))
        keys = ["beta", "tab_scale", "tab_n", "epsfac", "k_rf", "c_rf", "sh_ewald", "cpot_disp", "cpot_rep"]
        return dict(zip(keys, out.tolist()))

    def vdw_constants(self) -> dict:
        out = np.zeros(12)
        self.lib.orc_get_vdw_constants(self.h, _p(out))
        return dict(zip(VDW_CONSTANT_KEYS, out.tolist()))

    def tables(self):
        n = int(self.constants()["tab_n"])
        F = np.zeros(n, self.dt)
        V = np.zeros(n, self.dt)
        FDV0 = np.zeros(4 * n, self.dt)
        self.lib.orc_get_tables(self.h, _p(F), _p(V), _p(FDV0))
        return F, V, FDV0

    def setup(self, box, x, types, charges, nbfp, ntype, excl_off, excl_idx):
        """setNonBondedParameters, setParticlesOnGrid, constructPairList,
        setAtomProperties in nblib's order (gmxsetup.cpp:306-325)."""
        x = np.ascontiguousarray(x, self.dt).reshape(-1)
        n = len(x) // 3
        self.natoms = n
        nb = np.ascontiguousarray(nbfp, self.dt)
        self.lib.orc_set_nbfp(self.h, ntype, _p(nb))
        eo = np.ascontiguousarray(excl_off, np.int32)
        ei = np.ascontiguousarray(excl_idx, np.int32)
        if len(ei) == 0:
            ei = np.zeros(1, np.int32)
        self.lib.orc_set_exclusions(self.h, n, _p(eo), _p(ei))
        self.box = np.ascontiguousarray(box, self.dt)
        rc = self.lib.orc_put_on_grid(self.h, _p(self.box), n, _p(x))
        if rc != 0:
            raise ValueError(f"put_on_grid failed ({rc}): atoms must lie inside the box")
        self.nci = self.lib.orc_construct_pairlist(self.h)
        t = np.ascontiguousarray(types, np.int32)
        q = np.ascontiguousarray(charges, self.dt)
        self.lib.orc_set_atom_props(self.h, _p(t), _p(q))

    def grid(self) -> dict:
        dims = np.zeros(3, np.int32)
        self.lib.orc_get_grid_dims(self.h, _p(dims))
        ncx, ncy, ncells = (int(v) for v in dims)
        ai = np.zeros(ncells * 4, np.int32)
        cells = np.zeros(self.natoms, np.int32)
        cxy = np.zeros(ncx * ncy + 1, np.int32)
        bb = np.zeros(ncells * 6, np.float32)
        xnb = np.zeros(ncells * 12, self.dt)
        self.lib.orc_get_grid(self.h, _p(ai), _p(cells), _p(cxy), _p(bb), _p(xnb))
        return dict(ncx=ncx, ncy=ncy, ncells=ncells, atomIndices=ai, cells=cells, cxy_ind=cxy,
                    bb=bb.reshape(-1, 6), xnb=xnb.reshape(-1, 3))

    def pairlist(self):
        nci = self.lib.orc_num_ci(self.h)
        ncj = self.lib.orc_num_cj(self.h)
        ci = np.zeros(4 * max(nci, 1), np.int32)
        cj = np.zeros(max(ncj, 1), np.int32)
        ex = np.zeros(max(ncj, 1), np.uint32)
        self.lib.orc_get_list(self.h, _p(ci), _p(cj), _p(ex))
        return ci[:4 * nci].reshape(-1, 4), cj[:ncj], ex[:ncj]

    def prune(self, r_inner: float, x=None) -> int:
        """nbnxn_kernel_prune_ref (kernel_ref_prune.cpp:39-123): prune the
        list as searched to r_inner with the current coordinates (x, atom
        order, converted first when given). Returns the i-entry count."""
        if x is not None:
            xx = np.ascontiguousarray(x, self.dt).reshape(-1)
            self.lib.orc_convert_coordinates(self.h, _p(xx))
        return int(self.lib.orc_prune_pairlist(self.h, float(r_inner)))

    def shift_vectors(self) -> np.ndarray:
        out = np.zeros(45 * 3, self.dt)
        self.lib.orc_get_shiftvec(self.h, _p(out))
        return out.reshape(45, 3)

    def list_pairs(self, r: float) -> np.ndarray:
        """Atom pairs of the cluster list with canonical d^2 < r^2 (sorted keys)."""
        r2 = float(self.dt(r) * self.dt(r))
        n = self.lib.orc_list_pairs(self.h, r2, 0, 0)
        out = np.zeros(max(n, 1), np.int64)
        self.lib.orc_list_pairs(self.h, r2, _p(out), n)
        return np.sort(out[:n])

    def compute_threads(self, x, nthreads: int, energy=True):
        """The CPU baseline (precision "bench_f32" only): the kernel over
        the list with nthreads OpenMP threads and per-thread outputs reduced
        afterwards.  Returns dict(f, Vvdw, Vc)."""
        if self.precision != "bench_f32":
            raise ValueError("compute_threads needs the bench_f32 build")
        x = np.ascontiguousarray(x, self.dt).reshape(-1)
        f = np.zeros_like(x)
        e = np.zeros(2)
        self.lib.orc_compute_threads(self.h, _p(x), _p(f), 1 if energy else 0, int(nthreads), _p(e))
        return dict(f=f.reshape(-1, 3), Vvdw=float(e[0]), Vc=float(e[1]))

    def simd_available(self) -> bool:
        """Whether compute_simd runs here (precision "bench_f32", an AVX-512
        host, Ewald with potential-shifted LJ and geometric parameters)."""
        return self.precision == "bench_f32" and bool(self.lib.orc_baseline_simd_available(self.h))

    def compute_simd(self, x, nthreads: int, energy=True):
        """The AVX-512 CPU baseline (the reference's SIMD kernel scheme with
        the analytical Ewald correction), nthreads OpenMP threads.  Returns
        dict(f, Vvdw, Vc), or None where it does not apply."""
        if self.precision != "bench_f32":
            raise ValueError("compute_simd needs the bench_f32 build")
        x = np.ascontiguousarray(x, self.dt).reshape(-1)
        f = np.zeros_like(x)
        e = np.zeros(2)
        if not self.lib.orc_compute_simd(self.h, _p(x), _p(f), 1 if energy else 0, int(nthreads), _p(e)):
            return None
        return dict(f=f.reshape(-1, 3), Vvdw=float(e[0]), Vc=float(e[1]))

    def compute(self, x, energy=True):
        """GmxForceCalculator::compute plus energies, shift forces and virial.
        Returns dict(f[N,3], Vvdw, Vc, fshift[45,3], virial[3,3])."""
        x = np.ascontiguousarray(x, self.dt).reshape(-1)
        f = np.zeros_like(x)
        e = np.zeros(2)
        fsh = np.zeros(45 * 3, self.dt)
        vir = np.zeros(9)
        self.lib.orc_compute(self.h, _p(x), _p(f), 1 if energy else 0, _p(e), _p(fsh), _p(vir))
        return dict(f=f.reshape(-1, 3), Vvdw=float(e[0]), Vc=float(e[1]), fshift=fsh.reshape(45, 3),
                    virial=vir.reshape(3, 3))
