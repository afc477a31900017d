"""Python binding of the libnbx C ABI (include/nbx.h) through ctypes.

This is the layer the gmxapi-style Python code and the tests use.  It does no
arithmetic: every call goes to the HIP library.  If the library has not been
built the import of a Context fails with NbxLibraryMissing, never falling
back to anything else.

Error behaviour follows nblib: invalid input raises NbxInputError (nblib
throws nblib::InputException, api/nblib/exception.h), device failures raise
NbxDeviceError, calls out of order raise NbxStateError.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from . import build as _build

NBX_SUCCESS = 0
NBX_ERROR_INPUT = 1
NBX_ERROR_DEVICE = 2
NBX_ERROR_STATE = 3

COULOMB = {"ewald": 0, "pme": 0, "cutoff": 1, "reaction-field": 2, "rf": 2}
EWALD_CORRECTION = {"table": 0, "analytical": 1}
# LJ modifiers (interaction_const_t::vdw_modifier): nblib's potential shift,
# and the force and potential switches from rvdw_switch to the cut-off
VDW_MODIFIER = {"potential-shift": 0, "force-switch": 1, "potential-switch": 2}
# LJ type (interaction_const_t::vdwtype) and LJ-PME's grid combination rule
VDW_TYPE = {"cutoff": 0, "pme": 1}
LJPME_COMB_RULE = {"geometric": 0, "lorentz-berthelot": 1}

COMPUTE_FORCES = 1
COMPUTE_ENERGY = 2
COMPUTE_SHIFT_FORCES = 4
COMPUTE_VIRIAL = 8
COMPUTE_ADD_FORCES = 16  # add to f instead of overwriting it
COMPUTE_REPRODUCIBLE = 32  # window overflow summed in fixed point: bitwise reproducible in every case


class NbxError(RuntimeError):
    pass


class NbxInputError(NbxError, ValueError):
    """Invalid input (nblib::InputException)."""


class NbxDeviceError(NbxError):
    pass


class NbxStateError(NbxError):
    pass


class NbxLibraryMissing(NbxError, ImportError):
    pass


class _Options(C.Structure):
    _fields_ = [("coulomb_type", C.c_int), ("ewald_correction", C.c_int), ("pairlist_cutoff", C.c_double),
                ("coulomb_cutoff", C.c_double), ("ewald_rtol", C.c_double), ("epsilon_r", C.c_double),
                ("epsilon_rf", C.c_double), ("ewald_potential_shift", C.c_int), ("device", C.c_int),
                ("kernel_group_size", C.c_int), ("kernel_blocks_per_cu", C.c_int), ("vdw_modifier", C.c_int),
                ("rvdw_switch", C.c_double), ("vdw_type", C.c_int), ("ljpme_comb_rule", C.c_int),
                ("ewald_rtol_lj", C.c_double), ("vdw_cutoff", C.c_double)]


def _options(coulomb="ewald", ewald_correction="table", rlist=1.0, rcoulomb=None, ewald_rtol=1e-5, epsilon_r=1.0,
             epsilon_rf=1.0, ewald_potential_shift=False, device=0, kernel_group_size=0, kernel_blocks_per_cu=0,
             vdw_modifier="potential-shift", rvdw_switch=0.0, vdw_type="cutoff", ljpme_comb_rule="geometric",
             ewald_rtol_lj=1e-3, vdw_cutoff=None):
    o = _Options()
    library().nbx_default_options(C.byref(o))
    o.coulomb_type = COULOMB[coulomb]
    o.ewald_correction = EWALD_CORRECTION[ewald_correction]
    o.pairlist_cutoff = rlist
    o.coulomb_cutoff = 0.0 if rcoulomb is None else rcoulomb
    o.ewald_rtol = ewald_rtol
    o.epsilon_r = epsilon_r
    o.epsilon_rf = epsilon_rf
    o.ewald_potential_shift = 1 if ewald_potential_shift else 0
    o.device = device
    o.kernel_group_size = kernel_group_size
    o.kernel_blocks_per_cu = kernel_blocks_per_cu
    if vdw_modifier not in VDW_MODIFIER:
        raise NbxInputError(f"unknown LJ modifier {vdw_modifier!r}; one of {sorted(VDW_MODIFIER)}")
    o.vdw_modifier = VDW_MODIFIER[vdw_modifier]
    o.rvdw_switch = rvdw_switch
    if vdw_type not in VDW_TYPE:
        raise NbxInputError(f"unknown LJ type {vdw_type!r}; one of {sorted(VDW_TYPE)}")
    if ljpme_comb_rule not in LJPME_COMB_RULE:
        raise NbxInputError(f"unknown LJ-PME combination rule {ljpme_comb_rule!r}")
    o.vdw_type = VDW_TYPE[vdw_type]
    o.ljpme_comb_rule = LJPME_COMB_RULE[ljpme_comb_rule]
    o.ewald_rtol_lj = ewald_rtol_lj
    # rvdw: None or 0 means rcoulomb; below rcoulomb is the twin cut-off
    o.vdw_cutoff = 0.0 if vdw_cutoff is None else vdw_cutoff
    return o


class _Info(C.Structure):
    _fields_ = [("num_atoms", C.c_int64), ("num_clusters", C.c_int64), ("num_columns_x", C.c_int64),
                ("num_columns_y", C.c_int64), ("num_i_entries", C.c_int64), ("num_j_entries", C.c_int64)]


_lib = None


def header_symbols() -> list[str]:
    """Every function include/nbx.h declares."""
    hdr = os.path.join(os.path.dirname(_build.HERE), "include", "nbx.h")
    with open(hdr) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(nbx_[a-z0-9_]+)\s*\(", text)))


def _init_torch_device_runtime():
    """PyTorch wheels bring their own HIP runtime.  In one process the two
    runtimes work side by side only when torch's opens the device first (the
    other order leaves torch without devices), so when torch is installed
    and sees a GPU, its runtime is initialised before this library loads."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def library() -> C.CDLL:
    """Load libnbx.so (built in-tree by gmxapi_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    path = _build.library_path()
    if not os.path.exists(path):
        raise NbxLibraryMissing(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    _init_torch_device_runtime()
    lib = C.CDLL(path)
    vp, i, u, d = C.c_void_p, C.c_int, C.c_uint, C.c_double
    lib.nbx_last_error.restype = C.c_char_p
    lib.nbx_default_options.argtypes = [C.POINTER(_Options)]
    lib.nbx_default_options.restype = None
    lib.nbx_create.argtypes = [C.POINTER(_Options), C.POINTER(vp)]
    lib.nbx_destroy.argtypes = [vp]
    lib.nbx_set_stream.argtypes = [vp, vp]
    lib.nbx_set_nonbonded_parameters.argtypes = [vp, i, vp, vp]
    lib.nbx_set_topology.argtypes = [vp, i, vp, vp, vp, vp]
    lib.nbx_put_on_grid.argtypes = [vp, vp, vp, i]
    lib.nbx_construct_pairlist.argtypes = [vp]
    lib.nbx_prune_pairlist.argtypes = [vp, d]
    lib.nbx_prune_rolling.argtypes = [vp, d, i, i]
    lib.nbx_convert_coordinates.argtypes = [vp, vp, i]
    lib.nbx_dispatch_kernel.argtypes = [vp, u]
    lib.nbx_add_nbat_f_to_f.argtypes = [vp, vp, i]
    lib.nbx_compute.argtypes = [vp, vp, vp, u, i]
    lib.nbx_get_energies.argtypes = [vp, vp, vp]
    lib.nbx_get_shift_forces.argtypes = [vp, vp]
    lib.nbx_get_virial.argtypes = [vp, vp]
    lib.nbx_synchronize.argtypes = [vp]
    lib.nbx_get_pairlist_info.argtypes = [vp, C.POINTER(_Info)]
    lib.nbx_get_grid.argtypes = [vp, vp, vp]
    lib.nbx_get_pairlist.argtypes = [vp, vp, vp]
    lib.nbx_list_pairs.argtypes = [vp, d, vp, C.c_int64]
    lib.nbx_list_pairs.restype = C.c_int64
    lib.nbx_get_interaction_constants.argtypes = [vp, vp]
    lib.nbx_get_vdw_constants.argtypes = [vp, vp]
    lib.nbx_get_ewald_table.argtypes = [vp, vp, vp]
    lib.nbx_get_work_groups.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.nbx_set_kernel_timing.argtypes = [vp, C.c_int]
    lib.nbx_batch_create.argtypes = [vp, C.c_int, vp, vp, vp]
    lib.nbx_batch_destroy.argtypes = [vp]
    lib.nbx_batch_set_stream.argtypes = [vp, vp]
    lib.nbx_batch_add_system.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp]
    lib.nbx_batch_construct_pairlists.argtypes = [vp]
    lib.nbx_batch_get_info.argtypes = [vp, vp]
    lib.nbx_batch_compute.argtypes = [vp, vp, vp, C.c_uint]
    lib.nbx_batch_capture_graph.argtypes = [vp, vp, vp, C.c_uint]
    lib.nbx_batch_launch_graph.argtypes = [vp]
    lib.nbx_batch_get_energies.argtypes = [vp, vp]
    lib.nbx_batch_autotune.argtypes = [vp, vp, vp, C.c_uint, C.c_int, vp]
    lib.nbx_batch_get_system.argtypes = [vp, C.c_int, vp]
    lib.nbx_get_kernel_timing.argtypes = [vp, vp, vp]
    lib.nbx_set_charges.argtypes = [vp, vp, i]
    lib.nbx_compute_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, vp, i]
    _lib = lib
    return lib


def _raise(code: int, what: str):
    msg = library().nbx_last_error().decode(errors="replace")
    cls = {NBX_ERROR_INPUT: NbxInputError, NBX_ERROR_STATE: NbxStateError}.get(code, NbxDeviceError)
    raise cls(f"{what}: {msg}")


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Context:
    """One system on one GPU: the nonbonded_verlet_t replacement.

    Methods mirror the calls nblib makes (gmxsetup.cpp:306-325 and
    gmxcalculator.cpp:70-103).
    """

    def __init__(self, coulomb: str = "ewald", ewald_correction: str = "table", rlist: float = 1.0,
                 rcoulomb: float | None = None, ewald_rtol: float = 1e-5, epsilon_r: float = 1.0,
                 epsilon_rf: float = 1.0, ewald_potential_shift: bool = False, device: int = 0,
                 kernel_group_size: int = 0, kernel_blocks_per_cu: int = 0,
                 vdw_modifier: str = "potential-shift", rvdw_switch: float = 0.0, vdw_type: str = "cutoff",
                 ljpme_comb_rule: str = "geometric", ewald_rtol_lj: float = 1e-3, vdw_cutoff: float | None = None):
        lib = library()
        self._lib = lib
        o = _options(coulomb, ewald_correction, rlist, rcoulomb, ewald_rtol, epsilon_r, epsilon_rf,
                     ewald_potential_shift, device, kernel_group_size, kernel_blocks_per_cu, vdw_modifier,
                     rvdw_switch, vdw_type, ljpme_comb_rule, ewald_rtol_lj, vdw_cutoff)
        h = C.c_void_p()
        rc = lib.nbx_create(C.byref(o), C.byref(h))
        if rc != NBX_SUCCESS:
            _raise(rc, "nbx_create")
        self._h = h
        self._owner = None
        self.natoms = 0

    @classmethod
    def _view(cls, handle, owner, natoms):
        """A non-owning Context on a handle that `owner` (a Batch) keeps alive."""
        obj = cls.__new__(cls)
        obj._lib = library()
        obj._h = C.c_void_p(handle)
        obj._owner = owner
        obj.natoms = natoms
        return obj

    def close(self):
        if getattr(self, "_h", None) and getattr(self, "_owner", None) is None:
            self._lib.nbx_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != NBX_SUCCESS:
            _raise(rc, what)

    @property
    def handle(self) -> int:
        return self._h.value

    def set_stream(self, stream_ptr: int | None):
        """Queue work on this hipStream_t (an int handle); 0/None is the null stream."""
        self._check(self._lib.nbx_set_stream(self._h, stream_ptr or None), "nbx_set_stream")

    def set_nonbonded_parameters(self, c6: np.ndarray, c12: np.ndarray):
        c6 = np.ascontiguousarray(c6, np.float64)
        c12 = np.ascontiguousarray(c12, np.float64)
        t = c6.shape[0]
        self._check(self._lib.nbx_set_nonbonded_parameters(self._h, t, _ptr(c6), _ptr(c12)),
                    "nbx_set_nonbonded_parameters")
        self.num_types = t

    def set_topology(self, types, charges, excl_off, excl_idx):
        t = np.ascontiguousarray(types, np.int32)
        q = np.ascontiguousarray(charges, np.float32)
        eo = np.ascontiguousarray(excl_off, np.int32)
        ei = np.ascontiguousarray(excl_idx, np.int32)
        if len(ei) == 0:
            ei = np.zeros(1, np.int32)
        self.natoms = len(t)
        self._check(self._lib.nbx_set_topology(self._h, len(t), _ptr(t), _ptr(q), _ptr(eo), _ptr(ei)),
                    "nbx_set_topology")

    def put_on_grid(self, box, x=None, x_device_ptr: int | None = None):
        b = np.ascontiguousarray(box, np.float32)
        if x_device_ptr is not None:
            self._check(self._lib.nbx_put_on_grid(self._h, _ptr(b), x_device_ptr, 1), "nbx_put_on_grid")
        else:
            xa = np.ascontiguousarray(x, np.float32).reshape(-1)
            self._check(self._lib.nbx_put_on_grid(self._h, _ptr(b), _ptr(xa), 0), "nbx_put_on_grid")

    def construct_pairlist(self):
        self._check(self._lib.nbx_construct_pairlist(self._h), "nbx_construct_pairlist")

    def prune_pairlist(self, r_inner: float):
        """Dynamic pruning (dispatchPruneKernelCpu, nbnxm.h:312-313): the list
        as searched pruned to r_inner with the current coordinates."""
        self._check(self._lib.nbx_prune_pairlist(self._h, float(r_inner)), "nbx_prune_pairlist")

    def prune_rolling(self, r_inner: float, part: int = 0, num_parts: int = 1):
        """Rolling dynamic pruning inside the kernel's work layout
        (nbnxm.h:306-316, pairlistparams.h:136-138): the kernel wavefronts
        of `part` out of `num_parts` get the layout as built pruned to
        r_inner with the current coordinates; asynchronous."""
        self._check(self._lib.nbx_prune_rolling(self._h, float(r_inner), int(part), int(num_parts)),
                    "nbx_prune_rolling")

    def compute(self, x, energy=False, shift_forces=False, virial=False, reproducible=False) -> np.ndarray:
        """Forces in atom order (host arrays, synchronous)."""
        xa = np.ascontiguousarray(x, np.float32).reshape(-1)
        f = np.empty_like(xa)
        flags = COMPUTE_FORCES | (COMPUTE_ENERGY if energy else 0) | (COMPUTE_SHIFT_FORCES if shift_forces else 0)
        flags |= (COMPUTE_VIRIAL if virial else 0) | (COMPUTE_REPRODUCIBLE if reproducible else 0)
        self._check(self._lib.nbx_compute(self._h, _ptr(xa), _ptr(f), flags, 0), "nbx_compute")
        return f.reshape(-1, 3)

    def compute_device(self, x_ptr: int, f_ptr: int, flags: int = COMPUTE_FORCES):
        """Asynchronous compute on device buffers of the context's device."""
        self._check(self._lib.nbx_compute(self._h, x_ptr, f_ptr, flags, 1), "nbx_compute")

    def set_charges(self, charges):
        """New per-atom charges; the grid and pair list stay valid."""
        q = np.ascontiguousarray(charges, np.float32)
        self._check(self._lib.nbx_set_charges(self._h, _ptr(q), 0), "nbx_set_charges")

    def set_charges_device(self, q_ptr: int):
        self._check(self._lib.nbx_set_charges(self._h, q_ptr, 1), "nbx_set_charges")

    def compute_adjoint(self, x, g):
        """dL/dx (N, 3), dL/dq (N,) and dL/dC6, dL/dC12 (T, T) for a
        force cotangent g = dL/dF (host arrays, synchronous)."""
        xa = np.ascontiguousarray(x, np.float32).reshape(-1)
        ga = np.ascontiguousarray(g, np.float32).reshape(-1)
        gx = np.empty_like(xa)
        gq = np.empty(len(xa) // 3, np.float64)
        t = self.num_types
        g6 = np.empty((t, t), np.float64)
        g12 = np.empty((t, t), np.float64)
        self._check(self._lib.nbx_compute_adjoint(self._h, _ptr(xa), _ptr(ga), _ptr(gx), _ptr(gq), _ptr(g6),
                                                  _ptr(g12), 0), "nbx_compute_adjoint")
        return gx.reshape(-1, 3), gq, g6, g12

    def compute_adjoint_device(self, x_ptr: int, g_ptr: int, grad_x_ptr: int | None, grad_q_ptr: int | None,
                               grad_c6_ptr: int | None = None, grad_c12_ptr: int | None = None):
        """Device-pointer adjoint, asynchronous on the context's stream:
        grad_c6/grad_c12 are (T, T) float64 device arrays (or None)."""
        self._check(self._lib.nbx_compute_adjoint(self._h, x_ptr, g_ptr, grad_x_ptr, grad_q_ptr, grad_c6_ptr,
                                                  grad_c12_ptr, 1), "nbx_compute_adjoint")

    def convert_coordinates(self, x):
        """nonbonded_verlet_t::convertCoordinates (nbnxm.cpp:129-141) from a
        host array (N x 3), synchronous."""
        xx = np.ascontiguousarray(x, np.float32).reshape(-1)
        self._check(self._lib.nbx_convert_coordinates(self._h, _ptr(xx), 0), "nbx_convert_coordinates")
        self.synchronize()

    def convert_coordinates_device(self, x_ptr: int):
        self._check(self._lib.nbx_convert_coordinates(self._h, x_ptr, 1), "nbx_convert_coordinates")

    def dispatch_kernel(self, flags: int = COMPUTE_FORCES):
        self._check(self._lib.nbx_dispatch_kernel(self._h, flags), "nbx_dispatch_kernel")

    def add_nbat_f_to_f_device(self, f_ptr: int):
        self._check(self._lib.nbx_add_nbat_f_to_f(self._h, f_ptr, 1), "nbx_add_nbat_f_to_f")

    def synchronize(self):
        self._check(self._lib.nbx_synchronize(self._h), "nbx_synchronize")

    def energies(self) -> tuple[float, float]:
        vlj = C.c_double()
        vc = C.c_double()
        self._check(self._lib.nbx_get_energies(self._h, C.byref(vlj), C.byref(vc)), "nbx_get_energies")
        return vlj.value, vc.value

    def shift_forces(self) -> np.ndarray:
        out = np.zeros(45 * 3)
        self._check(self._lib.nbx_get_shift_forces(self._h, _ptr(out)), "nbx_get_shift_forces")
        return out.reshape(45, 3)

    def virial(self) -> np.ndarray:
        out = np.zeros(9)
        self._check(self._lib.nbx_get_virial(self._h, _ptr(out)), "nbx_get_virial")
        return out.reshape(3, 3)

    def info(self) -> dict:
        inf = _Info()
        self._check(self._lib.nbx_get_pairlist_info(self._h, C.byref(inf)), "nbx_get_pairlist_info")
        return {k: int(getattr(inf, k)) for k, _ in _Info._fields_}

    def grid(self) -> tuple[np.ndarray, np.ndarray]:
        inf = self.info()
        ai = np.zeros(4 * inf["num_clusters"], np.int32)
        cells = np.zeros(inf["num_atoms"], np.int32)
        self._check(self._lib.nbx_get_grid(self._h, _ptr(ai), _ptr(cells)), "nbx_get_grid")
        return ai, cells

    def pairlist(self) -> tuple[np.ndarray, np.ndarray]:
        inf = self.info()
        ie = np.zeros((max(inf["num_i_entries"], 1), 4), np.int32)
        je = np.zeros((max(inf["num_j_entries"], 1), 2), np.uint32)
        self._check(self._lib.nbx_get_pairlist(self._h, _ptr(ie), _ptr(je)), "nbx_get_pairlist")
        return ie[:inf["num_i_entries"]], je[:inf["num_j_entries"]]

    def count_pairs(self, r: float) -> int:
        """Number of atom pairs the list holds within r (canonical rounding)."""
        n = self._lib.nbx_list_pairs(self._h, r, None, 0)
        if n < 0:
            _raise(int(-n), "nbx_list_pairs")
        return int(n)

    def list_pairs(self, r: float) -> np.ndarray:
        n = self._lib.nbx_list_pairs(self._h, r, None, 0)
        if n < 0:
            _raise(int(-n), "nbx_list_pairs")
        out = np.zeros(max(n, 1), np.int64)
        m = self._lib.nbx_list_pairs(self._h, r, _ptr(out), n)
        if m < 0:
            _raise(int(-m), "nbx_list_pairs")
        return np.sort(out[:n])

    def work_groups(self) -> dict:
        """The kernel's work groups: see nbx_get_work_groups in include/nbx.h."""
        info = np.zeros(5, np.int32)
        self._check(self._lib.nbx_get_work_groups(self._h, _ptr(info), None, None, None, None),
                    "nbx_get_work_groups")
        G, ng, nch, slots, nwork = (int(v) for v in info)
        groups = np.zeros((max(ng, 1), 4), np.int32)
        chunks = np.zeros((max(nch, 1), 4), np.int32)
        self._check(self._lib.nbx_get_work_groups(self._h, _ptr(info), _ptr(groups), _ptr(chunks), None, None),
                    "nbx_get_work_groups")
        total = int(groups[ng - 1, 2] + groups[ng - 1, 3]) if ng else 0
        union = np.zeros(max(total, 1), np.int32)
        work = np.zeros((max(nwork, 1), 2), np.uint32)
        self._check(self._lib.nbx_get_work_groups(self._h, _ptr(info), None, None, _ptr(union), _ptr(work)),
                    "nbx_get_work_groups")
        return dict(group_size=G, num_groups=ng, num_chunks=nch, window_slots=slots, groups=groups[:ng],
                    chunks=chunks[:nch], union=union[:total], work=work[:nwork])

    def set_kernel_timing(self, max_launches: int):
        """Record HIP events around the next max_launches pair-kernel launches."""
        self._check(self._lib.nbx_set_kernel_timing(self._h, int(max_launches)), "nbx_set_kernel_timing")

    def kernel_timing(self):
        """(average ms, launches) of the recorded pair-kernel launches."""
        avg = C.c_double(0.0)
        n = C.c_int(0)
        self._check(self._lib.nbx_get_kernel_timing(self._h, C.byref(avg), C.byref(n)),
                    "nbx_get_kernel_timing")
        return avg.value, n.value

    def interaction_constants(self) -> dict:
        out = np.zeros(9)
        self._check(self._lib.nbx_get_interaction_constants(self._h, _ptr(out)), "nbx_get_interaction_constants")
        keys = ["beta", "tab_scale", "tab_n", "epsfac", "k_rf", "c_rf", "sh_ewald", "cpot_disp", "cpot_rep"]
        return dict(zip(keys, out.tolist()))

    def vdw_constants(self) -> dict:
        out = np.zeros(12)
        self._check(self._lib.nbx_get_vdw_constants(self._h, _ptr(out)), "nbx_get_vdw_constants")
        keys = ["rvdw", "rvdw_switch", "fsw_c2d", "fsw_c3d", "fsw_c2r", "fsw_c3r", "sw_c3", "sw_c4", "sw_c5",
                "lje_coeff2", "lje_coeff6_6", "sh_lj_ewald"]
        return dict(zip(keys, out.tolist()))

    def ewald_table(self) -> np.ndarray:
        """The FDV0 quadruplets the kernel reads, (n, 4) float32."""
        n = C.c_int(0)
        self._check(self._lib.nbx_get_ewald_table(self._h, C.byref(n), None), "nbx_get_ewald_table")
        out = np.zeros((max(n.value, 1), 4), np.float32)
        if n.value:
            self._check(self._lib.nbx_get_ewald_table(self._h, C.byref(n), _ptr(out)), "nbx_get_ewald_table")
        return out[:n.value]


def nblib_c6c12_matrices(c6_type, c12_type):
    """Geometric combination as nblib does it in float
    (api/nblib/interactions.cpp:83-93): C_ij = sqrt(C_i C_j)."""
    c6 = np.asarray(c6_type, np.float32)
    c12 = np.asarray(c12_type, np.float32)
    m6 = np.sqrt((c6[:, None] * c6[None, :]).astype(np.float32)).astype(np.float64)
    m12 = np.sqrt((c12[:, None] * c12[None, :]).astype(np.float32)).astype(np.float64)
    return m6, m12


def lorentz_berthelot_c6c12_matrices(sigma_type, epsilon_type):
    """C6/C12 matrices by the Lorentz-Berthelot rule, sigma_ij = (sigma_i +
    sigma_j) / 2 and epsilon_ij = sqrt(epsilon_i epsilon_j), C6 = 4 eps
    sigma^6, C12 = 4 eps sigma^12 (the rule nbnxm recognises as ljcrLB,
    atomdata.cpp:540-590; nblib declares it, kerneloptions.h:66-72, but its
    combineNonbondedParameters throws for it, interactions.cpp:83-93).  The
    pair kernel reads the full matrices, so any rule, or none, works."""
    sig = np.asarray(sigma_type, np.float64)
    eps = np.asarray(epsilon_type, np.float64)
    if sig.shape != eps.shape or sig.ndim != 1:
        raise NbxInputError("one sigma and one epsilon per particle type")
    if np.any(sig < 0) or np.any(eps < 0):
        raise NbxInputError("sigma and epsilon must be non-negative")
    s6 = (0.5 * (sig[:, None] + sig[None, :])) ** 6
    e = np.sqrt(eps[:, None] * eps[None, :])
    return 4.0 * e * s6, 4.0 * e * s6 * s6


class Batch:
    """A batched ensemble (include/nbx.h nbx_batch_*): independent systems
    sharing options and LJ parameters, computed by one kernel launch."""

    def __init__(self, c6_matrix, c12_matrix, **options):
        lib = library()
        self._lib = lib
        o = _options(**options)
        c6 = np.ascontiguousarray(c6_matrix, np.float64)
        c12 = np.ascontiguousarray(c12_matrix, np.float64)
        h = C.c_void_p()
        rc = lib.nbx_batch_create(C.byref(o), c6.shape[0], _ptr(c6), _ptr(c12), C.byref(h))
        if rc != NBX_SUCCESS:
            _raise(rc, "nbx_batch_create")
        self._h = h
        self.natoms = []

    def _check(self, rc, what):
        if rc != NBX_SUCCESS:
            _raise(rc, what)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nbx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_stream(self, stream_ptr: int):
        self._check(self._lib.nbx_batch_set_stream(self._h, C.c_void_p(stream_ptr)), "nbx_batch_set_stream")

    def add_system(self, types, charges, excl_off, excl_idx, box, x) -> int:
        types = np.ascontiguousarray(types, np.int32)
        charges = np.ascontiguousarray(charges, np.float32)
        eo = np.ascontiguousarray(excl_off, np.int32)
        ei = np.ascontiguousarray(excl_idx, np.int32)
        box = np.ascontiguousarray(box, np.float32)
        x = np.ascontiguousarray(x, np.float32).reshape(-1)
        idx = C.c_int(-1)
        self._check(self._lib.nbx_batch_add_system(self._h, len(types), _ptr(types), _ptr(charges), _ptr(eo),
                                                   _ptr(ei), _ptr(box), _ptr(x), C.byref(idx)),
                    "nbx_batch_add_system")
        self.natoms.append(len(types))
        return idx.value

    def construct_pairlists(self):
        self._check(self._lib.nbx_batch_construct_pairlists(self._h), "nbx_batch_construct_pairlists")

    def info(self) -> dict:
        out = np.zeros(7, np.int64)
        self._check(self._lib.nbx_batch_get_info(self._h, _ptr(out)), "nbx_batch_get_info")
        keys = ["num_systems", "num_atoms", "num_clusters", "num_j_entries", "num_groups", "group_size",
                "blocks_per_cu"]
        return dict(zip(keys, (int(v) for v in out)))

    def compute_device(self, x_ptr: int, f_ptr: int, flags: int = COMPUTE_FORCES):
        self._check(self._lib.nbx_batch_compute(self._h, C.c_void_p(x_ptr), C.c_void_p(f_ptr), flags),
                    "nbx_batch_compute")

    def capture_graph(self, x_ptr: int, f_ptr: int, flags: int = COMPUTE_FORCES):
        self._check(self._lib.nbx_batch_capture_graph(self._h, C.c_void_p(x_ptr), C.c_void_p(f_ptr), flags),
                    "nbx_batch_capture_graph")

    def launch_graph(self):
        self._check(self._lib.nbx_batch_launch_graph(self._h), "nbx_batch_launch_graph")

    def energies(self) -> np.ndarray:
        out = np.zeros((len(self.natoms), 2))
        self._check(self._lib.nbx_batch_get_energies(self._h, _ptr(out)), "nbx_batch_get_energies")
        return out

    def autotune(self, x_ptr: int, f_ptr: int, flags: int = COMPUTE_FORCES, repeats: int = 5):
        out = np.zeros(3, np.int32)
        self._check(self._lib.nbx_batch_autotune(self._h, C.c_void_p(x_ptr), C.c_void_p(f_ptr), flags, repeats,
                                                 _ptr(out)), "nbx_batch_autotune")
        return dict(group_size=int(out[0]), blocks_per_cu=int(out[1]), entry_weight=int(out[2]))

    def system(self, index: int) -> "Context":
        h = C.c_void_p()
        self._check(self._lib.nbx_batch_get_system(self._h, index, C.byref(h)), "nbx_batch_get_system")
        return Context._view(h.value, self, self.natoms[index])


def check_batch_members(systems):
    """Raise NbxInputError unless every system shares the first one's C6,
    C12, rlist and rcoulomb (host-side check, no device needed)."""
    systems = list(systems)
    if not systems:
        raise NbxInputError("a batch needs at least one system")
    s0 = systems[0]
    for k, s in enumerate(systems[1:], start=1):
        if (np.shape(s.c6) != np.shape(s0.c6) or np.shape(s.c12) != np.shape(s0.c12)
                or not np.array_equal(s.c6, s0.c6) or not np.array_equal(s.c12, s0.c12)):
            raise NbxInputError(f"batch member {k} has LJ parameters that differ from member 0's")
        if s.rlist != s0.rlist or s.rcoulomb != s0.rcoulomb:
            raise NbxInputError(f"batch member {k} has cut-offs (rlist {s.rlist}, rcoulomb {s.rcoulomb}) that "
                                f"differ from member 0's ({s0.rlist}, {s0.rcoulomb})")


def batch_for(systems, **options) -> Batch:
    """A Batch holding gmxapi_amd.systems.System objects that share LJ
    parameters (set up as nblib would set up each one).  Every member must
    have the first member's per-type C6/C12 and cut-offs: one launch uses one
    parameter matrix and one set of interaction constants."""
    check_batch_members(systems)
    m6, m12 = nblib_c6c12_matrices(systems[0].c6, systems[0].c12)
    b = Batch(m6, m12, rlist=systems[0].rlist, rcoulomb=systems[0].rcoulomb, **options)
    for s in systems:
        b.add_system(s.types, s.charges, s.excl_off, s.excl_idx, s.box, s.x)
    b.construct_pairlists()
    return b


def context_for(system, coulomb="ewald", ewald_correction="table", device=0, potential_shift=False,
                kernel_group_size=0, kernel_blocks_per_cu=0, vdw_modifier="potential-shift",
                rvdw_switch=0.0, vdw_type="cutoff", ljpme_comb_rule="geometric", ewald_rtol_lj=1e-3,
                vdw_cutoff=None) -> Context:
    """Context set up on a gmxapi_amd.systems.System the way nblib sets up its
    calculator: parameters, topology, grid, pair list."""
    ctx = Context(coulomb=coulomb, ewald_correction=ewald_correction, rlist=system.rlist,
                  rcoulomb=system.rcoulomb, ewald_potential_shift=potential_shift, device=device,
                  kernel_group_size=kernel_group_size, kernel_blocks_per_cu=kernel_blocks_per_cu,
                  vdw_modifier=vdw_modifier, rvdw_switch=rvdw_switch, vdw_type=vdw_type,
                  ljpme_comb_rule=ljpme_comb_rule, ewald_rtol_lj=ewald_rtol_lj, vdw_cutoff=vdw_cutoff)
    m6, m12 = nblib_c6c12_matrices(system.c6, system.c12)
    ctx.set_nonbonded_parameters(m6, m12)
    ctx.set_topology(system.types, system.charges, system.excl_off, system.excl_idx)
    ctx.put_on_grid(system.box, system.x)
    ctx.construct_pairlist()
    return ctx
