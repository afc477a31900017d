"""Shared test helpers: build oracle calculators from synthetic systems."""
import json
import os

import numpy as np

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        d = json.load(fh)
    off = [0]
    idx = []
    for g in d["exclusions"]:
        idx += g
        off.append(len(idx))
    d["excl_off"] = np.array(off, np.int32)
    d["excl_idx"] = np.array(idx, np.int32)
    return d


def oracle_for(system, precision="f64", coulomb="ewald", ewald_exact=False, potential_shift=False,
               rlist=None, rcoulomb=None, vdw=("potential-shift", 0.0), lj_pme=None, rvdw=None):
    """An oracle calculator set up on `system` the way nblib sets one up;
    vdw = (LJ modifier, rvdw_switch); lj_pme = None or LJ-PME's grid rule;
    rvdw = None (= rcoulomb) or the twin cut-off's LJ radius."""
    o = O.Oracle(precision, coulomb, ewald_exact=ewald_exact,
                 rlist=system.rlist if rlist is None else rlist,
                 rcoulomb=system.rcoulomb if rcoulomb is None else rcoulomb,
                 ewald_potential_shift=potential_shift, vdw_modifier=vdw[0], rvdw_switch=vdw[1], lj_pme=lj_pme,
                 rvdw=rvdw)
    nb = O.nblib_nbfp(system.c6, system.c12, precision)
    o.setup(system.box, system.x, system.types, system.charges, nb, system.ntypes, system.excl_off,
            system.excl_idx)
    return o


def nbfp_f64(system):
    return O.nblib_nbfp(system.c6, system.c12, "f64")


def rms(a):
    a = np.asarray(a, np.float64)
    return float(np.sqrt(np.mean(np.sum(a * a, axis=-1))))


def parallel_map(fn, items, workers=8):
    """fn over items on a thread pool: the oracle's ctypes calls release the
    GIL, so independent oracle computations run in parallel."""
    from concurrent.futures import ThreadPoolExecutor

    items = list(items)
    if len(items) <= 1:
        return [fn(it) for it in items]
    with ThreadPoolExecutor(max_workers=min(workers, len(items))) as ex:
        return list(ex.map(fn, items))


def translated(system, shift):
    """A copy of `system` rigidly translated by `shift` and wrapped into
    [0, L) as float32 sees it."""
    import dataclasses

    box = np.asarray(system.box, np.float64)
    x = np.mod(system.x + shift, box)
    x[x.astype(np.float32) >= box.astype(np.float32)] = 0.0
    return dataclasses.replace(system, x=x.astype(np.float32).astype(np.float64))


def fp32_energy_noise(system, copies=6, seed=17, workers=8, **oracle_kwargs):
    """RMS deviation from FP64 of the reference FP32 kernel (the oracle's
    single-precision restatement) over rigid translations of the system.

    A translation leaves every energy unchanged in exact arithmetic but moves
    every coordinate's rounding, so these deviations are samples of the FP32
    kernel's rounding noise on this system; one sample alone can land
    arbitrarily close to zero."""
    rng = np.random.default_rng(seed)
    box = np.asarray(system.box, np.float64)
    shifts = [np.zeros(3)] + [rng.uniform(0.0, 1.0, 3) * box for _ in range(copies - 1)]

    def dev(shift):
        t = translated(system, shift)
        r64 = oracle_for(t, "f64", **oracle_kwargs).compute(t.x)
        kw32 = {k: v for k, v in oracle_kwargs.items() if k != "ewald_exact"}
        r32 = oracle_for(t, "f32", **kw32).compute(t.x)
        return r32["Vvdw"] - r64["Vvdw"], r32["Vc"] - r64["Vc"]

    d = np.array(parallel_map(dev, shifts, workers))
    return {"Vvdw": float(np.sqrt(np.mean(d[:, 0] ** 2))), "Vc": float(np.sqrt(np.mean(d[:, 1] ** 2)))}
