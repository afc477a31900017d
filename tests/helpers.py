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
               rlist=None, rcoulomb=None, vdw=("potential-shift", 0.0), lj_pme=None):
    """An oracle calculator set up on `system` the way nblib sets one up;
    vdw = (LJ modifier, rvdw_switch); lj_pme = None or LJ-PME's grid rule."""
    o = O.Oracle(precision, coulomb, ewald_exact=ewald_exact,
                 rlist=system.rlist if rlist is None else rlist,
                 rcoulomb=system.rcoulomb if rcoulomb is None else rcoulomb,
                 ewald_potential_shift=potential_shift, vdw_modifier=vdw[0], rvdw_switch=vdw[1], lj_pme=lj_pme)
    nb = O.nblib_nbfp(system.c6, system.c12, precision)
    o.setup(system.box, system.x, system.types, system.charges, nb, system.ntypes, system.excl_off,
            system.excl_idx)
    return o


def nbfp_f64(system):
    return O.nblib_nbfp(system.c6, system.c12, "f64")


def rms(a):
    a = np.asarray(a, np.float64)
    return float(np.sqrt(np.mean(np.sum(a * a, axis=-1))))
