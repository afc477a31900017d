"""Batched nonbonded forces as an ensemble operation in the shape of gmxapi's
operations.

gmxapi's `function_wrapper` (python_packaging/src/gmxapi/operation.py:2635-2700)
turns a function into an operation factory: keyword inputs, a returned
handle whose ``output`` attribute holds named results that ``result()``
resolves; an input given as a list makes an ensemble whose width is the list
length, and each output then resolves to a list with one entry per member
(`ResultDescription.width`, operation.py:77-98).  gmxapi runs the wrapped
function once per member; `nonbonded_ensemble` computes every member's step
in one batched launch (`nbx.Batch`: concatenated work lists, one static
schedule, optional HIP-graph replay), which is what makes an ensemble of
small systems fill the GPU.

    op = nonbonded_ensemble(system=replicas)             # width len(replicas)
    forces = op.output.forces.result()                   # list of (N_r, 3) arrays
    e_lj = op.output.energy_lj.result()                  # list of floats

A single system (not a list) gives width 1 and plain (non-list) results, as
in gmxapi.  The members must share the LJ type parameters and cut-offs (one
batch holds one parameter set).
"""
from __future__ import annotations

import numpy as np

from . import nbx

OUTPUTS = ("forces", "energy_lj", "energy_coulomb")


class _Result:
    """A named output, resolved on first use (gmxapi's Future.result())."""

    def __init__(self, op: "NonbondedEnsemble", name: str):
        self._op = op
        self._name = name

    def result(self):
        values = self._op._resolve()[self._name]
        return values if self._op.width > 1 else values[0]


class _Outputs:
    def __init__(self, op: "NonbondedEnsemble"):
        self._op = op

    def __getattr__(self, name: str) -> _Result:
        if name not in OUTPUTS:
            raise AttributeError(f"no output {name!r}; outputs are {OUTPUTS}")
        return _Result(self._op, name)


class NonbondedEnsemble:
    """Handle of one `nonbonded_ensemble` operation (gmxapi's OperationHandle
    with ``output`` and ``run()``)."""

    def __init__(self, systems, coordinates=None, energy=True, **options):
        if not systems:
            raise nbx.NbxInputError("an ensemble needs at least one member")
        if coordinates is not None and len(coordinates) != len(systems):
            raise nbx.NbxInputError("one coordinate array per ensemble member")
        nbx.check_batch_members(systems)
        self.width = len(systems)
        self._systems = list(systems)
        self._coordinates = coordinates
        self._energy = bool(energy)
        self._options = options
        self._results = None
        self.output = _Outputs(self)

    def run(self):
        self._resolve()
        return self

    def _resolve(self) -> dict:
        if self._results is not None:
            return self._results
        import torch

        batch = nbx.batch_for(self._systems, **self._options)
        xs = self._coordinates if self._coordinates is not None else [s.x for s in self._systems]
        natoms = [len(s.types) for s in self._systems]
        for x, n in zip(xs, natoms):
            if np.shape(x) != (n, 3):
                raise nbx.NbxInputError(f"coordinates of shape {np.shape(x)} for a member of {n} atoms")
        dev = torch.device("cuda", torch.cuda.current_device())
        stream = torch.cuda.current_stream(dev)
        batch.set_stream(stream.cuda_stream)
        x_all = torch.from_numpy(np.concatenate([np.asarray(x, np.float32).reshape(-1, 3) for x in xs])).to(dev)
        f_all = torch.empty_like(x_all)
        flags = nbx.COMPUTE_FORCES | (nbx.COMPUTE_ENERGY if self._energy else 0)
        batch.compute_device(x_all.data_ptr(), f_all.data_ptr(), flags)
        torch.cuda.synchronize(dev)
        f = f_all.cpu().numpy()
        offs = np.concatenate([[0], np.cumsum(natoms)])
        e = batch.energies() if self._energy else np.full((self.width, 2), np.nan)
        self._results = {
            "forces": [f[offs[r]:offs[r + 1]].copy() for r in range(self.width)],
            "energy_lj": [float(v) for v in e[:, 0]],
            "energy_coulomb": [float(v) for v in e[:, 1]],
        }
        batch.close()
        return self._results


def nonbonded_ensemble(system, coordinates=None, energy: bool = True, **options) -> NonbondedEnsemble:
    """Operation factory: `system` is one gmxapi_amd.systems.System (width 1)
    or a list of them (an ensemble); `coordinates` optionally replaces each
    member's coordinates (same nesting); options as for nbx.Context (Ewald
    correction flavour, kernel tuning)."""
    if isinstance(system, (list, tuple)):
        systems = list(system)
        coords = coordinates
    else:
        systems = [system]
        coords = None if coordinates is None else [coordinates]
    return NonbondedEnsemble(systems, coords, energy=energy, **options)
