"""The gmxapi-shaped ensemble operation (gmxapi_amd.ensemble): width from
list inputs, per-member results (operation.py:77-98, 2635-2700 semantics),
and results equal to computing each member with its own context."""
import numpy as np
import pytest

from gmxapi_amd import ensemble, nbx, systems as S
from tests.helpers import rms


def test_width_and_input_errors():
    # handle construction touches no device: width, output names, errors
    reps = [S.small_water_box(216, seed=5 + k) for k in range(3)]
    op = ensemble.nonbonded_ensemble(system=reps)
    assert op.width == 3
    with pytest.raises(AttributeError):
        op.output.virial
    assert ensemble.nonbonded_ensemble(system=reps[0]).width == 1
    with pytest.raises(nbx.NbxInputError):
        ensemble.nonbonded_ensemble(system=[])
    with pytest.raises(nbx.NbxInputError):
        ensemble.nonbonded_ensemble(system=reps, coordinates=[reps[0].x])


@pytest.mark.gpu
def test_ensemble_results_match_single_systems():
    print("[ensemble] 4 water boxes, one batched launch", flush=True)
    reps = [S.small_water_box(n, seed=30 + n) for n in (216, 300, 400, 512)]
    op = ensemble.nonbonded_ensemble(system=reps)
    forces = op.output.forces.result()
    e_lj = op.output.energy_lj.result()
    e_c = op.output.energy_coulomb.result()
    assert len(forces) == len(e_lj) == len(e_c) == 4
    for r, s in enumerate(reps):
        ctx = nbx.context_for(s)
        f = ctx.compute(s.x, energy=True)
        vlj, vc = ctx.energies()
        assert forces[r].shape == (s.natoms, 3)
        assert rms(forces[r] - f) <= 1e-6 * rms(f)
        assert abs(e_lj[r] - vlj) <= 1e-9 * abs(vlj) + 1e-9 and abs(e_c[r] - vc) <= 1e-9 * abs(vc)
    single = ensemble.nonbonded_ensemble(system=reps[0])
    assert single.output.forces.result().shape == (reps[0].natoms, 3)
    assert isinstance(single.output.energy_lj.result(), float)
