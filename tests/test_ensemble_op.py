"""The gmxapi-shaped ensemble operation (gmxapi_amd.ensemble): width from
list inputs, per-member results (operation.py:77-98, 2635-2700 semantics),
results against the FP64 oracle per member and equal to computing each
member with its own context."""
import numpy as np
import pytest

from gmxapi_amd import ensemble, nbx, systems as S
from tests.helpers import fp32_energy_noise, oracle_for, rms


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
        # the FP64 oracle: forces to the path's 5e-4 RMS bar, energies to
        # 1e-6 or three times the FP32 reference kernel's rounding noise on
        # these small boxes
        r64 = oracle_for(s, "f64").compute(s.x)
        assert forces[r].shape == (s.natoms, 3)
        assert rms(forces[r] - r64["f"]) <= 5e-4 * rms(r64["f"])
        noise = fp32_energy_noise(s, copies=6)
        scale = abs(r64["Vvdw"]) + abs(r64["Vc"])
        for got, key in ((e_lj[r], "Vvdw"), (e_c[r], "Vc")):
            assert abs(got - r64[key]) <= max(1e-6 * scale, 3.0 * noise[key]), (r, key, got, r64[key])
        # and a single context on the same member gives the same results
        ctx = nbx.context_for(s)
        f = ctx.compute(s.x, energy=True)
        vlj, vc = ctx.energies()
        assert rms(forces[r] - f) <= 1e-6 * rms(f)
        assert abs(e_lj[r] - vlj) <= 1e-9 * abs(vlj) + 1e-9 and abs(e_c[r] - vc) <= 1e-9 * abs(vc)
    single = ensemble.nonbonded_ensemble(system=reps[0])
    assert single.output.forces.result().shape == (reps[0].natoms, 3)
    assert isinstance(single.output.energy_lj.result(), float)
