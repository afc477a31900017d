"""The batched ensemble (nbx_batch_*) on the GPU: one launch over several
independent water boxes gives each system the forces and energies of its own
single-system computation; a captured HIP graph replays to the eager result;
the autotuner's choice leaves the results unchanged."""
import numpy as np
import pytest

from gmxapi_amd import nbx, systems as S
from tests.helpers import rms

pytestmark = pytest.mark.gpu

FLAGS = nbx.COMPUTE_FORCES | nbx.COMPUTE_ENERGY


@pytest.fixture(scope="module")
def replicas():
    return S.ensemble(4, seed=777, min_molecules=400, max_molecules=1500)


def _device_arrays(replicas):
    import torch

    x = np.concatenate([s.x.astype(np.float32).reshape(-1) for s in replicas])
    xd = torch.from_numpy(x).cuda()
    return xd, torch.zeros_like(xd)


def _split(f, replicas):
    out, o = [], 0
    for s in replicas:
        out.append(f[o:o + 3 * s.natoms].reshape(-1, 3))
        o += 3 * s.natoms
    return out


def _single_results(replicas, **options):
    res = []
    for s in replicas:
        ctx = nbx.context_for(s, **options)
        f = ctx.compute(s.x, energy=True)
        res.append((f, ctx.energies()))
    return res


def test_batch_matches_single_systems(replicas):
    import torch

    print("[batch] 4 replicas", [s.natoms for s in replicas], flush=True)
    b = nbx.batch_for(replicas)
    xd, fd = _device_arrays(replicas)
    b.compute_device(xd.data_ptr(), fd.data_ptr(), FLAGS)
    torch.cuda.synchronize()
    e = b.energies()
    fs = _split(fd.cpu().numpy(), replicas)
    for k, (f_ref, e_ref) in enumerate(_single_results(replicas)):
        assert rms(fs[k] - f_ref) <= 1e-6 * rms(f_ref)
        np.testing.assert_allclose(e[k], e_ref, rtol=1e-10)


@pytest.mark.parametrize("opts", [{"vdw_modifier": "force-switch", "rvdw_switch": 0.7},
                                  {"vdw_modifier": "potential-switch", "rvdw_switch": 0.7},
                                  {"vdw_type": "pme", "ljpme_comb_rule": "lorentz-berthelot"},
                                  {"vdw_cutoff": 0.7}],
                         ids=["force-switch", "potential-switch", "lj-pme", "twin-cutoff"])
def test_batch_with_lj_flavours_matches_single_systems(replicas, opts):
    # the batched launch takes the other LJ flavours like a single context
    # (LJ-PME's self energies included per system)
    import torch

    b = nbx.batch_for(replicas, **opts)
    xd, fd = _device_arrays(replicas)
    b.compute_device(xd.data_ptr(), fd.data_ptr(), FLAGS)
    torch.cuda.synchronize()
    e = b.energies()
    fs = _split(fd.cpu().numpy(), replicas)
    plain = _single_results(replicas)
    for k, (f_ref, e_ref) in enumerate(_single_results(replicas, **opts)):
        assert rms(fs[k] - f_ref) <= 1e-6 * rms(f_ref)
        np.testing.assert_allclose(e[k], e_ref, rtol=1e-10)
        # the flavour changes the LJ energy
        assert abs(e_ref[0] - plain[k][1][0]) > 1e-4 * abs(plain[k][1][0])


def test_graph_replays_match_eager(replicas):
    import torch

    b = nbx.batch_for(replicas)
    stream = torch.cuda.Stream()
    b.set_stream(stream.cuda_stream)
    xd, fd = _device_arrays(replicas)
    with torch.cuda.stream(stream):
        b.compute_device(xd.data_ptr(), fd.data_ptr(), FLAGS)
        stream.synchronize()
        f_eager = fd.cpu().numpy().copy()
        e_eager = b.energies().copy()
        b.capture_graph(xd.data_ptr(), fd.data_ptr(), FLAGS)
        stream.synchronize()
        for k in range(3):
            fd.zero_()
            stream.synchronize()
            b.launch_graph()
            stream.synchronize()
            assert rms(fd.cpu().numpy() - f_eager) <= 1e-6 * rms(f_eager), k
            np.testing.assert_allclose(b.energies(), e_eager, rtol=1e-12)
            # eager work between the replays
            b.compute_device(xd.data_ptr(), fd.data_ptr(), FLAGS)
            stream.synchronize()


def test_autotune_keeps_results(replicas):
    import torch

    b = nbx.batch_for(replicas)
    xd, fd = _device_arrays(replicas)
    b.compute_device(xd.data_ptr(), fd.data_ptr(), FLAGS)
    torch.cuda.synchronize()
    f0, e0 = fd.cpu().numpy().copy(), b.energies().copy()
    choice = b.autotune(xd.data_ptr(), fd.data_ptr(), FLAGS, repeats=2)
    print("[batch] autotune", choice, b.info(), flush=True)
    assert choice["group_size"] in (1, 2, 4, 8, 16) and choice["blocks_per_cu"] >= 1
    b.compute_device(xd.data_ptr(), fd.data_ptr(), FLAGS)
    torch.cuda.synchronize()
    assert rms(fd.cpu().numpy() - f0) <= 1e-6 * rms(f0)
    np.testing.assert_allclose(b.energies(), e0, rtol=1e-10)


def test_batch_input_errors(replicas):
    import torch

    b = nbx.batch_for(replicas[:1])
    xd, fd = _device_arrays(replicas[:1])
    with pytest.raises(nbx.NbxInputError):
        b.compute_device(xd.data_ptr(), fd.data_ptr(), nbx.COMPUTE_FORCES | nbx.COMPUTE_VIRIAL)
    with pytest.raises(nbx.NbxStateError):
        b.energies()
    torch.cuda.synchronize()


def test_reproducible_flag_gives_identical_forces(replicas):
    # with NBX_COMPUTE_REPRODUCIBLE, j-forces beyond a group's window are
    # summed in fixed point: eager calls and graph replays are bit-identical
    # (the fixed-point sums are cleared by the step's set-up kernel, not a
    # memset node, so the graph replays them too)
    import torch

    flags = FLAGS | nbx.COMPUTE_REPRODUCIBLE
    b = nbx.batch_for(replicas)
    stream = torch.cuda.Stream()
    b.set_stream(stream.cuda_stream)
    xd, fd = _device_arrays(replicas)
    with torch.cuda.stream(stream):
        runs = []
        for _ in range(3):
            b.compute_device(xd.data_ptr(), fd.data_ptr(), flags)
            stream.synchronize()
            runs.append((fd.cpu().numpy().copy(), b.energies().copy()))
        for f, e in runs[1:]:
            np.testing.assert_array_equal(f, runs[0][0])
            np.testing.assert_array_equal(e, runs[0][1])
        b.capture_graph(xd.data_ptr(), fd.data_ptr(), flags)
        stream.synchronize()
        for _ in range(2):
            fd.zero_()
            stream.synchronize()
            b.launch_graph()
            stream.synchronize()
            np.testing.assert_array_equal(fd.cpu().numpy(), runs[0][0])
    fs = _split(runs[0][0], replicas)
    for k, (f_ref, _) in enumerate(_single_results(replicas)):
        assert rms(fs[k] - f_ref) <= 1e-6 * rms(f_ref)
