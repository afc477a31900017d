"""The seeded synthetic systems (no GPU): the chain generator keeps every
non-bonded pair at least min_dist apart, chains are neutral, and batches and
ensembles refuse members whose parameters differ."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from gmxapi_amd import ensemble, nbx, systems as S


def _excluded(system, pairs):
    """Whether each pair (i, j) is in i's exclusion list."""
    out = np.zeros(len(pairs), bool)
    for k, (i, j) in enumerate(pairs):
        out[k] = j in system.excl_idx[system.excl_off[i]:system.excl_off[i + 1]]
    return out


def _closest_nonexcluded(system, r):
    tree = cKDTree(system.x, boxsize=system.box)
    pairs = tree.query_pairs(r, output_type="ndarray")
    if len(pairs) == 0:
        return pairs
    return pairs[~_excluded(system, pairs)]


@pytest.mark.parametrize("index", [4, 2])
def test_chain_configs_have_no_clashes(index):
    s = S.config(index)
    assert s.natoms == {2: 150_000, 4: 24_000}[index]
    # bonded neighbours sit at exactly 0.15 nm; every other pair at least
    # min_dist = 0.15 nm apart (the wrap to float32 moves atoms by < 1e-6 nm)
    close = _closest_nonexcluded(s, 0.15 - 1e-5)
    assert len(close) == 0, f"{len(close)} non-excluded pairs closer than 0.15 nm"
    q = s.charges.reshape(-1, 10).sum(axis=1)
    assert np.max(np.abs(q)) < 1e-12


def test_chain_generator_restarts_blocked_chains():
    # a dense small box forces blocked walks: the restart keeps the distance bound
    s = S.hetero_chains(2_000, box_length=2.8, ntypes=4, seed=99)
    assert len(_closest_nonexcluded(s, 0.15 - 1e-5)) == 0
    v = s.x[1::10] - s.x[0::10]
    v -= s.box * np.round(v / s.box)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 0.15, atol=1e-5)


def test_batch_members_must_share_parameters():
    a = S.small_water_box(216, seed=1)
    b = S.small_water_box(300, seed=2)
    nbx.check_batch_members([a, b])
    c = S.hetero_chains(1_000, box_length=2.2, ntypes=3, seed=5)
    with pytest.raises(nbx.NbxInputError):
        nbx.check_batch_members([a, c])
    d = S.small_water_box(216, seed=3)
    d.rlist = 1.1
    with pytest.raises(nbx.NbxInputError):
        nbx.check_batch_members([a, d])
    with pytest.raises(nbx.NbxInputError):
        ensemble.nonbonded_ensemble(system=[a, c])


def test_lj_matrices_gradients_finite_with_zero_epsilon():
    torch = pytest.importorskip("torch")
    from gmxapi_amd.autograd import lj_matrices

    sigma = torch.tensor([0.3166, 0.1, 0.25], dtype=torch.float64, requires_grad=True)
    eps = torch.tensor([0.65, 0.0, 0.2], dtype=torch.float64, requires_grad=True)
    c6, c12 = lj_matrices(sigma, eps)
    # values: the geometric rule, zero rows for the epsilon-0 type
    c6i = 4 * eps.detach() * sigma.detach() ** 6
    torch.testing.assert_close(c6.detach(), torch.sqrt(torch.outer(c6i, c6i)), rtol=1e-14, atol=0)
    assert float(c6[1].abs().sum()) == 0.0
    w6 = torch.arange(9, dtype=torch.float64).reshape(3, 3)
    (c6 * w6).sum().add((c12 * w6).sum() * 1e4).backward()
    assert torch.isfinite(sigma.grad).all() and torch.isfinite(eps.grad).all()
    assert float(sigma.grad[1]) == 0.0 and float(eps.grad[1]) == 0.0
    # the types with LJ: against central differences of the same function
    h = 1e-7

    def loss(s, e):
        a, b = lj_matrices(s, e)
        return float((a * w6).sum() + (b * w6).sum() * 1e4)

    for k in (0, 2):
        for var, grad in ((sigma, sigma.grad), (eps, eps.grad)):
            p = var.detach().clone()
            m = var.detach().clone()
            p[k] += h * p[k]
            m[k] -= h * m[k]
            args_p = (p, eps.detach()) if var is sigma else (sigma.detach(), p)
            args_m = (m, eps.detach()) if var is sigma else (sigma.detach(), m)
            fd = (loss(*args_p) - loss(*args_m)) / (2 * h * float(var[k]))
            assert abs(float(grad[k]) - fd) <= 1e-6 * abs(fd)
