"""Seeded synthetic systems for the nonbonded path.

These are the inputs the benchmark and the parity tests run on.  They play
the role of the reference's benchmark system (src/gromacs/nbnxm/benchmark/
bench_system.cpp:67-202: a rigid 3-site water box with SPC/E charges and LJ
on oxygen only), generated here from a seed instead of a stored 1000-molecule
box so any size can be built reproducibly.

Every builder returns a :class:`System` with coordinates inside the box
[0, L)^3 (nblib requires atoms on the grid, gmxcalculator.cpp:85-103),
per-atom type and charge, per-type C6/C12, and the exclusions as a CSR list
that includes each atom itself, as nblib's topology does
(api/nblib/molecules.cpp:86-88, topology.cpp:60-101).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# SPC/E-like water used by the reference benchmark (bench_system.cpp:67-82)
Q_OXYGEN = -0.8476
Q_HYDROGEN = 0.4238
SIGMA_OW = 0.3166
EPSILON_OW = 0.650
R_OH = 0.1
ANGLE_HOH_DEG = 109.47
WATER_SEED = 20261202


@dataclasses.dataclass
class System:
    """A periodic rectangular system in nblib's terms."""

    name: str
    box: np.ndarray  # (3,) float64, edge lengths in nm
    x: np.ndarray  # (N, 3) float64, inside [0, L)
    types: np.ndarray  # (N,) int32, 0 .. ntypes-1
    charges: np.ndarray  # (N,) float64, e
    c6: np.ndarray  # (ntypes,) float64, kJ mol^-1 nm^6
    c12: np.ndarray  # (ntypes,) float64, kJ mol^-1 nm^12
    excl_off: np.ndarray  # (N+1,) int32
    excl_idx: np.ndarray  # (nnz,) int32, includes self
    rcoulomb: float = 1.0
    rlist: float = 1.0

    @property
    def natoms(self) -> int:
        return len(self.x)

    @property
    def ntypes(self) -> int:
        return len(self.c6)

    def c6_matrix(self) -> np.ndarray:
        """Geometric combination, nblib's only rule (api/nblib/interactions.cpp:83-93)."""
        return np.sqrt(np.outer(self.c6, self.c6))

    def c12_matrix(self) -> np.ndarray:
        return np.sqrt(np.outer(self.c12, self.c12))

    def useful_pairs_estimate(self) -> float:
        """The reference benchmark's "useful pairs" estimate
        N/2 * (rho * 4/3 pi rc^3 + 1) (bench_setup.cpp:273-276)."""
        rho = self.natoms / float(np.prod(self.box))
        return self.natoms / 2.0 * (rho * 4.0 / 3.0 * math.pi * self.rcoulomb ** 3 + 1.0)


def _csr(groups: list[list[int]], natoms: int) -> tuple[np.ndarray, np.ndarray]:
    off = np.zeros(natoms + 1, np.int32)
    for a, g in enumerate(groups):
        off[a + 1] = off[a] + len(g)
    idx = np.zeros(off[-1], np.int32)
    for a, g in enumerate(groups):
        idx[off[a]:off[a + 1]] = g
    return off, idx


def _rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices from normalised Gaussian quaternions."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    r = np.empty((n, 3, 3))
    r[:, 0, 0] = 1 - 2 * (y * y + z * z)
    r[:, 0, 1] = 2 * (x * y - z * w)
    r[:, 0, 2] = 2 * (x * z + y * w)
    r[:, 1, 0] = 2 * (x * y + z * w)
    r[:, 1, 1] = 1 - 2 * (x * x + z * z)
    r[:, 1, 2] = 2 * (y * z - x * w)
    r[:, 2, 0] = 2 * (x * z - y * w)
    r[:, 2, 1] = 2 * (y * z + x * w)
    r[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return r


def _wrap(x: np.ndarray, box: np.ndarray) -> np.ndarray:
    """Put coordinates in [0, L) as float32 will see them (no value rounds to L)."""
    x = x - np.floor(x / box) * box
    x32 = x.astype(np.float32)
    b32 = box.astype(np.float32)
    over = x32 >= b32
    x32 = np.where(over, x32 - b32, x32)
    x32 = np.where(x32 < 0, np.float32(0), x32)
    return x32.astype(np.float64)


def water_box(n_molecules: int, box_length: float | None = None, seed: int = WATER_SEED,
              jitter: float = 0.05, rc: float = 1.0, rlist: float | None = None,
              density: float = 33.4) -> System:
    """Rigid 3-site water on a jittered cubic lattice.

    Sites: m^3 lattice with m = ceil(n^(1/3)) and spacing L/m, the first n
    sites in x-major order (x slowest), each jittered uniformly by +-jitter
    per axis; orientations are uniform random rotations.  O-H = 0.1 nm,
    H-O-H = 109.47 deg; q_O = -0.8476, q_H = +0.4238; LJ on oxygen only
    (sigma 0.3166 nm, epsilon 0.650 kJ/mol); intramolecular pairs excluded.
    Default box: cubic at `density` molecules/nm^3.
    """
    if box_length is None:
        box_length = (n_molecules / density) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    m = int(math.ceil(n_molecules ** (1.0 / 3.0) - 1e-9))
    while m ** 3 < n_molecules:
        m += 1
    a = box_length / m
    ix, iy, iz = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    sites = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)[:n_molecules]
    centers = (sites + 0.5) * a
    centers = centers + rng.uniform(-jitter, jitter, size=(n_molecules, 3))
    rot = _rotations(rng, n_molecules)
    half = math.radians(ANGLE_HOH_DEG) / 2
    local = np.array([[0.0, 0.0, 0.0],
                      [R_OH * math.sin(half), 0.0, R_OH * math.cos(half)],
                      [-R_OH * math.sin(half), 0.0, R_OH * math.cos(half)]])
    xyz = centers[:, None, :] + np.einsum("nij,kj->nki", rot, local)
    box = np.array([box_length] * 3)
    x = _wrap(xyz.reshape(-1, 3), box)
    n = 3 * n_molecules
    types = np.tile(np.array([0, 1, 1], np.int32), n_molecules)
    charges = np.tile(np.array([Q_OXYGEN, Q_HYDROGEN, Q_HYDROGEN]), n_molecules)
    c6 = np.array([4 * EPSILON_OW * SIGMA_OW ** 6, 0.0])
    c12 = np.array([4 * EPSILON_OW * SIGMA_OW ** 12, 0.0])
    groups = []
    for mol in range(n_molecules):
        b = 3 * mol
        for k in range(3):
            groups.append([b + k] + [b + j for j in range(3) if j != k])
    off, idx = _csr(groups, n)
    return System(f"water{n_molecules}", box, x, types, charges, c6, c12, off, idx, rc,
                  rc if rlist is None else rlist)


def hetero_chains(natoms: int = 150_000, box_length: float = 11.45, ntypes: int = 20, chain: int = 10,
                  bond: float = 0.15, rc: float = 1.2, rlist: float = 1.25, seed: int = WATER_SEED + 2,
                  min_dist: float = 0.15) -> System:
    """Random bonded chains of `chain` atoms packed in a cubic box.

    Types: sigma ~ U[0.25, 0.40] nm, epsilon ~ logU[0.05, 1.0] kJ/mol;
    charges ~ U[-0.8, 0.8] e shifted to make each chain neutral; types drawn
    uniformly per atom.  Each chain is a random walk with bond length `bond`;
    steps are rejected when the new atom would come closer than `min_dist` to
    any earlier atom other than its bonded predecessor (a cell-list check),
    which keeps the LJ forces finite.  A chain whose walk finds no free step
    in 1,000 tries is removed and started again from a new random head, so
    no two atoms other than bonded neighbours are ever closer than
    `min_dist`.  1-2 and 1-3 pairs are excluded.
    """
    if natoms % chain:
        raise ValueError("natoms must be a multiple of the chain length")
    rng = np.random.default_rng(seed)
    sigma = rng.uniform(0.25, 0.40, ntypes)
    eps = np.exp(rng.uniform(math.log(0.05), math.log(1.0), ntypes))
    c6 = 4 * eps * sigma ** 6
    c12 = 4 * eps * sigma ** 12
    types = rng.integers(0, ntypes, natoms).astype(np.int32)
    q = rng.uniform(-0.8, 0.8, natoms)
    q = (q.reshape(-1, chain) - q.reshape(-1, chain).mean(axis=1, keepdims=True)).ravel()
    L = box_length
    ncell = max(1, int(L / min_dist))
    cs = L / ncell
    head = -np.ones(ncell ** 3, np.int64)
    nxt = -np.ones(natoms, np.int64)
    x = np.zeros((natoms, 3))
    mind2 = min_dist * min_dist

    def cell_of(p):
        c = np.floor(p / cs).astype(np.int64) % ncell
        return c

    def clashes(p, skip):
        c = cell_of(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    k = (((c[0] + dx) % ncell) * ncell + (c[1] + dy) % ncell) * ncell + (c[2] + dz) % ncell
                    a = head[k]
                    while a >= 0:
                        if a != skip:
                            d = p - x[a]
                            d -= L * np.round(d / L)
                            if d @ d < mind2:
                                return True
                        a = nxt[a]
        return False

    def cell_key(p):
        c = cell_of(p)
        return (c[0] * ncell + c[1]) * ncell + c[2]

    def insert(a, p):
        x[a] = p
        k = cell_key(p)
        nxt[a] = head[k]
        head[k] = a

    def remove_last(a):
        # atoms are inserted at the head of their cell list and only the
        # chain being grown is removed, newest first, so `a` is its cell's head
        k = cell_key(x[a])
        head[k] = nxt[a]
        nxt[a] = -1

    def grow(a0):
        """Grow one chain from atom a0; False when a step finds no free
        position (the caller removes the chain and starts it elsewhere)."""
        while True:
            p = rng.uniform(0, L, 3)
            if not clashes(p, -1):
                break
        insert(a0, p)
        for a in range(a0 + 1, a0 + chain):
            for _ in range(1000):
                v = rng.normal(size=3)
                p = x[a - 1] + bond * v / np.linalg.norm(v)
                p -= L * np.floor(p / L)
                if not clashes(p, a - 1):
                    insert(a, p)
                    break
            else:
                for b in range(a - 1, a0 - 1, -1):
                    remove_last(b)
                return False
        return True

    nchain = natoms // chain
    for ch in range(nchain):
        while not grow(ch * chain):
            pass
    box = np.array([L, L, L])
    x = _wrap(x, box)
    groups = []
    for i in range(natoms):
        c0 = (i // chain) * chain
        g = [i] + [j for j in range(max(c0, i - 2), min(c0 + chain, i + 3)) if j != i]
        groups.append(g)
    off, idx = _csr(groups, natoms)
    return System(f"chains{natoms}", box, x, types, q, c6, c12, off, idx, rc, rlist)


def small_water_box(n_molecules: int = 216, seed: int = WATER_SEED + 7, box_length: float | None = None) -> System:
    """A small water box for tests (default density 33.4 /nm^3)."""
    L = box_length if box_length is not None else max((n_molecules / 33.4) ** (1 / 3), 2.05)
    return water_box(n_molecules, L, seed=seed)


# A displacement (dx, dy, 0) whose float32 squared length rounds to exactly
# 1.25^2 when each product and sum is rounded (fl(fl(dx^2) + fl(dy^2)) =
# 1.5625) but to just below it when either square is fused into the sum
# (fma(dx, dx, fl(dy^2)) and fma(dy, dy, fl(dx^2)) both give 1.5624999).
BOUNDARY_DX = 0.6729803085327148
BOUNDARY_DY = 1.0533742904663086


def boundary_pair_system(seed: int = 17) -> System:
    """A pair at exactly rlist = 1.25 nm in float32 among 124 lattice atoms
    in a 5 nm box (rc 1.2): the pair is outside the list by the rounding of
    the reference's distance test, and inside it if the multiply-adds are
    fused.  Atoms 0 and 1 are the pair; both coordinates and their
    difference are exact in float32."""
    rng = np.random.default_rng(seed)
    L = 5.0
    g = np.arange(5) + 0.5
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    lat = lat[np.linalg.norm(lat - 2.5, axis=1) > 0.6]  # keep the middle free
    lat = lat + rng.uniform(-0.05, 0.05, lat.shape)
    pair = np.array([[2.0, 2.0, 2.0], [2.0 + BOUNDARY_DX, 2.0 + BOUNDARY_DY, 2.0]])
    x = np.concatenate([pair, lat])
    n = len(x)
    types = np.zeros(n, np.int32)
    charges = np.zeros(n)
    c6 = np.array([4 * EPSILON_OW * SIGMA_OW ** 6])
    c12 = np.array([4 * EPSILON_OW * SIGMA_OW ** 12])
    off, idx = _csr([[a] for a in range(n)], n)
    return System("boundary_pair", np.full(3, L), x, types, charges, c6, c12, off, idx, 1.2, 1.25)


def ensemble(n_replicas: int = 64, seed: int = WATER_SEED + 3, min_molecules: int = 1000,
             max_molecules: int = 16000, density: float = 33.4) -> list[System]:
    """The batched-ensemble workload: water boxes built like the 3,000-water
    box, with molecule counts drawn log-uniformly in [min, max] (3k-48k
    atoms), each box cubic at `density` molecules/nm^3; replica r uses seed
    seed + 1000 r."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(min_molecules), math.log(max_molecules), size=n_replicas))
    counts = np.round(counts).astype(int)
    return [water_box(int(n), None, seed=seed + 1000 * r, density=density) for r, n in enumerate(counts)]


def config(index: int) -> System:
    """The benchmark workloads: 0 = 3,000 waters in 4.48 nm, 1 = 96,000
    waters in 14.2 nm, 2 = 150,000-atom chains in 11.45 nm, 4 = the
    force-matching system, 24,000 atoms of chains with 16 LJ types in
    6.214 nm (3 is the batched ensemble, see `ensemble`)."""
    if index == 0:
        return water_box(3000, 4.48, seed=WATER_SEED + 0)
    if index == 1:
        return water_box(96000, 14.2, seed=WATER_SEED + 1)
    if index == 2:
        return hetero_chains()
    if index == 4:
        return hetero_chains(24_000, 6.214, ntypes=16, seed=WATER_SEED + 4)
    raise ValueError(f"unknown config {index}")
