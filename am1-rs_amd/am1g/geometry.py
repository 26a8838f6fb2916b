"""Seeded synthetic molecules for the benchmark workloads and the parity tests.

Every generator is deterministic in its seed (numpy ``PCG64``) and clash-checked: a draw that
puts two non-bonded atoms closer than ``min_contact`` Angstrom is redrawn, and the number of
redraws is returned with the geometry so a result file can record it.

Workloads (sizes as in the benchmark description):
  * water clusters — O on an (nx, ny, nz) grid at 3.0 A spacing with U(-0.2, 0.2) A jitter per
    coordinate, O-H 0.96 A, H-O-H 104.5 deg, a Haar-uniform orientation per molecule;
  * branched saturated chains C_nH_(2n+2), n ~ U{3..20}, one methyl branch with p = 0.3, ideal
    tetrahedral geometry (C-C 1.54, C-H 1.09 A), torsions from {60, 180, 300} deg, and each
    terminal CH3 replaced by OH (C-O 1.43, O-H 0.96) or NH2 (C-N 1.47, N-H 1.01) with p = 0.2.

Coordinates are returned in Angstrom; multiply by ``ANGSTROM_TO_BOHR`` (1/0.529167, the MOPAC7
Bohr radius of am1-rs constants.rs:16-18) for the SCF entry points, which take Bohr.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

ANGSTROM_TO_BOHR = 1.0 / 0.529167
TETRA = math.degrees(math.acos(-1.0 / 3.0))  # 109.4712 deg


@dataclass
class Molecule:
    z: np.ndarray          # int32 [n]
    xyz: np.ndarray        # float64 [n, 3], Angstrom
    charge: float = 0.0
    multiplicity: int = 1
    name: str = ""
    redraws: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def xyz_bohr(self) -> np.ndarray:
        return self.xyz * ANGSTROM_TO_BOHR

    @property
    def n_atoms(self) -> int:
        return len(self.z)

    @property
    def nao(self) -> int:
        return int(sum(1 if zz == 1 else 4 for zz in self.z))


def _haar_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniformly random rotation from a normalised Gaussian quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def _water_template() -> np.ndarray:
    r, ang = 0.96, math.radians(104.5)
    return np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0], [r * math.cos(ang), r * math.sin(ang), 0.0]])


def water_cluster(grid, seed: int, spacing: float = 3.0, jitter: float = 0.2, min_contact: float = 1.6,
                  count: int | None = None, max_redraws: int = 20000) -> Molecule:
    """Water molecules on a grid (the first ``count`` grid points, row-major x-fastest order)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = grid
    sites = [(i, j, k) for k in range(nz) for j in range(ny) for i in range(nx)]
    if count is not None:
        sites = sites[:count]
    tmpl = _water_template()
    mols = []
    redraws = 0
    for (i, j, k) in sites:
        centre = np.array([i, j, k], dtype=float) * spacing + rng.uniform(-jitter, jitter, size=3)
        for attempt in range(max_redraws):
            rot = _haar_rotation(rng)
            cand = centre + tmpl @ rot.T
            ok = True
            for m in mols:
                d = np.linalg.norm(cand[:, None, :] - m[None, :, :], axis=-1)
                if d.min() < min_contact:
                    ok = False
                    break
            if ok:
                break
            redraws += 1
        else:
            raise RuntimeError("water cluster: could not place a molecule without a clash")
        mols.append(cand)
    xyz = np.concatenate(mols, axis=0)
    z = np.tile(np.array([8, 1, 1], dtype=np.int32), len(mols))
    return Molecule(z=z, xyz=xyz, name=f"(H2O){len(mols)}", redraws=redraws,
                    meta={"grid": list(grid), "seed": seed})


def _cross(u, v):
    return (u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0])


def _place(a, b, c, bond, angle_deg, torsion_deg):
    """Position of atom d with |cd| = bond, angle(b, c, d) and dihedral(a, b, c, d) (NeRF).

    Plain-tuple arithmetic: the generator places ~10^5 atoms per batch and numpy's per-call
    overhead on 3-vectors dominated it."""
    ang = math.radians(angle_deg)
    tor = math.radians(torsion_deg)
    bc = (c[0] - b[0], c[1] - b[1], c[2] - b[2])
    l = math.sqrt(bc[0] ** 2 + bc[1] ** 2 + bc[2] ** 2)
    bc = (bc[0] / l, bc[1] / l, bc[2] / l)
    n = _cross((b[0] - a[0], b[1] - a[1], b[2] - a[2]), bc)
    nn = math.sqrt(n[0] ** 2 + n[1] ** 2 + n[2] ** 2)
    if nn < 1e-8:  # colinear reference: pick any perpendicular
        n = _cross(bc, (1.0, 0.0, 0.0))
        nn = math.sqrt(n[0] ** 2 + n[1] ** 2 + n[2] ** 2)
        if nn < 1e-8:
            n = _cross(bc, (0.0, 1.0, 0.0))
            nn = math.sqrt(n[0] ** 2 + n[1] ** 2 + n[2] ** 2)
    n = (n[0] / nn, n[1] / nn, n[2] / nn)
    m = _cross(n, bc)
    x = -bond * math.cos(ang)
    y = bond * math.sin(ang) * math.cos(tor)
    zz = bond * math.sin(ang) * math.sin(tor)
    return np.array([c[k] + x * bc[k] + y * m[k] + zz * n[k] for k in range(3)])


def _bond_len(z1, z2):
    key = tuple(sorted((z1, z2)))
    return {(6, 6): 1.54, (1, 6): 1.09, (6, 8): 1.43, (1, 8): 0.96, (6, 7): 1.47, (1, 7): 1.01}[key]


def branched_chain(rng: np.random.Generator, n_carbon: int | None = None, min_contact: float = 1.6,
                   max_redraws: int = 2000) -> Molecule:
    """One branched saturated chain with optional OH/NH2 terminal substitutions."""
    n = int(rng.integers(3, 21)) if n_carbon is None else int(n_carbon)
    branched = n >= 4 and rng.random() < 0.3
    # carbon skeleton: main chain of m carbons plus (optionally) one methyl branch
    m = n - 1 if branched else n
    parent = [-1] + list(range(m - 1))
    if branched:
        at = int(rng.integers(1, m - 1))
        parent.append(at)
    nc = len(parent)
    children = [[] for _ in range(nc)]
    for c in range(1, nc):
        children[parent[c]].append(c)
    # terminal carbons (one carbon neighbour) may become O or N
    nbrs = [len(children[c]) + (1 if parent[c] >= 0 else 0) for c in range(nc)]
    heavy_z = [6] * nc
    for c in range(nc):
        if nbrs[c] == 1 and rng.random() < 0.2:
            heavy_z[c] = 8 if rng.random() < 0.5 else 7
    redraws = 0
    for attempt in range(max_redraws):
        pos = _build_skeleton(rng, parent, children, heavy_z)
        if pos is not None:
            z, xyz, bonded = pos
            if _clash_free(xyz, bonded, min_contact):
                return Molecule(z=np.array(z, dtype=np.int32), xyz=np.array(xyz), name=f"chain C{n}",
                                redraws=redraws, meta={"n_carbon": n, "branched": branched})
        redraws += 1
    raise RuntimeError("chain generator: no clash-free conformation")


def _build_skeleton(rng, parent, children, heavy_z):
    nc = len(parent)
    zs: list[int] = []
    xyz: list[np.ndarray] = []
    bonded: list[tuple[int, int]] = []
    idx = [-1] * nc
    # C0 at origin, C1 (first child) along +x, a fixed reference point for the first dihedral
    root = 0
    idx[root] = 0
    zs.append(heavy_z[root])
    xyz.append(np.zeros(3))
    ghost_ref = [None] * nc  # a point defining the dihedral frame behind each heavy atom
    order = [root]
    placed_dir = {}
    # place the root's substituents relative to a virtual frame
    root_sub_dirs = None
    q = [root]
    while q:
        c = q.pop(0)
        pc = parent[c]
        kids = children[c]
        # substituent slots: heavy children first, then hydrogens fill to the valence
        valence = {6: 4, 7: 3, 8: 2}[heavy_z[c]]
        n_h = valence - len(kids) - (1 if pc >= 0 else 0)
        if pc < 0:
            # root: its first child lies along +x; other substituents tetrahedral around it
            a_ref = np.array([0.0, 1.0, 0.0])
            b_ref = np.array([1.0, 0.0, 0.0])  # virtual atom on the bond axis
            base = xyz[idx[c]]
            slots = []
            if kids:
                first = kids[0]
                bl = _bond_len(heavy_z[c], heavy_z[first])
                p = base + np.array([bl, 0.0, 0.0])
                slots.append((first, p))
                # remaining tetrahedral positions around root relative to the first bond
                tors0 = float(rng.choice([60.0, 180.0, 300.0]))
                rem = kids[1:]
                k = 0
                for t in range(1, 4):
                    tor = tors0 + 120.0 * (t - 1)
                    if k < len(rem):
                        bl2 = _bond_len(heavy_z[c], heavy_z[rem[k]])
                        pp = _place(a_ref + p, p, base, bl2, TETRA, tor)
                        slots.append((rem[k], pp))
                        k += 1
                    elif n_h > 0:
                        pp = _place(a_ref + p, p, base, _bond_len(heavy_z[c], 1), TETRA, tor)
                        slots.append((None, pp))
                        n_h -= 1
            for child, p in slots:
                if child is None:
                    zs.append(1)
                    xyz.append(p)
                    bonded.append((idx[c], len(zs) - 1))
                else:
                    idx[child] = len(zs)
                    zs.append(heavy_z[child])
                    xyz.append(p)
                    bonded.append((idx[c], idx[child]))
                    ghost_ref[child] = xyz[idx[c]] + np.array([0.0, 1.0, 0.0])
                    q.append(child)
            continue
        # non-root: frame from (grand-reference, parent, self)
        a = ghost_ref[c] if parent[pc] < 0 and ghost_ref[c] is not None else (
            xyz[idx[parent[pc]]] if parent[pc] >= 0 else ghost_ref[c])
        b = xyz[idx[pc]]
        cpos = xyz[idx[c]]
        tors0 = float(rng.choice([60.0, 180.0, 300.0]))
        subs = list(kids) + [None] * max(n_h, 0)
        # OH / NH2 hydrogens: keep the tetrahedral slots (ideal directions)
        for t, child in enumerate(subs[:3]):
            tor = tors0 + 120.0 * t
            if child is None:
                bl = _bond_len(heavy_z[c], 1)
                p = _place(a, b, cpos, bl, TETRA, tor)
                zs.append(1)
                xyz.append(p)
                bonded.append((idx[c], len(zs) - 1))
            else:
                bl = _bond_len(heavy_z[c], heavy_z[child])
                p = _place(a, b, cpos, bl, TETRA, tor)
                idx[child] = len(zs)
                zs.append(heavy_z[child])
                xyz.append(p)
                bonded.append((idx[c], idx[child]))
                q.append(child)
    return zs, xyz, bonded


def _clash_free(xyz, bonded, min_contact):
    """True when every pair more than two bonds apart is at least ``min_contact`` apart."""
    n = len(xyz)
    P = np.asarray(xyz)
    A = np.eye(n, dtype=np.int32)
    for a, b in bonded:
        A[a, b] = A[b, a] = 1
    near = (A @ A) > 0  # within two bonds (1-2 and 1-3 separations are fixed by construction)
    d2 = ((P[:, None, :] - P[None, :, :]) ** 2).sum(-1)
    return not np.any((d2 < min_contact * min_contact) & ~near)


def _chain_at(args):
    seed, index = args
    return branched_chain(np.random.default_rng([seed, index]))


def chain_batch(count: int, seed: int, workers: int = 0) -> list[Molecule]:
    """The batched workload: ``count`` chains from one seed. Molecule ``i`` is drawn from its own
    stream ``PCG64([seed, i])``, so the batch is the same however it is split across workers."""
    jobs = [(seed, i) for i in range(count)]
    if workers <= 0:
        import os
        workers = min(16, os.cpu_count() or 1)
    if count >= 256 and workers > 1:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            return pool.map(_chain_at, jobs, chunksize=32)
    return [_chain_at(j) for j in jobs]


def water_cluster_config(which: int) -> Molecule:
    """The water-cluster workloads: 32 H2O (4x4x2, seed 0), 512 H2O (8x8x8, seed 2),
    2000 H2O (20x10x10, seed 3)."""
    if which == 0:
        return water_cluster((4, 4, 2), seed=0)
    if which == 2:
        return water_cluster((8, 8, 8), seed=2)
    if which == 3:
        return water_cluster((20, 10, 10), seed=3)
    raise ValueError(which)


def small_water_cluster(k: int, rng: np.random.Generator) -> Molecule:
    side = max(1, math.ceil(k ** (1.0 / 3.0)))
    grid = (side, side, math.ceil(k / (side * side)))
    return water_cluster(grid, seed=int(rng.integers(0, 2**31)), count=k)


@dataclass
class ServingRequest:
    molecule: Molecule
    method: str
    bcc: bool
    arrival_s: float


def serving_stream(count: int, rate_per_s: float, seed_arrivals: int = 4, seed_molecules: int = 5):
    """Poisson request stream: 70 % chains, 30 % (H2O)_k with k ~ U{2..40}; 10 % lose one H and
    run as doublet UHF; AM1 80 % / RM1 20 %; 30 % also ask for AM1-BCC charges."""
    ra = np.random.default_rng(seed_arrivals)
    rm = np.random.default_rng(seed_molecules)
    t = 0.0
    out = []
    for _ in range(count):
        t += ra.exponential(1.0 / rate_per_s)
        if rm.random() < 0.7:
            mol = branched_chain(rm)
        else:
            mol = small_water_cluster(int(rm.integers(2, 41)), rm)
        if rm.random() < 0.1:
            hs = np.nonzero(mol.z == 1)[0]
            drop = int(hs[rm.integers(0, len(hs))])
            keep = np.ones(len(mol.z), dtype=bool)
            keep[drop] = False
            mol = Molecule(z=mol.z[keep], xyz=mol.xyz[keep], multiplicity=2, name=mol.name + " radical",
                           redraws=mol.redraws)
        method = "am1" if rm.random() < 0.8 else "rm1"
        bcc = bool(rm.random() < 0.3)
        out.append(ServingRequest(mol, method, bcc, t))
    return out
