"""Deterministic synthetic inputs for the on-grid Bader/NCI workloads.

The reference ships no density grids that this path can run on, so inputs are
generated: a periodic cell, seeded atoms, and the Slater-sum model density

    rho(r) = sum_{atoms, images} Z * zeta^3/(8 pi) * exp(-zeta |r - R|),
    zeta = 2 Z^(1/3)  (bohr^-1),

summed over every periodic image within 8 angstrom of the point.  Each atom
integrates to Z electrons.  Lengths are bohr internally; cells are specified
in angstrom and converted with 1 A = 1/0.52917720859 bohr (critic2's
``bohrtoa``, param.F90:42).

Random numbers come from splitmix64 so that every host reproduces the same
atoms bit for bit.  Atom order is generation order; it becomes the atom /
attractor numbering, as the cube reader keeps file order
(crystalseedmod@proc.f90:2909-2932).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

BOHR_PER_ANG = 1.0 / 0.52917720859
MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014), 64-bit state."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, n: int) -> int:
        return int(self.uniform() * n)


@dataclass
class Cell:
    """A periodic cell with atoms.

    x2c: 3x3 array, lattice vectors as columns, bohr.
    xat: (nat, 3) fractional coordinates.  zat: (nat,) atomic numbers.
    """

    x2c: np.ndarray
    xat: np.ndarray
    zat: np.ndarray
    n: tuple = (64, 64, 64)
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def omega(self) -> float:
        return abs(float(np.linalg.det(self.x2c)))

    @property
    def npts(self) -> int:
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])


def lattice_from_params(a, b, c, alpha, beta, gamma) -> np.ndarray:
    """Lattice vectors (columns, same length unit as a,b,c) with a along x and
    b in the xy plane; angles in degrees."""
    al, be, ga = (math.radians(t) for t in (alpha, beta, gamma))
    ca, cb, cg, sg = math.cos(al), math.cos(be), math.cos(ga), math.sin(ga)
    v = math.sqrt(1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg)
    m = np.zeros((3, 3))
    m[:, 0] = (a, 0.0, 0.0)
    m[:, 1] = (b * cg, b * sg, 0.0)
    m[:, 2] = (c * cb, c * (ca - cb * cg) / sg, c * v / sg)
    return m


def _min_periodic_dist(x2c, xa, xs):
    """Shortest distance between fractional point xa and any of the points xs
    under lattice translations (brute force over -1..1 after folding)."""
    if len(xs) == 0:
        return float("inf")
    d = np.asarray(xs) - np.asarray(xa)[None, :]
    d -= np.round(d)
    best = np.full(len(xs), np.inf)
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                dc = (d + np.array([i, j, k])) @ x2c.T
                best = np.minimum(best, np.sqrt((dc * dc).sum(1)))
    return float(best.min())


def random_cell(seed: int, cell_ang, angles, nat: int, zset, n, min_sep_ang=1.5,
                name="") -> Cell:
    """Atoms uniform in fractional [0,1)^3, rejection-sampled to a minimum
    periodic separation; Z uniform over zset."""
    rng = SplitMix64(seed)
    x2c = lattice_from_params(*(np.array(cell_ang) * BOHR_PER_ANG), *angles)
    xs, zs = [], []
    min_sep = min_sep_ang * BOHR_PER_ANG
    tries = 0
    while len(xs) < nat:
        tries += 1
        if tries > 1000000:
            raise RuntimeError("could not place atoms")
        x = np.array([rng.uniform(), rng.uniform(), rng.uniform()])
        if _min_periodic_dist(x2c, x, xs) < min_sep:
            continue
        xs.append(x)
        zs.append(zset[rng.randint(len(zset))])
    return Cell(x2c=x2c, xat=np.array(xs), zat=np.array(zs, dtype=np.int32), n=tuple(n),
                name=name)


def config(cid: int, n_override=None) -> Cell:
    """The five workloads: 1) 64^3 cubic a=10 A, 8 atoms; 2) 256^3
    orthorhombic 20x20x24 A, 64 atoms; 3) 512^3 cubic 30 A, 216 atoms;
    5) 768^3 triclinic a=b=c=40 A, 500 atoms.  (4, the 512-cell batch, is
    built by batch_cells.)"""
    if cid == 1:
        c = random_cell(1, (10, 10, 10), (90, 90, 90), 8, (1, 6, 7, 8), (64, 64, 64),
                        name="cubic a=10A 64^3 8 atoms")
    elif cid == 2:
        c = random_cell(2, (20, 20, 24), (90, 90, 90), 64, (1, 6, 7, 8), (256, 256, 256),
                        name="orthorhombic 20x20x24A 256^3 64 atoms")
    elif cid == 3:
        c = random_cell(3, (30, 30, 30), (90, 90, 90), 216, (1, 6, 7, 8), (512, 512, 512),
                        name="cubic a=30A 512^3 216 atoms")
    elif cid == 5:
        rng = SplitMix64(5 ^ 0xA5A5)
        ang = [70 + 40 * rng.uniform() for _ in range(3)]
        c = random_cell(5, (40, 40, 40), ang, 500, (1, 6, 7, 8, 14, 26), (768, 768, 768),
                        name="triclinic a=b=c=40A 768^3 500 atoms")
    else:
        raise ValueError(f"unknown config {cid}")
    if n_override is not None:
        c.n = tuple(n_override)
    return c


GRID_EDGES = (48, 64, 96, 128, 160)


def batch_cells(ncell: int = 512, base_seed: int = 1000):
    """Dimer cells: two monomers of 3-20 atoms each, grown as random chains at
    1.0-1.5 A spacing, separated by U(2.5, 4.0) A along a random axis; cell =
    bounding box + 8 A per axis (orthorhombic); grid edges drawn per axis from
    {48, 64, 96, 128, 160}."""
    cells = []
    for c in range(ncell):
        rng = SplitMix64(base_seed + c)
        mons = []
        for _m in range(2):
            na = 3 + rng.randint(18)
            pos = [np.zeros(3)]
            while len(pos) < na:
                base = pos[rng.randint(len(pos))]
                u = np.array([rng.uniform() * 2 - 1 for _ in range(3)])
                nrm = np.linalg.norm(u)
                if nrm < 1e-3:
                    continue
                cand = base + u / nrm * (1.0 + 0.5 * rng.uniform())
                if min(np.linalg.norm(cand - q) for q in pos) < 0.95:
                    continue
                pos.append(cand)
            mons.append(np.array(pos))
        axis = rng.randint(3)
        gap = 2.5 + 1.5 * rng.uniform()
        a, b = mons
        shift = a[:, axis].max() - b[:, axis].min() + gap
        b = b.copy()
        b[:, axis] += shift
        allp = np.vstack([a, b])
        lo, hi = allp.min(0), allp.max(0)
        box = hi - lo + 8.0
        allp = allp - lo + 4.0
        x2c = np.diag(box * BOHR_PER_ANG)
        xat = allp / box
        zat = np.array([(1, 6, 7, 8)[rng.randint(4)] for _ in range(len(allp))], dtype=np.int32)
        n = tuple(GRID_EDGES[rng.randint(len(GRID_EDGES))] for _ in range(3))
        cells.append(Cell(x2c=x2c, xat=xat, zat=zat, n=n, name=f"dimer cell {c}"))
    return cells
