"""Promolecular radial density tables (critic2's grid1 type).

``build_grid1`` restates how critic2 turns sum(occ*psi^2) on a logarithmic
radial grid into the density table and its first and second radial
derivatives (read_critic, grid1mod@proc.f90:206-331): the grid is cut where
rho < 1e-8 (core_cutdens), the derivatives come from 6-point finite
differences in the log variable, and everything is divided by 4 pi.  The same
builder serves critic2's atomic wfc files and synthetic Slater densities, so
both reach the GPU evaluator through identical table semantics.
"""
from __future__ import annotations

import math

import numpy as np

BOHR_PER_ANG = 1.0 / 0.52917720859

# cutoff radii (bohr) from global.f90:53 (cutrad), for the elements used here
CUTRAD = {1: 2.149886192475e+01, 6: 2.167675592180e+01, 7: 1.749805708313e+01,
          8: 1.465173060207e+01, 14: 2.847261278601e+01, 26: 2.609916866690e+01}

_NOEF = np.array([[0, 1, 2, 3, 4, 5], [-2, -1, 0, 1, 2, 3], [-5, -4, -3, -2, -1, 0]])
_COEF1 = np.array([[-274, 600, -600, 400, -150, 24], [6, -60, -40, 120, -30, 4],
                   [-24, 150, -400, 600, -600, 274]], dtype=np.float64)
_COEF2 = np.array([[225, -770, 1070, -780, 305, -50], [-5, 80, -150, 80, -5, 0],
                   [-50, 305, -780, 1070, -770, 225]], dtype=np.float64)
CORE_CUTDENS = 1e-8


class Grid1:
    """A radial table: nodes r, density f, derivatives fp, fpp; a, b, rmax."""

    def __init__(self, r, f, fp, fpp, z=0):
        self.r, self.f, self.fp, self.fpp = r, f, fp, fpp
        self.ngrid = len(r)
        self.a = float(r[0])
        self.b = math.log(r[1] / r[0])
        self.rmax = float(r[-1])
        self.z = z

    @property
    def cutrad(self):
        return CUTRAD.get(self.z, self.rmax)


def truncate(r, rr0):
    """Number of nodes kept by read_critic: stop at the first node i > 1
    where rr0/(4 pi r^2) < core_cutdens (grid1mod@proc.f90:258-266)."""
    for i in range(len(r)):
        if rr0[i] / (4.0 * math.pi * r[i] ** 2) < CORE_CUTDENS and i > 0:
            return i + 1
    return len(r)


def build_grid1(r, rr0, z=0) -> Grid1:
    """f, fp, fpp from rr0 = 4 pi r^2 rho on a log grid (already truncated),
    with the reference's operation order (grid1mod@proc.f90:281-313)."""
    r = np.asarray(r, dtype=np.float64)
    rr0 = np.asarray(rr0, dtype=np.float64)
    ng = len(r)
    b = math.log(r[1] / r[0])
    f = np.zeros(ng)
    fp = np.zeros(ng)
    fpp = np.zeros(ng)
    fac1, fac2 = 1.0 / 120.0, 2.0 / 120.0
    delta = 1.0 / b
    delta2 = delta * delta
    for i in range(ng):
        i1 = i + 1
        ic = 0 if i1 <= 2 else (2 if i1 >= ng - 2 else 1)
        d1 = 0.0
        d2 = 0.0
        for j in range(6):
            d1 = d1 + _COEF1[ic, j] * rr0[i + _NOEF[ic, j]]
            d2 = d2 + _COEF2[ic, j] * rr0[i + _NOEF[ic, j]]
        d1 = d1 * fac1
        d2 = d2 * fac2
        rv = r[i]
        r1 = 1.0 / rv
        r2 = r1 * r1
        r3 = r2 * r1
        r4 = r3 * r1
        f[i] = rr0[i] * r2
        fp[i] = (d1 * delta - 2.0 * rr0[i]) * r3
        fpp[i] = (d2 * delta2 - 5.0 * d1 * delta + 6.0 * rr0[i]) * r4
    fourpi = 4.0 * math.pi
    return Grid1(r, f / fourpi, fp / fourpi, fpp / fourpi, z=z)


def slater_grid1(z: int, ngrid=2000, rmin_ang=1e-4, rmax_ang=12.0) -> Grid1:
    """Table of the Slater model rho = Z zeta^3/(8 pi) exp(-zeta r),
    zeta = 2 Z^(1/3), on ngrid log-spaced nodes from rmin to rmax (angstrom),
    built as critic2 builds a table from an atomic wfc file."""
    a = rmin_ang * BOHR_PER_ANG
    rmax = rmax_ang * BOHR_PER_ANG
    b = math.log(rmax / a) / (ngrid - 1)
    r = a * np.exp(b * np.arange(ngrid))
    zeta = 2.0 * z ** (1.0 / 3.0)
    rho = z * zeta ** 3 / (8.0 * math.pi) * np.exp(-zeta * r)
    rr0 = 4.0 * math.pi * r * r * rho
    ng = truncate(r, rr0)
    return build_grid1(r[:ng], rr0[:ng], z=z)


def read_wfc(path: str, nelec: int):
    """Parse a critic2 atomic wfc file (grid1mod@proc.f90:233-262):
    returns (r, rr0) truncated as read_critic does."""
    with open(path) as fh:
        toks = fh.read().split()
    pos = 0
    norb = int(toks[pos]); pos += 1
    pos += norb  # labels
    occ = np.array([float(t) for t in toks[pos:pos + norb]]); pos += norb
    pos += norb  # energies
    ngrid = int(toks[pos]); pos += 1
    if occ.sum() != nelec:
        ns = 0
        for i in range(norb):
            if ns + occ[i] > nelec:
                occ[i] = nelec - ns
                occ[i + 1:] = 0
                break
            ns += int(round(occ[i]))
    vals = np.array([float(t.replace("D", "E")) for t in toks[pos:pos + ngrid * (norb + 1)]])
    vals = vals.reshape(ngrid, norb + 1)
    r = vals[:, 0].copy()
    psi = vals[:, 1:]
    rr0 = np.zeros(ngrid)
    for i in range(ngrid):
        acc = 0.0
        for j in range(norb):
            acc = acc + occ[j] * psi[i, j] ** 2
        rr0[i] = acc
    ng = truncate(r, rr0)
    return r[:ng].copy(), rr0[:ng].copy()
