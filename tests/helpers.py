"""Shared test inputs: urea golden case, small synthetic cells."""
import json
import os

import numpy as np

from critic2_amd.radial import build_grid1
from critic2_amd.synth import Cell, lattice_from_params, random_cell, BOHR_PER_ANG

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def urea():
    d = json.load(open(os.path.join(GOLDEN, "urea_nci.json")))
    T = np.load(os.path.join(GOLDEN, "urea_tables.npz"))
    a, b, c = d["cell_bohr_deg"][:3]
    x2c = np.diag([a, b, c])
    syms = ["h", "c", "n", "o"]
    zs = {"h": 1, "c": 6, "n": 7, "o": 8}
    tables = [build_grid1(T["r_" + s], T["rr0_" + s], z=zs[s]) for s in syms]
    xat = np.array([at["x"] for at in d["atoms"]])
    species = np.array([syms.index(at["sym"].lower()) for at in d["atoms"]], dtype=np.int32)
    return d, x2c, tables, xat, species


def fmt6(v):
    """Value as the cube writer prints it, 6 significant digits (1p,e13.5)."""
    return float(f"{v:.5E}")


def small_cell(seed, n, nat=4, ang=(90, 90, 90), edge=(8.0, 8.0, 8.0), zset=(1, 6, 7, 8)):
    return random_cell(seed, edge, ang, nat, zset, n)


def images(x2c, xat, lim=3):
    pos, idx = [], []
    for t, x in enumerate(np.asarray(xat)):
        for i in range(-lim, lim + 1):
            for j in range(-lim, lim + 1):
                for k in range(-lim, lim + 1):
                    pos.append(np.asarray(x2c) @ (x + [i, j, k]))
                    idx.append(t)
    return np.array(pos), np.array(idx)


def node_points(x2c, n):
    i, j, k = np.meshgrid(np.arange(n[0]) / n[0], np.arange(n[1]) / n[1], np.arange(n[2]) / n[2],
                          indexing="ij")
    frac = np.stack([i, j, k], -1).reshape(-1, 3)  # i slowest, k fastest (z-fastest order)
    return frac @ np.asarray(x2c).T
