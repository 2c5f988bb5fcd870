"""GPU basin multipoles (c2gpu_multipoles) against the oracle's restatement
through spherical angles and genylm: agreement to rounding per (basin, l)
block, on Bader labels of orthorhombic and triclinic cells."""
import numpy as np
import pytest

from critic2_amd import bader as gb
from critic2_amd.synth import random_cell
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,edge,ang,n,lmax,device", [
    (71, (8.0, 8.5, 9.0), (90, 90, 90), (32, 34, 36), 5, False),
    (72, (8.5, 9.0, 8.0), (80, 97, 105), (34, 32, 36), 5, True),
    (73, (9.0, 9.0, 9.5), (75, 100, 110), (30, 32, 34), 8, False),
])
def test_multipoles_match_oracle(ctx, seed, edge, ang, n, lmax, device):
    cell = random_cell(seed, edge, ang, 4, (1, 6, 7, 8), n)
    print(f"[case] multipoles seed={seed} n={cell.n} lmax={lmax}", flush=True)
    f = gb.synth_slater(ctx, cell)
    idg, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat)
    nattr = info["nattr"]
    xa = np.asarray(xattr).reshape(-1, 3)[:nattr]
    ref = oracle.multipoles(idg, f, cell.n, cell.x2c, xa, lmax)
    if device:
        got = gb.multipoles(ctx, ctx.to_device(idg.astype(np.int32)), ctx.to_device(f), cell.n, cell.x2c, xa, lmax)
    else:
        got = gb.multipoles(ctx, idg, f, cell.n, cell.x2c, xa, lmax)
    for l in range(lmax + 1):
        s = slice(l * l, (l + 1) * (l + 1))
        scale = np.max(np.abs(ref[:, s]), axis=1, keepdims=True)
        err = np.abs(got[:, s] - ref[:, s]) / np.maximum(scale, 1e-300)
        assert err.max() < 1e-10, (l, err.max())
