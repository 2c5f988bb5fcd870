"""Batches of cells (c2gpu_batch_run): every cell's Bader labels, attractor
count and Volume/Pop sums equal the oracle's, and its FP32 promolecular NCI
equals the single-cell GPU result bit for bit (the work-queue kernel runs the
same per-node code over the same image lists)."""
import numpy as np
import pytest

from critic2_amd import batch as gbt
from critic2_amd import bader as gb
from critic2_amd import nci as gn
from critic2_amd.radial import slater_grid1
from critic2_amd.synth import random_cell
from oracle import oracle

pytestmark = pytest.mark.gpu

SHAPES = [(24, 28, 32), (32, 32, 32), (20, 36, 28), (40, 24, 24), (28, 28, 28), (36, 32, 20), (30, 26, 34)]


def _cells():
    cells = []
    for q, n in enumerate(SHAPES):
        ang = (85 + 2 * q, 90, 95 - q) if q % 2 else (90, 90, 90)
        cells.append(random_cell(700 + q, (7.0 + 0.3 * q, 8.0, 7.5 + 0.2 * q), ang, 3 + q % 4, (1, 6, 7, 8), n))
    return cells


def _tables(cells):
    zs = sorted({int(z) for c in cells for z in c.zat})
    tables = [slater_grid1(z) for z in zs]
    species = [np.array([zs.index(int(z)) for z in c.zat], dtype=np.int32) for c in cells]
    return tables, species


def test_batch_matches_single_cells_and_oracle(ctx):
    cells = _cells()
    print(f"[case] batch of {len(cells)} cells", flush=True)
    grids = [gb.synth_slater(ctx, c) for c in cells]
    tables, species = _tables(cells)
    res, st = gbt.batch_run(ctx, cells, grids, tables, species, nstreams=3)
    print(f"  stats {st}", flush=True)
    assert st["points"] == sum(g.size for g in grids)
    for q, (cell, f, r) in enumerate(zip(cells, grids, res)):
        ref, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat)
        assert int(np.count_nonzero(r["labels"] != ref)) == 0, f"cell {q}"
        assert r["nattr"] == info["nattr"]
        np.testing.assert_array_equal(r["xattr"], xattr)
        counts, sums = oracle.basin_sums(ref, info["nattr"], [f])
        np.testing.assert_array_equal(r["counts"], counts)
        np.testing.assert_allclose(r["sums"], sums[0], rtol=1e-10)
        c1, s1 = gn.nciplot_promol(ctx, cell.n, cell.x2c, tables, cell.xat, species[q], precision="fp32")
        assert np.array_equal(r["crho"], c1) and np.array_equal(r["cgrad"], s1), f"cell {q} NCI"


def test_batch_attractor_rule_and_partial_outputs(ctx):
    """NOATOMS numbering per cell, and cells that ask only for NCI or only for
    Bader."""
    cells = _cells()[:4]
    grids = [gb.synth_slater(ctx, c) for c in cells]
    tables, species = _tables(cells)
    res, _ = gbt.batch_run(ctx, cells, grids, tables, species, noatoms=True, ratom=1.0, nstreams=2)
    for cell, f, r in zip(cells, grids, res):
        ref, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat, mode=2, ratom=1.0)
        assert int(np.count_nonzero(r["labels"] != ref)) == 0
        np.testing.assert_array_equal(r["xattr"], xattr)
    only_nci, _ = gbt.batch_run(ctx, cells, grids, tables, species, bader=False)
    assert all("labels" not in r and r["crho"].size == f.size for r, f in zip(only_nci, grids))
    only_bader, st = gbt.batch_run(ctx, cells, grids, bader=True, nci=False)
    assert st["ms_nci"] == 0.0 and all("crho" not in r for r in only_bader)
