"""The benchmark workloads on the GPU, each against the oracle at a size the
oracle finishes in seconds and, at the full size, through properties that
hold whatever the size:

- 512^3 cubic cell (30 A, 216 atoms): at 128^3 exact labels, Volume/Pop
  sums, FP64 promolecular NCI and the FP32 path's errors; at 512^3 the
  counts sum to N, the populations to the grid sum, every atom's nearest
  node lies in its own basin, and a second run gives identical labels.
- 512-cell dimer batch: 16 of its cells, every grid edge of
  {48, 64, 96, 128, 160} among them, through c2gpu_batch_run against the
  oracle cell by cell.
- 768^3 triclinic cell (40 A, 500 atoms, Z up to 26 with Si and Fe tables):
  at 96^3 exact labels, sums and NCI against the oracle; at 768^3 the same
  properties as the 512^3 cell.
- 256^3 orthorhombic cell: the NCI FFT scheme against the oracle's
  grid3%fft restatement.
"""
import numpy as np
import pytest

from critic2_amd import bader as gb
from critic2_amd import batch as gbt
from critic2_amd import nci as gn
from critic2_amd.radial import slater_grid1
from critic2_amd.synth import GRID_EDGES, batch_cells, config
from oracle import oracle
from tests.helpers import images, node_points

pytestmark = pytest.mark.gpu


def _tables(cell):
    zs = sorted(set(int(z) for z in cell.zat))
    return [slater_grid1(z) for z in zs], np.array([zs.index(int(z)) for z in cell.zat], dtype=np.int32)


def _bader_vs_oracle(ctx, cell):
    f = gb.synth_slater(ctx, cell)
    bas = gb.bader_integrate(ctx, f, cell.n, cell.x2c, cell.xat)
    ref, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat)
    assert int(np.count_nonzero(bas.idg != ref)) == 0
    assert bas.nattr == info["nattr"]
    np.testing.assert_array_equal(bas.xattr, xattr)
    c_ref, s_ref = oracle.basin_sums(ref, info["nattr"], [f])
    c_gpu, s_gpu = gb.basin_sums(ctx, bas.idg, bas.nattr, [f])
    np.testing.assert_array_equal(c_gpu, c_ref)
    np.testing.assert_allclose(s_gpu, s_ref, rtol=1e-10)
    return f, bas


def _nci_vs_oracle(ctx, cell, sample):
    """FP64 promolecular NCI at every node against the oracle on `sample`
    node indices (z-fastest order), and the FP32 path's error report."""
    tables, species = _tables(cell)
    c64, s64 = gn.nciplot_promol(ctx, cell.n, cell.x2c, tables, cell.xat, species, precision="fp64")
    c32, s32 = gn.nciplot_promol(ctx, cell.n, cell.x2c, tables, cell.xat, species, precision="fp32")
    otabs = [oracle.Table(g.r, g.f, g.fp, g.fpp, g.cutrad) for g in tables]
    pos, idx = images(cell.x2c, cell.xat, lim=1)
    pts = node_points(cell.x2c, cell.n)[sample]
    ref = oracle.promol_nci_batch(pts, pos, species[idx], otabs)
    np.testing.assert_allclose(np.abs(c64[sample]), np.abs(ref[:, 3]), rtol=1e-12, atol=1e-300)
    ok = ref[:, 0] > 1e-10
    np.testing.assert_allclose(s64[sample][ok], ref[ok, 2], rtol=1e-9)
    assert np.mean(np.sign(c64[sample]) == np.sign(ref[:, 3])) > 0.999
    zmax = 2.0 * max(int(z) for z in cell.zat) ** (1.0 / 3.0)
    rep = gn.precision_report(c32, s32, c64, s64, zeta_max=zmax)
    print(f"  FP32 vs FP64 in s<2: {rep}", flush=True)
    assert rep["points"] > 0
    # rho: a sum of positive terms, FP32 round-off only
    assert rep["rho_max_rel"] < 1e-5
    # s and |grad rho|: within a fixed multiple of the cancellation bound
    # (one rounding per atom term; a term's FP32 evaluation -- distance,
    # spline, direction scaling -- and the running sum over the atoms in
    # range take about a dozen roundings, so 16 bounds are allowed), and
    # within 1e-5 wherever that bound is below 1e-6
    assert rep["s_over_bound_max"] < 16.0
    assert rep["s_max_rel_well_conditioned"] < 1e-5
    assert rep["s_below_floor_max_abs"] < 1e-5
    return rep


def test_cubic_512_workload_cell_at_128(ctx):
    cell = config(3, n_override=(128, 128, 128))
    print(f"[case] 30 A / 216 atoms at {cell.n}", flush=True)
    _bader_vs_oracle(ctx, cell)
    N = cell.npts
    sample = np.arange(0, N, 97)
    _nci_vs_oracle(ctx, cell, sample)


def _nearest_nodes(cell):
    """Linear x-fastest index of each atom's nearest grid node."""
    n = np.asarray(cell.n)
    ijk = np.rint(np.asarray(cell.xat) * n).astype(np.int64) % n
    return ijk[:, 0] + n[0] * (ijk[:, 1] + n[1] * ijk[:, 2])


def _full_size_properties(ctx, cell, repeat):
    N = cell.npts
    rho = ctx.empty((N,), np.float64)
    gb.synth_slater(ctx, cell, device_out=rho)
    lab = ctx.empty((N,), np.int32)
    bas = gb.bader_integrate(ctx, rho, cell.n, cell.x2c, cell.xat, labels_out=lab)
    print(f"  stats {bas.stats}", flush=True)
    counts, sums = gb.basin_sums(ctx, lab, bas.nattr, [rho])
    f = rho.numpy()
    L = lab.numpy()
    assert int(counts.sum()) == N
    assert int(np.count_nonzero((L < 1) | (L > bas.nattr))) == 0
    tot = float(np.sum(f))
    assert abs(float(np.sum(sums)) - tot) <= 1e-12 * tot
    # every atom is an attractor (nearest-atom rule) and its nearest node is in its basin
    near = _nearest_nodes(cell)
    assert np.array_equal(L[near], np.arange(1, len(cell.zat) + 1))
    if repeat:
        lab2 = ctx.empty((N,), np.int32)
        gb.bader_integrate(ctx, rho, cell.n, cell.x2c, cell.xat, labels_out=lab2)
        assert np.array_equal(lab2.numpy(), L)
        lab2.free()
    rho.free()
    lab.free()


@pytest.mark.slow
def test_cubic_512_workload_properties(ctx):
    cell = config(3)
    print(f"[case] 30 A / 216 atoms at {cell.n}", flush=True)
    _full_size_properties(ctx, cell, repeat=True)


def _batch_selection():
    """16 cells of the 512-cell batch, chosen in order so that every grid
    edge value appears, preferring small cells."""
    cells = batch_cells(512)
    order = sorted(range(len(cells)), key=lambda q: cells[q].npts)
    chosen, seen = [], set()
    for q in order:  # the smallest cell that brings a new edge value
        new = set(cells[q].n) - seen
        if new:
            chosen.append(q)
            seen |= set(cells[q].n)
        if seen == set(GRID_EDGES):
            break
    for q in order:
        if len(chosen) >= 16:
            break
        if q not in chosen:
            chosen.append(q)
    return [cells[q] for q in sorted(chosen)]


def test_batch_workload_cells(ctx):
    cells = _batch_selection()
    assert len(cells) == 16 and set(e for c in cells for e in c.n) == set(GRID_EDGES)
    print(f"[case] 16 dimer cells, {sum(c.npts for c in cells)} points", flush=True)
    zs = (1, 6, 7, 8)
    tables = [slater_grid1(z) for z in zs]
    species = [np.array([zs.index(int(z)) for z in c.zat], dtype=np.int32) for c in cells]
    grids = [gb.synth_slater(ctx, c) for c in cells]
    res, st = gbt.batch_run(ctx, cells, grids, tables, species)
    for q, (cell, f, r) in enumerate(zip(cells, grids, res)):
        ref, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat)
        assert int(np.count_nonzero(r["labels"] != ref)) == 0, f"cell {q} {cell.n}"
        assert r["nattr"] == info["nattr"]
        counts, sums = oracle.basin_sums(ref, info["nattr"], [f])
        np.testing.assert_array_equal(r["counts"], counts)
        np.testing.assert_allclose(r["sums"], sums[0], rtol=1e-10)
        c1, s1 = gn.nciplot_promol(ctx, cell.n, cell.x2c, tables, cell.xat, species[q], precision="fp32")
        assert np.array_equal(r["crho"], c1) and np.array_equal(r["cgrad"], s1), f"cell {q} NCI"


def test_triclinic_768_workload_cell_at_96(ctx):
    cell = config(5, n_override=(96, 96, 96))
    assert {14, 26} <= set(int(z) for z in cell.zat)
    print(f"[case] triclinic 40 A / 500 atoms at {cell.n}", flush=True)
    _bader_vs_oracle(ctx, cell)
    sample = np.arange(0, cell.npts, 211)
    _nci_vs_oracle(ctx, cell, sample)


@pytest.mark.slow
def test_triclinic_768_workload_properties(ctx):
    cell = config(5)
    print(f"[case] triclinic 40 A / 500 atoms at {cell.n}", flush=True)
    _full_size_properties(ctx, cell, repeat=False)


@pytest.mark.slow
def test_orthorhombic_256_nci_fft(ctx):
    cell = config(2)
    print(f"[case] NCI FFT at {cell.n}", flush=True)
    f = gb.synth_slater(ctx, cell)
    cf, sf = gn.nciplot_grid(ctx, f, cell.n, cell.x2c, zfast=False)
    cr, sr = oracle.nci_grid_fft(f, cell.n, cell.x2c)
    np.testing.assert_allclose(np.abs(cf), np.abs(cr), rtol=1e-12)
    # FFT round-off is absolute, about 1e-15 of the largest |grad rho| (the
    # oracle's numpy FFT and the reference's fftpack differ by as much): the
    # gradient magnitudes s * const * rho^(4/3) agree to 1e-12 of the largest,
    # and s itself to 1e-8 where rho is not tiny
    ok = f > 1e-8
    gg = sf[ok] * gn.RDG_CONST * f[ok] ** (4.0 / 3.0)
    go = sr[ok] * gn.RDG_CONST * f[ok] ** (4.0 / 3.0)
    assert np.max(np.abs(gg - go)) <= 1e-12 * np.max(np.abs(go))
    dense = f > 1e-3
    np.testing.assert_allclose(sf[dense], sr[dense], rtol=1e-8)
    assert np.mean(np.sign(cf) == np.sign(cr)) > 0.9999
