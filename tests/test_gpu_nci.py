"""GPU parity of the NCI path: promolecular FP64 against the reference's
golden urea values and the oracle; FP32 against FP64 (error reported);
FFT scheme against the oracle's grid3%fft restatement."""
import math
import os

import numpy as np
import pytest

from critic2_amd import _lib
from critic2_amd import bader as gb
from critic2_amd import nci as gn
from critic2_amd.radial import slater_grid1
from critic2_amd.synth import config, random_cell
from oracle import oracle
from tests.cubetext import fortran_e
from tests.helpers import GOLDEN, urea, fmt6, images, node_points

pytestmark = pytest.mark.gpu


def test_urea_golden_on_gpu_fp64(ctx):
    """tests/015_grdplot/005_nciplot_basic through the GPU promolecular path."""
    d, x2c, tables, xat, species = urea()
    crho, s = gn.nciplot_promol(ctx, d["nstep"], x2c, tables, xat, species, precision="fp64")
    o = d["options"]
    cg = s.copy()
    text, _ = gn.cutoffs(ctx, crho, cg, rhocut=o["rhocut"], dimcut=o["dimcut"], rhoplot=o["rhoplot"],
                         srhorange=o["srhorange"])
    dat = gn.dat_rows(text)
    assert [fmt6(v) for v in crho] == d["dens_cube"]
    # the .dat file of the reference: the density column byte for byte, the
    # RDG column where it is not round-off noise at a symmetric node
    ref_lines = open(os.path.join(GOLDEN, "cube", "005_nciplot_basic-gen.dat"), "rb").read().splitlines()
    got_lines = text.splitlines()
    assert len(got_lines) == len(ref_lines)
    for g, r in zip(got_lines, ref_lines):
        assert g[:15] == r[:15]
        if float(r[15:]) > 1e-10:
            assert g[15:] == r[15:]
    for v, ref in zip(cg, d["grad_cube"]):
        if abs(ref) > 1e-10:
            assert fmt6(v) == ref
        else:
            assert abs(v - ref) < d["cube_atol"]
    np.testing.assert_allclose(dat[:, 0], np.array(d["dat"])[:, 0], rtol=1e-7)


def _slater_setup(cell):
    zs = sorted(set(int(z) for z in cell.zat))
    tables = [slater_grid1(z) for z in zs]
    species = np.array([zs.index(int(z)) for z in cell.zat], dtype=np.int32)
    return tables, species


def test_promol_fp64_matches_oracle(ctx):
    cell = random_cell(31, (6.0, 6.5, 7.0), (90, 90, 90), 5, (1, 6, 7, 8), (10, 11, 12))
    tables, species = _slater_setup(cell)
    crho, s = gn.nciplot_promol(ctx, cell.n, cell.x2c, tables, cell.xat, species, precision="fp64")
    otabs = [oracle.Table(g.r, g.f, g.fp, g.fpp, g.cutrad) for g in tables]
    pos, idx = images(cell.x2c, cell.xat, lim=3)
    ref = oracle.promol_nci(node_points(cell.x2c, cell.n), pos, species[idx], otabs)
    np.testing.assert_allclose(np.abs(crho), np.abs(ref[:, 3]), rtol=1e-12, atol=1e-300)
    ok = ref[:, 0] > 1e-12
    np.testing.assert_allclose(s[ok], ref[ok, 2], rtol=1e-9)
    # lambda2 sign: identical except where lambda2 is round-off close to zero
    agree = np.mean(np.sign(crho) == np.sign(ref[:, 3]))
    assert agree > 0.999


def test_promol_fp32_error_against_fp64(ctx):
    """Mixed precision: rho, |grad rho| and s in FP32 against the FP64 path
    in the NCI region (s < 2): rho to 1e-5 (max, not RMS); s within a fixed
    multiple of the cancellation bound (cancellation_bound: the atom terms of
    the gradient cancel near critical points and in the low-density tail)
    and within 1e-5 wherever that bound is below 1e-6; below s = 0.1, s
    absolutely to 1e-5."""
    cell = random_cell(32, (8.0, 8.0, 8.0), (90, 90, 90), 8, (1, 6, 7, 8), (48, 48, 48))
    tables, species = _slater_setup(cell)
    c64, s64 = gn.nciplot_promol(ctx, cell.n, cell.x2c, tables, cell.xat, species, precision="fp64")
    c32, s32 = gn.nciplot_promol(ctx, cell.n, cell.x2c, tables, cell.xat, species, precision="fp32")
    r = gn.precision_report(c32, s32, c64, s64, zeta_max=2.0 * 8 ** (1.0 / 3.0))
    print(f"  FP32 vs FP64 in s<2 ({r['points']} pts): rho max {r['rho_max_rel']:.2e} rms {r['rho_rms_rel']:.2e}; "
          f"|grad rho| max {r['gradmod_max_rel']:.2e}; s (s>=0.1) max {r['s_max_rel']:.2e} rms "
          f"{r['s_rms_rel']:.2e}; s<0.1 ({r['s_below_floor_points']} pts) max abs {r['s_below_floor_max_abs']:.2e}")
    assert r["points"] > 1000
    assert r["rho_max_rel"] < 1e-5
    assert r["s_over_bound_max"] < 16.0  # about a dozen FP32 roundings per atom term
    assert r["s_max_rel_well_conditioned"] < 1e-5
    assert r["s_below_floor_max_abs"] < 1e-5


@pytest.mark.parametrize("grid", [(24, 30, 36), (25, 27, 31), (24, 27, 32)])
@pytest.mark.parametrize("kind,name", [(_lib.FT_GRAD, "grad"), (_lib.FT_XX, "xx"), (_lib.FT_YY, "yy"),
                                       (_lib.FT_ZZ, "zz"), (_lib.FT_LAP, "lap")])
def test_fft_deriv_matches_oracle(ctx, kind, name, grid):
    """Even grids (a Nyquist plane on every axis), odd grids (none) and mixed."""
    cell = random_cell(33, (7.0, 8.0, 9.0), (80, 100, 110), 4, (1, 6, 7, 8), grid)
    f = gb.synth_slater(ctx, cell)
    g = gb.fft_deriv(ctx, f, cell.n, cell.x2c, kind)
    o = oracle.fft_deriv(f, cell.n, cell.x2c, name)
    np.testing.assert_allclose(g, o, rtol=1e-9, atol=1e-9 * np.abs(o).max())


@pytest.mark.parametrize("which", ["cubic64", "triclinic_odd"])
def test_nci_fft_scheme_matches_oracle(ctx, which):
    if which == "cubic64":
        cell = config(1)
    else:
        cell = random_cell(64, (6.5, 7.5, 8.5), (75, 95, 115), 5, (1, 6, 7, 8), (27, 25, 31))
    print(f"[case] nci fft scheme {which} n={cell.n}", flush=True)
    f = gb.synth_slater(ctx, cell)
    crho, s = gn.nciplot_grid(ctx, f, cell.n, cell.x2c, zfast=False)
    cr_ref, s_ref = oracle.nci_grid_fft(f, cell.n, cell.x2c)
    np.testing.assert_allclose(np.abs(crho), np.abs(cr_ref), rtol=1e-15)
    # rocFFT and the oracle's FFT differ by round-off, about 1e-15 of the
    # largest gradient; s amplifies it where rho is small.  NCI region
    # (s < 2): 1e-6 relative; elsewhere the implied gradient error is bounded.
    ok = f > 1e-8
    nci = ok & (s_ref < 2.0)
    np.testing.assert_allclose(s[nci], s_ref[nci], rtol=1e-6)
    cst = 2 * (3 * math.pi ** 2) ** (1 / 3)
    g, g_ref = s * cst * f ** (4 / 3), s_ref * cst * f ** (4 / 3)
    assert np.abs(g - g_ref)[ok].max() <= 1e-12 * g_ref.max()
    # the lambda2 sign may differ only where a diagonal Hessian element is
    # round-off close to zero
    assert np.count_nonzero(np.sign(crho) != np.sign(cr_ref)) <= max(2, 1e-4 * f.size)
    # z-fastest layout of the reference's crho/cgrad arrays
    cz, sz = gn.nciplot_grid(ctx, f, cell.n, cell.x2c, zfast=True)
    n = cell.n
    assert np.array_equal(cz.reshape(n[0], n[1], n[2]), crho.reshape(n[2], n[1], n[0]).transpose(2, 1, 0))


@pytest.mark.parametrize("which", ["cubic64", "triclinic", "triclinic_odd"])
def test_nci_grid_interpolation_scheme_matches_oracle(ctx, which):
    """INTERPOLATION GRID (nci@proc.f90:565-571): tricubic value, gradient
    and Hessian at the node as field%grd locates it.  The node search is
    reproduced bit for bit (so the same cell's polynomial is evaluated; on
    the cell's origin planes, rounding below zero folds to 1.0 and both
    evaluate that cell's cubic at u = n, as the reference does); the
    polynomial itself is evaluated in a different but equivalent form, so
    values agree to round-off: 1e-8 relative (the worst case is the u = n
    extrapolation), and the lambda2 sign everywhere."""
    if which == "cubic64":
        cell = config(1)
    elif which == "triclinic":
        cell = random_cell(61, (7.0, 7.5, 8.0), (80, 100, 110), 4, (1, 6, 7, 8), (24, 26, 28))
    else:
        cell = random_cell(63, (6.5, 7.5, 8.5), (75, 95, 115), 5, (1, 6, 7, 8), (23, 25, 29))
    print(f"[case] nci grid scheme {which} n={cell.n}", flush=True)
    f = gb.synth_slater(ctx, cell)
    crho, s = gn.nciplot_grid(ctx, f, cell.n, cell.x2c, interpolation="grid")
    cr_ref, s_ref = oracle.nci_grid_interp(f, cell.n, cell.x2c)
    np.testing.assert_allclose(np.abs(crho), np.abs(cr_ref), rtol=1e-8)
    np.testing.assert_allclose(s, s_ref, rtol=1e-8)
    assert np.array_equal(np.sign(crho), np.sign(cr_ref))
    # x-fastest layout is the same data
    cx, sx = gn.nciplot_grid(ctx, f, cell.n, cell.x2c, interpolation="grid", zfast=False)
    n = cell.n
    assert np.array_equal(cx.reshape(n[2], n[1], n[0]).transpose(2, 1, 0), crho.reshape(n[0], n[1], n[2]))


def test_nci_grid_scheme_rejects_fp32_and_tiny_grids(ctx):
    cell = random_cell(62, (5.0, 5.0, 5.0), (90, 90, 90), 2, (1, 8), (8, 8, 8))
    f = gb.synth_slater(ctx, cell)
    with pytest.raises(_lib.C2GpuError):
        gn._run(ctx, f, cell.n, cell.x2c, _lib.NCI_GRID, _lib.FP32, None, True, False)
    with pytest.raises(_lib.C2GpuError):
        gn.nciplot_grid(ctx, np.ones(8), (8, 1, 1), cell.x2c, interpolation="grid")


@pytest.mark.parametrize("dtype,device,opts", [
    (np.float64, False, {}),
    (np.float64, True, {"onlyneg": True, "srhorange": (-0.3, 0.2)}),
    (np.float32, False, {"rho_void": 0.05, "rhoplot": 0.2}),
    (np.float32, True, {"rho_void": 0.02, "rhocut": 0.1, "dimcut": 1.0}),
])
def test_cutoff_pass_matches_oracle(ctx, dtype, device, opts):
    """c2gpu_nci_cutoff against the oracle's restatement: the same rows in
    node order (each field the correctly rounded 1p,E15.7 of the oracle's
    value), the same masked RDG, the void sums."""
    rng = np.random.default_rng(3)
    npts = 20000
    crho = (rng.standard_normal(npts) * 10.0 ** rng.uniform(-5, 1.5, npts)).astype(dtype)
    cg = (rng.uniform(0, 3, npts) * 10.0 ** rng.uniform(-3, 0.5, npts)).astype(dtype)
    rhoat = (10.0 ** rng.uniform(-4, 0, npts)).astype(dtype)
    kw = {"rhocut": 0.2, "dimcut": 2.0, "rhoplot": 0.1, "srhorange": (-1e30, 1e30), "onlyneg": False,
          "rho_void": -1.0}
    kw.update(opts)
    dat, cgo, sums_o = oracle.nci_cutoffs(crho.astype(np.float64), cg.astype(np.float64), rhoat=rhoat, **kw)
    want = "".join(fortran_e(float(a), 3) + fortran_e(float(b), 3) + "\n" for a, b in dat).encode()
    g = cg.copy()
    if device:
        dc, dg, da = ctx.to_device(crho), ctx.to_device(g), ctx.to_device(rhoat)
        text, sums = gn.cutoffs(ctx, dc, dg, rhoat=da, **kw)
        g = dg.numpy()
    else:
        text, sums = gn.cutoffs(ctx, crho, g, rhoat=rhoat, **kw)
    assert text == want
    assert np.array_equal(g.astype(np.float64), cgo)
    np.testing.assert_allclose(sums, sums_o, rtol=1e-12)
