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
    (74, (7.0, 7.5, 8.0), (85, 95, 100), (24, 26, 28), 0, False),   # monopole only
    (75, (7.5, 8.0, 7.0), (90, 90, 90), (26, 24, 28), 10, True),    # the largest lmax
    (76, (5.0, 6.0, 7.0), (80, 100, 95), (7, 11, 13), 3, False),    # small odd grid
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



def test_monopole_is_basin_sum(ctx):
    """lmax = 0: Q_00 of each basin is its plain sum of f (R_00 = 1)."""
    cell = random_cell(77, (8.0, 8.0, 8.5), (90, 90, 90), 4, (1, 6, 7, 8), (28, 28, 30))
    f = gb.synth_slater(ctx, cell)
    bas = gb.bader_integrate(ctx, f, cell.n, cell.x2c, cell.xat)
    q = gb.multipoles(ctx, bas.idg, f, cell.n, cell.x2c, bas.xattr[:bas.nattr], 0)
    _, sums = gb.basin_sums(ctx, bas.idg, bas.nattr, [f])
    np.testing.assert_allclose(q[:, 0], np.asarray(sums).reshape(-1), rtol=1e-12)


@pytest.mark.parametrize("seed,edge,ang,nat,n,lmax,device", [
    (81, (8.0, 8.5, 9.0), (90, 90, 90), 4, (32, 34, 36), 5, False),
    (82, (8.5, 8.0, 9.5), (80, 100, 105), 5, (34, 32, 36), 3, True),
    (83, (9.0, 9.5, 10.0), (85, 95, 100), 40, (40, 44, 48), 2, False),  # surface points with > 32 basins
    (84, (7.0, 7.5, 8.0), (85, 95, 100), 4, (24, 26, 28), 0, False),   # monopole only
])
def test_yt_multipoles_match_oracle(ctx, seed, edge, ang, nat, n, lmax, device):
    """YT multipoles (interior points weight 1, surface points their
    yt_weights weight) against the oracle's loop over every basin's weight
    field; the monopole is the YT sum of f."""
    from critic2_amd import _lib
    from critic2_amd import yt as gy
    cell = random_cell(seed, edge, ang, nat, (1, 6, 7, 8), n)
    print(f"[case] yt multipoles seed={seed} n={cell.n} atoms={nat} lmax={lmax} device={device}", flush=True)
    f = gb.synth_slater(ctx, cell)
    vec, area = _lib.ws_facets(np.asarray(cell.x2c) / np.array(cell.n)[None, :])
    idg, nattr, xattr, vol, sums, w = oracle.yt(f, cell.n, cell.x2c, cell.xat, vec, area, fields=[f],
                                                weights=True)
    xa = np.asarray(xattr).reshape(-1, 3)[:nattr]
    ref = oracle.yt_multipoles(w, f, cell.n, cell.x2c, xa, lmax)
    if device:
        df = ctx.to_device(f)
        got_yt = gy.yt_integrate(ctx, df, cell.n, cell.x2c, cell.xat, fields=[df])
        got = gy.yt_multipoles(ctx, got_yt["labels"], df, cell.n, cell.x2c, got_yt["xattr"], lmax)
    else:
        got_yt = gy.yt_integrate(ctx, f, cell.n, cell.x2c, cell.xat, fields=[f])
        got = gy.yt_multipoles(ctx, got_yt["labels"], f, cell.n, cell.x2c, got_yt["xattr"], lmax)
    assert got_yt["nattr"] == nattr
    for l in range(lmax + 1):
        s = slice(l * l, (l + 1) * (l + 1))
        scale = np.max(np.abs(ref[:, s]), axis=1, keepdims=True)
        err = np.abs(got[:, s] - ref[:, s]) / np.maximum(scale, 1e-300)
        assert err.max() < 1e-10, (l, err.max())
    np.testing.assert_allclose(got[:, 0], got_yt["sums"][0], rtol=1e-12)


def test_yt_multipoles_need_a_yt_call(ctx):
    """Without a c2gpu_yt call on the context, or with another grid, the
    entry point refuses."""
    from critic2_amd import _lib
    from critic2_amd import yt as gy
    fresh = _lib.Context(0)
    try:
        lab = np.ones(8 * 8 * 8, dtype=np.int32)
        f = np.ones(8 * 8 * 8)
        with pytest.raises(_lib.C2GpuError) as e:
            gy.yt_multipoles(fresh, lab, f, (8, 8, 8), np.eye(3) * 10.0, np.zeros((1, 3)), 2)
        assert e.value.code == 1
        cell = random_cell(85, (7.0, 7.0, 7.0), (90, 90, 90), 2, (1, 6), (20, 22, 24))
        g = gb.synth_slater(fresh, cell)
        r = gy.yt_integrate(fresh, g, cell.n, cell.x2c, cell.xat)
        with pytest.raises(_lib.C2GpuError):
            gy.yt_multipoles(fresh, r["labels"][:20 * 22 * 22], g[:20 * 22 * 22], (20, 22, 22), cell.x2c,
                             r["xattr"], 2)
    finally:
        fresh.close()


@pytest.mark.parametrize("nattr,lmax,device", [
    (600, 5, False),    # 600 * 36 bins: 169 KiB, more than the LDS holds
    (130, 10, True),    # 130 * 121 bins: 123 KiB, just past it
    (106, 10, False),   # 100 KiB: the LDS bins, for the same labels
])
def test_multipoles_many_basins(ctx, nattr, lmax, device):
    """More basin bins than fit in the LDS: each workgroup accumulates into
    a row of global memory. Labels are 4^3 bricks dealt to nattr basins
    (some basins get no point), attractors at random positions."""
    rng = np.random.default_rng(nattr + lmax)
    n = (40, 36, 44)
    x2c = np.array([[9.0, 0.4, -0.3], [0.0, 8.5, 0.6], [0.0, 0.0, 9.5]])
    i, j, k = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    brick = (i // 4) + 10 * (j // 4) + 100 * (k // 4)
    deal = rng.integers(1, nattr + 1, size=int(brick.max()) + 1)
    idg = deal[brick].transpose(2, 1, 0).reshape(-1).astype(np.int32)   # x fastest
    f = rng.random(idg.size) + 0.1
    xa = rng.random((nattr, 3))
    print(f"[case] multipoles bricks nattr={nattr} lmax={lmax} device={device}", flush=True)
    ref = oracle.multipoles(idg, f, n, x2c, xa, lmax)
    if device:
        got = gb.multipoles(ctx, ctx.to_device(idg), ctx.to_device(f), n, x2c, xa, lmax)
    else:
        got = gb.multipoles(ctx, idg, f, n, x2c, xa, lmax)
    assert got.shape == ref.shape
    for l in range(lmax + 1):
        s = slice(l * l, (l + 1) * (l + 1))
        scale = np.maximum(np.max(np.abs(ref[:, s]), axis=1, keepdims=True), 1e-300)
        err = np.abs(got[:, s] - ref[:, s]) / scale
        assert err.max() < 1e-10, (l, err.max())
    empty = np.setdiff1d(np.arange(1, nattr + 1), idg)
    assert np.all(got[empty - 1] == 0.0)


def test_yt_multipoles_many_basins(ctx):
    """YT multipoles of 140 basins at lmax 10 (138 KiB of bins, past the
    LDS): interior and surface kernels both accumulate in global rows."""
    from critic2_amd import _lib
    from critic2_amd import yt as gy
    cell = random_cell(91, (13.0, 13.5, 14.0), (85, 95, 100), 140, (1, 6, 7, 8), (56, 56, 60))
    print(f"[case] yt multipoles seed=91 n={cell.n} atoms=140 lmax=10", flush=True)
    f = gb.synth_slater(ctx, cell)
    vec, area = _lib.ws_facets(np.asarray(cell.x2c) / np.array(cell.n)[None, :])
    idg, nattr, xattr, vol, sums, w = oracle.yt(f, cell.n, cell.x2c, cell.xat, vec, area, fields=[f],
                                                weights=True)
    assert nattr * 121 * 8 > 120 * 1024
    xa = np.asarray(xattr).reshape(-1, 3)[:nattr]
    ref = oracle.yt_multipoles(w, f, cell.n, cell.x2c, xa, 10)
    got_yt = gy.yt_integrate(ctx, f, cell.n, cell.x2c, cell.xat, fields=[f])
    assert got_yt["nattr"] == nattr
    got = gy.yt_multipoles(ctx, got_yt["labels"], f, cell.n, cell.x2c, got_yt["xattr"], 10)
    for l in range(11):
        s = slice(l * l, (l + 1) * (l + 1))
        scale = np.maximum(np.max(np.abs(ref[:, s]), axis=1, keepdims=True), 1e-300)
        err = np.abs(got[:, s] - ref[:, s]) / scale
        assert err.max() < 1e-10, (l, err.max())
    np.testing.assert_allclose(got[:, 0], got_yt["sums"][0], rtol=1e-12)
