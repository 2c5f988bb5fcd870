"""GPU Yu-Trinkle integration (c2gpu_yt) against the oracle's restatement of
yt_integrate / yt_weights and the imtype_yt sums: labels bit-exact, the same
attractors, weight sums within 1e-10 relative."""
import numpy as np
import pytest

from critic2_amd import _lib
from critic2_amd import bader as gb
from critic2_amd import yt as gy
from critic2_amd.synth import random_cell
from oracle import oracle

pytestmark = pytest.mark.gpu


def _run_both(ctx, cell, f, mode_kw, device=False):
    vec, area = _lib.ws_facets(np.asarray(cell.x2c) / np.array(cell.n)[None, :])
    mode, ra = gb._mode(mode_kw.get("nnm", False), mode_kw.get("ratom"), mode_kw.get("noatoms", False))
    ref = oracle.yt(f, cell.n, cell.x2c, cell.xat, vec, area, fields=[f, f * f], mode=mode, ratom=ra)
    if device:
        df = ctx.to_device(f)
        df2 = ctx.to_device(f * f)
        got = gy.yt_integrate(ctx, df, cell.n, cell.x2c, cell.xat, fields=[df, df2], **mode_kw)
        got["labels"] = got["labels"].numpy()
    else:
        got = gy.yt_integrate(ctx, f, cell.n, cell.x2c, cell.xat, fields=[f, f * f], **mode_kw)
    return ref, got


@pytest.mark.parametrize("seed,edge,ang,nat,n,mode_kw,device", [
    (61, (8.0, 8.5, 9.0), (90, 90, 90), 4, (36, 40, 44), {}, False),
    (62, (8.5, 8.0, 9.5), (80, 100, 105), 5, (40, 36, 44), {}, True),
    (63, (9.0, 9.0, 9.0), (75, 95, 110), 6, (44, 44, 40), {"nnm": True, "ratom": 1.0}, False),
    (64, (8.0, 9.0, 8.5), (88, 92, 97), 4, (36, 40, 40), {"noatoms": True, "ratom": 1.0}, False),
])
def test_yt_matches_oracle(ctx, seed, edge, ang, nat, n, mode_kw, device):
    cell = random_cell(seed, edge, ang, nat, (1, 6, 7, 8), n)
    print(f"[case] yt seed={seed} n={cell.n} mode={mode_kw} device={device}", flush=True)
    f = gb.synth_slater(ctx, cell)
    ref, got = _run_both(ctx, cell, f, mode_kw, device)
    idg, nattr, xattr, vol, sums = ref
    assert got["nattr"] == nattr
    np.testing.assert_array_equal(got["xattr"], xattr)
    assert int(np.count_nonzero(got["labels"] != idg)) == 0
    np.testing.assert_allclose(got["vol"], vol, rtol=1e-10)
    np.testing.assert_allclose(got["sums"], sums, rtol=1e-10)
    st = got["stats"]
    assert st["nias"] == int(np.count_nonzero(idg == 0))
    print(f"  nattr {nattr} surface points {st['nias']} label passes {st['label_passes']} "
          f"weight passes {st['weight_passes']}", flush=True)


def test_yt_many_basins_meet_wide_path(ctx):
    """40 atoms in a small cell: surface points where more than 32 basins
    meet take the workgroup accumulator; still the oracle's weights."""
    cell = random_cell(66, (9.0, 9.5, 10.0), (85, 95, 100), 40, (1, 6, 7, 8), (40, 44, 48))
    print(f"[case] yt wide n={cell.n} atoms=40", flush=True)
    f = gb.synth_slater(ctx, cell)
    ref, got = _run_both(ctx, cell, f, {})
    idg, nattr, xattr, vol, sums = ref
    assert got["nattr"] == nattr
    assert int(np.count_nonzero(got["labels"] != idg)) == 0
    np.testing.assert_allclose(got["vol"], vol, rtol=1e-10)
    np.testing.assert_allclose(got["sums"], sums, rtol=1e-10)
    print(f"  wide points {got['stats']['wide']}", flush=True)
    assert got["stats"]["wide"] > 0


def test_yt_thousands_of_basins(ctx):
    """Smoothed noise with a tiny ratom: 1858 maxima, each its own basin,
    and four fields, so the surface bins (5 x 1858 doubles) pass the LDS
    budget and add into the global accumulator directly."""
    rng = np.random.default_rng(5)
    n = (36, 36, 36)
    x2c = np.diag([12.0, 12.0, 12.0])
    g = rng.random(n)
    for ax in range(3):
        g = (g + np.roll(g, 1, ax) + np.roll(g, -1, ax)) / 3
    f = np.exp(3 * g).transpose(2, 1, 0).reshape(-1)
    fl = [f, f * f, np.sqrt(f), np.ones_like(f)]
    xat = np.zeros((1, 3))
    print(f"[case] yt noise n={n} ratom=0.01 fields=4", flush=True)
    vec, area = _lib.ws_facets(x2c / np.array(n)[None, :])
    idg, nattr, xattr, vol, sums = oracle.yt(f, n, x2c, xat, vec, area, fields=fl, mode=_lib.ATTR_RATOM,
                                             ratom=0.01)
    assert 8 * 5 * nattr > 60000
    got = gy.yt_integrate(ctx, f, n, x2c, xat, fields=fl, ratom=0.01)
    assert got["nattr"] == nattr
    np.testing.assert_array_equal(got["xattr"], xattr)
    assert int(np.count_nonzero(got["labels"] != idg)) == 0
    np.testing.assert_allclose(got["vol"], vol, rtol=1e-10)
    np.testing.assert_allclose(got["sums"], sums, rtol=1e-10)


@pytest.mark.parametrize("digits,seed,n,ranked", [(3, 65, (24, 26, 28), 1), (6, 67, (32, 34, 36), None),
                                                   (2, 68, (20, 22, 24), 1)])
def test_yt_equal_values_follow_the_reference_sort(ctx, digits, seed, n, ranked):
    """Grids rounded to a few significant digits, as cube files print them
    (ties between neighbours, between the higher neighbours of one point and
    between maxima): the points are ranked by the reference's quicksort and
    the labels, attractors and weight sums equal the oracle's, which ranks
    them with its own restatement of that sort."""
    cell = random_cell(seed, (8.0, 8.5, 9.0), (90, 95, 100), 4, (1, 6, 7, 8), n)
    print(f"[case] yt ties digits={digits} n={cell.n}", flush=True)
    f = gb.synth_slater(ctx, cell)
    f = np.array([float(f"{v:.{digits}g}") for v in f])
    ref, got = _run_both(ctx, cell, f, {})
    idg, nattr, xattr, vol, sums = ref
    print(f"  ranked {got['stats']['ranked']} nattr {nattr}", flush=True)
    if ranked is not None:  # 6 digits: ties only where they may not matter
        assert got["stats"]["ranked"] == ranked
    assert got["nattr"] == nattr
    np.testing.assert_array_equal(got["xattr"], xattr)
    assert int(np.count_nonzero(got["labels"] != idg)) == 0
    np.testing.assert_allclose(got["vol"], vol, rtol=1e-10)
    np.testing.assert_allclose(got["sums"], sums, rtol=1e-10)


def test_yt_plateau_and_tied_maxima(ctx):
    """A flat plateau with two equal maxima: every order decision is a tie,
    decided by the quicksort's ranks on both sides."""
    cell = random_cell(69, (8.0, 8.0, 8.0), (90, 90, 90), 2, (1, 6), (16, 16, 16))
    f = np.ones(16 ** 3)
    f[100] = 2.0
    f[2000] = 2.0
    f[3000] = 1.5
    ref, got = _run_both(ctx, cell, f, {"noatoms": True, "ratom": 0.5})
    idg, nattr, xattr, vol, sums = ref
    assert got["stats"]["ranked"] == 1
    assert got["nattr"] == nattr
    np.testing.assert_array_equal(got["xattr"], xattr)
    assert int(np.count_nonzero(got["labels"] != idg)) == 0
    np.testing.assert_allclose(got["vol"], vol, rtol=1e-10)
    np.testing.assert_allclose(got["sums"], sums, rtol=1e-10)


def test_yt_distinct_values_need_no_ranking(ctx):
    cell = random_cell(61, (8.0, 8.5, 9.0), (90, 90, 90), 4, (1, 6, 7, 8), (24, 26, 28))
    f = gb.synth_slater(ctx, cell)
    got = gy.yt_integrate(ctx, f, cell.n, cell.x2c, cell.xat)
    assert got["stats"]["ranked"] == 0


YT_EDGE = [  # name, seed, cell edges (A), angles, atoms, grid
    ("tiny-3x5x7", 71, (4.0, 5.0, 6.0), (90, 90, 90), 2, (3, 5, 7)),
    ("odd-17x3x11", 72, (7.0, 4.0, 6.0), (80, 100, 95), 3, (17, 3, 11)),
    ("prime-13x19x23", 73, (7.0, 8.0, 9.0), (85, 92, 101), 4, (13, 19, 23)),
    ("cube-5", 74, (5.0, 5.0, 5.0), (90, 90, 90), 1, (5, 5, 5)),
]


@pytest.mark.parametrize("name,seed,edge,ang,nat,n", YT_EDGE, ids=[e[0] for e in YT_EDGE])
def test_yt_edge_grids(ctx, name, seed, edge, ang, nat, n):
    """The smallest grids (3 points on an axis, where opposite lattice
    neighbours wrap onto nearby points), odd and prime sizes, against the
    oracle."""
    cell = random_cell(seed, edge, ang, nat, (1, 6, 7, 8), n)
    print(f"[case] yt {name} n={cell.n}", flush=True)
    f = gb.synth_slater(ctx, cell)
    ref, got = _run_both(ctx, cell, f, {})
    idg, nattr, xattr, vol, sums = ref
    assert got["nattr"] == nattr
    np.testing.assert_array_equal(got["xattr"], xattr)
    assert int(np.count_nonzero(got["labels"] != idg)) == 0
    np.testing.assert_allclose(got["vol"], vol, rtol=1e-10)
    np.testing.assert_allclose(got["sums"], sums, rtol=1e-10)
