"""Known critical points as attractors and int_reorder_gridout on real
non-nuclear-maxima runs (integration@proc.f90:1027-1166): labels from
c2gpu_bader / c2gpu_yt with the field's other critical points preloaded,
remapped on the GPU by c2gpu_relabel (and the YT surface weights by
c2gpu_yt_remap), against the oracle's sweep, its restatement of the reorder
and the `where` loop; then Volume/Pop sums and YT multipoles on the new
attractors."""
import numpy as np
import pytest

from critic2_amd import _lib
from critic2_amd import bader as gb
from critic2_amd import yt as gy
from critic2_amd.synth import random_cell
from oracle import oracle

pytestmark = pytest.mark.gpu

RATOM = 0.05  # bohr: below every atom's distance to its grid maximum here


def _cells():
    return [random_cell(91, (8.0, 8.5, 9.0), (90, 90, 90), 4, (1, 6, 7, 8), (32, 34, 36)),
            random_cell(92, (8.5, 8.0, 9.0), (80, 100, 105), 5, (1, 6, 7, 8), (34, 32, 36))]


def _critical_points(f, cell, nat):
    """A bond-type point (typ -1) and a maximum-type point (typ -3) placed
    on the grid maxima of the first two atoms (found by a run without them):
    the sweep assigns those maxima to the preloaded points (are_lclose), and
    the reorder then renumbers the bond point's basin."""
    _, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat, mode=1, ratom=RATOM)
    nnm = np.asarray(xattr)[nat:info["nattr"]]
    assert len(nnm) >= 2
    return nnm[:2].copy(), [-3] * nat + [-1, -3]


@pytest.mark.parametrize("q", [0, 1])
def test_bader_known_points_and_reorder(ctx, q):
    cell = _cells()[q]
    nat = len(cell.zat)
    print(f"[case] reorder bader cell {q} n={cell.n} atoms={nat}", flush=True)
    f = gb.synth_slater(ctx, cell)
    cps, cptyp = _critical_points(f, cell, nat)
    cpcel = np.concatenate([np.asarray(cell.xat), cps])
    for mode_kw, mode, ra in (({}, 0, 1e40), ({"ratom": RATOM}, 1, RATOM)):
        bas = gb.bader_integrate(ctx, f, cell.n, cell.x2c, cell.xat, cps=cps, **mode_kw)
        ref, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat, mode=mode, ratom=ra, cps=cps)
        assert int(np.count_nonzero(bas.idg != ref)) == 0, mode_kw
        assert bas.nattr == info["nattr"]
        np.testing.assert_array_equal(bas.xattr, xattr)
        if mode == 0:
            assert ref.max() <= nat  # the nearest-atom rule matches atoms only
            continue
        # the reorder: relabel on the GPU (device-resident labels)
        lab = ctx.to_device(bas.idg.astype(np.int32))
        r = gb.reorder_gridout(ctx, lab, bas.xattr, cell.x2c, cpcel, cptyp, nat, ra)
        assigned, nattr, icp, xnew = oracle.reorder_gridout(xattr, cell.x2c, cpcel, cptyp, nat, ra)
        assert r["assigned"].tolist() == assigned.tolist()
        assert np.any(assigned != np.arange(1, len(assigned) + 1))
        want = oracle.relabel(ref, assigned)
        assert int(np.count_nonzero(want != ref)) > 0  # the map moves real points
        got = lab.numpy()
        assert int(np.count_nonzero(got != want)) == 0
        # sums on the new attractors
        c_ref, s_ref = oracle.basin_sums(want, max(nattr, int(want.max())), [f])
        c_gpu, s_gpu = gb.basin_sums(ctx, got, max(nattr, int(want.max())), [f])
        np.testing.assert_array_equal(c_gpu, c_ref)
        np.testing.assert_allclose(s_gpu, s_ref, rtol=1e-10)
        lab.free()
        print(f"  assigned {assigned.tolist()} nattr {nattr} icp {icp.tolist()}", flush=True)


def test_yt_reorder_and_multipoles(ctx):
    """YT with the same preloaded points; after the reorder the multipoles
    over the remapped weights equal the oracle's over the merged per-basin
    weight fields (yt_weights is linear in the basin indicators, so merged
    basins' weights add), with the new attractor positions."""
    cell = _cells()[0]
    nat = len(cell.zat)
    print(f"[case] reorder yt n={cell.n}", flush=True)
    f = gb.synth_slater(ctx, cell)
    cps, cptyp = _critical_points(f, cell, nat)
    cpcel = np.concatenate([np.asarray(cell.xat), cps])
    vec, area = _lib.ws_facets(np.asarray(cell.x2c) / np.array(cell.n)[None, :])
    idg, nattr0, xattr, vol, sums, w = oracle.yt(f, cell.n, cell.x2c, cell.xat, vec, area, fields=[f], mode=1,
                                                 ratom=RATOM, weights=True, cps=cps)
    got = gy.yt_integrate(ctx, f, cell.n, cell.x2c, cell.xat, fields=[f], ratom=RATOM, cps=cps)
    assert got["nattr"] == nattr0 and int(np.count_nonzero(got["labels"] != idg)) == 0
    lab = got["labels"]
    r = gb.reorder_gridout(ctx, lab, got["xattr"], cell.x2c, cpcel, cptyp, nat, RATOM, yt=True)
    assigned, nattr, icp, xnew = oracle.reorder_gridout(xattr, cell.x2c, cpcel, cptyp, nat, RATOM)
    assert r["assigned"].tolist() == assigned.tolist() and r["nattr"] == nattr
    np.testing.assert_array_equal(lab, oracle.relabel(idg, assigned))
    nb = max(nattr, int(assigned.max()))
    w2 = np.zeros((nb, w.shape[1]))
    for i, a in enumerate(assigned):
        w2[a - 1] += w[i]
    xa2 = np.zeros((nb, 3))
    xa2[:nattr] = xnew
    lmax = 3
    ref = oracle.yt_multipoles(w2, f, cell.n, cell.x2c, xa2, lmax)
    if nb > nattr:  # a label beyond the new count (the reference's cpcel ids) needs a slot
        pytest.skip("labels beyond nattr after the reorder")
    q = gy.yt_multipoles(ctx, lab, f, cell.n, cell.x2c, xnew, lmax)
    for l in range(lmax + 1):
        s = slice(l * l, (l + 1) * (l + 1))
        scale = np.max(np.abs(ref[:, s]), axis=1, keepdims=True)
        err = np.abs(q[:, s] - ref[:, s]) / np.maximum(scale, 1e-300)
        assert err.max() < 1e-10, (l, err.max())
    # monopoles: the YT sums merged through the map
    merged = np.zeros(nb)
    np.add.at(merged, assigned - 1, got["sums"][0])
    np.testing.assert_allclose(q[:, 0], merged[:nattr], rtol=1e-11)
