"""CPU checks of the Yu-Trinkle pieces: the Wigner-Seitz facets of the grid
lattice (closed forms and the volume identity) and the oracle's restatement
of yt_integrate / yt_weights (partition of unity, one-basin limit)."""
import numpy as np
import pytest

from critic2_amd import _lib


def _vol_from_facets(m, vec, area):
    # the cell is the union of pyramids facet x (|v|/2): V = sum A |v| / 6
    return float(np.sum(area * np.linalg.norm(vec @ np.asarray(m).T, axis=1)) / 6.0)


def test_ws_orthorhombic_closed_form():
    vec, area = _lib.ws_facets(np.diag([1.0, 2.0, 3.0]))
    got = {tuple(int(x) for x in v): a for v, a in zip(vec, area)}
    want = {(1, 0, 0): 6.0, (-1, 0, 0): 6.0, (0, 1, 0): 3.0, (0, -1, 0): 3.0, (0, 0, 1): 2.0, (0, 0, -1): 2.0}
    assert got.keys() == want.keys()
    for k in want:
        assert got[k] == pytest.approx(want[k], rel=1e-13)


def test_ws_fcc_bcc_hexagonal_closed_forms():
    fcc = 0.5 * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], float).T
    vec, area = _lib.ws_facets(fcc)  # rhombic dodecahedron
    assert len(vec) == 12
    np.testing.assert_allclose(area, np.sqrt(2) / 8, rtol=1e-12)
    bcc = 0.5 * np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]], float).T
    vec, area = _lib.ws_facets(bcc)  # truncated octahedron: 6 squares, 8 hexagons
    assert len(vec) == 14
    a = np.sort(area)
    np.testing.assert_allclose(a[:6], 1 / 8, rtol=1e-12)
    np.testing.assert_allclose(a[6:], 3 * np.sqrt(3) / 16, rtol=1e-12)
    hexl = np.array([[1, 0, 0], [-0.5, np.sqrt(3) / 2, 0], [0, 0, 1.6]]).T
    vec, area = _lib.ws_facets(hexl)  # hexagonal prism
    assert len(vec) == 8
    a = np.sort(area)
    np.testing.assert_allclose(a[:2], 3 * np.sqrt(3) / 2 * (1 / np.sqrt(3)) ** 2, rtol=1e-12)
    np.testing.assert_allclose(a[2:], 1.6 / np.sqrt(3), rtol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_ws_triclinic_volume_identity(seed):
    rng = np.random.default_rng(seed)
    from critic2_amd.synth import lattice_from_params
    m = lattice_from_params(*rng.uniform(3, 6, 3), *rng.uniform(70, 110, 3)) / np.array([40, 44, 48])[None, :]
    vec, area = _lib.ws_facets(m)
    assert len(vec) in (6, 8, 10, 12, 14) and len(vec) % 2 == 0
    assert _vol_from_facets(m, vec, area) == pytest.approx(abs(np.linalg.det(m)), rel=1e-10)
    # facets come in +-v pairs of equal area
    d = {tuple(int(x) for x in v): a for v, a in zip(vec, area)}
    for v, a in d.items():
        assert d[tuple(-x for x in v)] == pytest.approx(a, rel=1e-12)


def test_oracle_yt_partition_of_unity_and_single_basin():
    from oracle import oracle
    from tests.helpers import small_cell
    from critic2_amd.synth import BOHR_PER_ANG
    cell = small_cell(5, (16, 18, 20), nat=3, ang=(80, 95, 105))
    f = oracle.slater_grid(cell.x2c, cell.n, cell.xat, cell.zat, 8.0 * BOHR_PER_ANG)
    vec, area = _lib.ws_facets(np.asarray(cell.x2c) / np.array(cell.n)[None, :])
    idg, nattr, xattr, vol, sums = oracle.yt(f, cell.n, cell.x2c, cell.xat, vec, area, fields=[f])
    N = f.size
    assert vol.sum() == pytest.approx(N, rel=1e-12)  # every point's weights sum to one
    assert sums[0].sum() == pytest.approx(f.sum(), rel=1e-12)
    assert np.all(vol >= np.bincount(idg, minlength=nattr + 1)[1:] - 1e-9)
    # one attractor: everything in it
    single = np.ones(N)
    x = np.indices(cell.n).reshape(3, -1).T / np.array(cell.n)
    d = x - np.array([0.5, 0.5, 0.5])
    single = np.exp(-np.sum((d @ np.asarray(cell.x2c).T) ** 2, axis=1) + 1e-9 * np.arange(N) / N)
    # x fastest: the indices above are (i, j, k) with i slowest -> reorder
    single = single.reshape(cell.n).transpose(2, 1, 0).reshape(-1)
    idg1, n1, _, vol1, s1 = oracle.yt(single, cell.n, cell.x2c, [[0.5, 0.5, 0.5]], vec, area, fields=[single])
    assert n1 == 1 and np.all(idg1 == 1)
    assert vol1[0] == pytest.approx(N, rel=1e-14)


def test_oracle_yt_weights_and_multipoles():
    """The per-basin weight fields behind the YT multipoles: 1 at each
    basin's interior points and 0 at other basins' interiors, a partition of
    unity at every point, their sums the YT volumes; the multipoles with the
    interior indicator as weights are the (independently pinned) Bader-branch
    multipoles of the labels, and the monopole is the YT sum of f."""
    from oracle import oracle
    from tests.helpers import small_cell
    from critic2_amd.synth import BOHR_PER_ANG
    cell = small_cell(6, (16, 18, 20), nat=3, ang=(80, 95, 105))
    f = oracle.slater_grid(cell.x2c, cell.n, cell.xat, cell.zat, 8.0 * BOHR_PER_ANG)
    vec, area = _lib.ws_facets(np.asarray(cell.x2c) / np.array(cell.n)[None, :])
    idg, nattr, xattr, vol, sums, w = oracle.yt(f, cell.n, cell.x2c, cell.xat, vec, area, fields=[f],
                                                weights=True)
    assert w.shape == (nattr, f.size)
    for a in range(nattr):
        assert np.all(w[a][idg == a + 1] == 1.0)
        assert np.all(w[a][(idg != a + 1) & (idg > 0)] == 0.0)
    np.testing.assert_allclose(w.sum(axis=0), 1.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(w.sum(axis=1), vol, rtol=1e-13)
    surf = idg == 0
    assert surf.any() and np.all(w[:, surf] >= 0.0)
    xa = np.asarray(xattr).reshape(-1, 3)[:nattr]
    lmax = 4
    q = oracle.yt_multipoles(w, f, cell.n, cell.x2c, xa, lmax)
    np.testing.assert_allclose(q[:, 0], sums[0], rtol=1e-12)
    ind = np.stack([(idg == a + 1).astype(np.float64) for a in range(nattr)])
    lab = np.where(surf, 0, idg).astype(np.int32)
    np.testing.assert_allclose(oracle.yt_multipoles(ind, f, cell.n, cell.x2c, xa, lmax),
                               oracle.multipoles(lab, f, cell.n, cell.x2c, xa, lmax), rtol=1e-12, atol=1e-12)


def test_qcksort_hand_traced():
    """The oracle's qcksort_r8 (tools@proc.f90:83-168) on arrays traced by
    hand.  Five elements: insertion sort only (keeps the index order of
    equal keys).  Eight elements: one partition, fx = 1663, pivot position
    int(8 * 1663 * 1.2698413e-4) = 1 (value 1), the hole scheme moves index 5
    to the front and leaves [5, 1 | 2, 3, 4, 0, 6, 7]; the right part is then
    insertion-sorted to [3, 6, 2, 4, 0, 7]: the tie order of the 1s and the
    3s is not the index order."""
    from oracle import oracle
    assert oracle.qcksort_order([2, 1, 2, 1, 2]).tolist() == [1, 3, 0, 2, 4]
    assert oracle.qcksort_order([3, 1, 3, 2, 3, 1, 2, 3]).tolist() == [5, 1, 3, 6, 2, 4, 0, 7]


@pytest.mark.parametrize("n,levels", [(1000, 7), (20000, 50), (200000, 1000)])
def test_qcksort_sorts_with_ties(n, levels):
    """Larger arrays with many ties (several partitions, the pivot sequence
    advancing): a permutation that sorts ascending."""
    from oracle import oracle
    rng = np.random.default_rng(n)
    a = rng.integers(0, levels, n).astype(np.float64)
    io = oracle.qcksort_order(a)
    assert sorted(io.tolist()) == list(range(n))
    assert np.all(np.diff(a[io]) >= 0)
    # equal keys are not left in index order (the sort is not stable)
    assert not all(np.all(np.diff(io[a[io] == v]) > 0) for v in range(levels))


def test_oracle_yt_ranks_equal_values_by_qcksort():
    """A grid quantised to 3 significant digits (ties everywhere): the YT
    oracle runs (it no longer refuses), its weights still form a partition
    of unity, and its labels follow the quicksort's ranks: a point is a
    maximum exactly when no Voronoi neighbour is ranked above it."""
    from oracle import oracle
    from tests.helpers import small_cell
    from critic2_amd.synth import BOHR_PER_ANG
    cell = small_cell(7, (14, 16, 18), nat=3, ang=(90, 90, 90))
    f = oracle.slater_grid(cell.x2c, cell.n, cell.xat, cell.zat, 8.0 * BOHR_PER_ANG)
    f = np.array([float(f"{v:.3g}") for v in f])
    vec, area = _lib.ws_facets(np.asarray(cell.x2c) / np.array(cell.n)[None, :])
    idg, nattr, xattr, vol, sums = oracle.yt(f, cell.n, cell.x2c, cell.xat, vec, area, fields=[f])
    assert vol.sum() == pytest.approx(f.size, rel=1e-12)
    io = oracle.qcksort_order(f)
    rk = np.empty(f.size, dtype=np.int64)
    rk[io] = np.arange(f.size)
    n = np.array(cell.n)
    ijk = np.stack(np.unravel_index(np.arange(f.size), cell.n, order="F"), axis=1)
    above = np.zeros(f.size, dtype=bool)
    for v in np.asarray(vec).reshape(-1, 3):
        nb = (ijk + v) % n
        jn = nb[:, 0] + n[0] * (nb[:, 1] + n[1] * nb[:, 2])
        above |= rk[jn] > rk
    maxima = np.nonzero(~above)[0]
    assert len(maxima) >= 1 and all(idg[m] > 0 for m in maxima)
    ties = sum(int(np.count_nonzero(f == v)) > 1 for v in np.unique(f))
    assert ties > 10
