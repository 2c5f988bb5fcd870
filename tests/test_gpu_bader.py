"""GPU parity of the Bader path: labels bit-exact against the oracle (zero
disagreeing points), per-basin sums within 1e-10 relative."""
import numpy as np
import pytest

from critic2_amd import _lib
from critic2_amd import bader as gb
from critic2_amd.synth import config, random_cell
from oracle import oracle

pytestmark = pytest.mark.gpu

RIMG = 8.0 / 0.52917720859


def _grid(ctx, cell):
    """The same FP64 grid for both sides: generated on the GPU, copied back."""
    return gb.synth_slater(ctx, cell)


CASES = {
    "cubic-a10-64^3-8atoms": lambda: config(1),
    "triclinic-40x44x48-6atoms": lambda: random_cell(11, (9.0, 10.0, 11.0), (78, 97, 105), 6, (1, 6, 7, 8),
                                                    (40, 44, 48)),
    "ortho-36x52x60-12atoms": lambda: random_cell(12, (8.0, 11.0, 12.5), (90, 90, 90), 12, (1, 6, 7, 8),
                                                 (36, 52, 60)),
    "cubic-a10-96^3-8atoms": lambda: config(1, n_override=(96, 96, 96)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_bader_labels_bit_exact(ctx, name):
    cell = CASES[name]()
    print(f"[case] {name} n={cell.n} nat={len(cell.zat)}", flush=True)
    f = _grid(ctx, cell)
    ref, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat)
    bas = gb.bader_integrate(ctx, f, cell.n, cell.x2c, cell.xat)
    mism = int(np.count_nonzero(bas.idg != ref))
    print(f"  oracle refine iters {info['refine_iters']}; gpu stats {bas.stats}", flush=True)
    assert mism == 0, f"{mism} disagreeing labels"
    assert bas.nattr == info["nattr"]
    assert bas.stats["refine_iters"] == info["refine_iters"]
    np.testing.assert_array_equal(bas.xattr, xattr)


@pytest.mark.parametrize("rule,ratom", [("nnm", 0.1), ("nnm", 1.0), ("noatoms", 1.0), ("noatoms", 0.05),
                                        ("noatoms", 100.0)])
def test_bader_attractor_rules_follow_discovery_order(ctx, rule, ratom):
    """NNM / RATOM r / NOATOMS (bader@proc.f90:159-199): a new maximum is an
    atom within ratom, else a known attractor within ratom, else a new
    attractor numbered in the order the serial sweep discovers it.  Labels,
    the attractor count and their positions must equal the oracle's."""
    cell = random_cell(13, (8.0, 9.0, 10.0), (85, 95, 100), 5, (1, 6, 7, 8), (32, 36, 40))
    print(f"[case] attractor rule {rule} ratom={ratom} n={cell.n}", flush=True)
    f = _grid(ctx, cell)
    omode = _lib.ATTR_RATOM if rule == "nnm" else _lib.ATTR_NOATOMS
    ref, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat, mode=omode, ratom=ratom)
    kw = dict(nnm=True, ratom=ratom) if rule == "nnm" else dict(noatoms=True, ratom=ratom)
    bas = gb.bader_integrate(ctx, f, cell.n, cell.x2c, cell.xat, **kw)
    assert int(np.count_nonzero(bas.idg != ref)) == 0
    assert bas.nattr == info["nattr"]
    np.testing.assert_array_equal(bas.xattr, xattr)
    if rule == "noatoms" and ratom > 50:
        assert bas.nattr == 1
    if rule == "nnm" and ratom < 0.2:
        assert bas.nattr > len(cell.zat)  # maxima off the nuclei became non-nuclear attractors


@pytest.mark.parametrize("rule,ratom,side", [("noatoms", 0.05, 0), ("noatoms", 0.05, 1), ("nnm", 0.1, 2),
                                             ("nnm", 0.1, 0)])
def test_bader_discard_matches_oracle(ctx, rule, ratom, side):
    """DISCARD (bader@proc.f90:185-191): maxima that would become new
    attractors are offered to the expression (here: a half-space of the
    cell, by Cartesian coordinate `side`); discarded ones get no attractor
    and their basins keep label 0.  Labels, attractors and positions equal
    the oracle's, which evaluates the expression at every visit."""
    cell = random_cell(13, (8.0, 9.0, 10.0), (85, 95, 100), 5, (1, 6, 7, 8), (32, 36, 40))
    print(f"[case] discard {rule} ratom={ratom} side={side} n={cell.n}", flush=True)
    f = _grid(ctx, cell)
    mid = 0.5 * np.asarray(cell.x2c).sum(1)[side]

    def discard(x):
        return x[side] < mid

    omode = _lib.ATTR_RATOM if rule == "nnm" else _lib.ATTR_NOATOMS
    ref, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat, mode=omode, ratom=ratom, discard=discard)
    kw = dict(nnm=True, ratom=ratom) if rule == "nnm" else dict(noatoms=True, ratom=ratom)
    bas = gb.bader_integrate(ctx, f, cell.n, cell.x2c, cell.xat, discard=discard, **kw)
    assert int(np.count_nonzero(bas.idg != ref)) == 0
    assert bas.nattr == info["nattr"]
    np.testing.assert_array_equal(bas.xattr, xattr)
    if rule == "noatoms":
        assert np.count_nonzero(ref == 0) > 0  # some basins were discarded


@pytest.mark.parametrize("mode", ["C2GPU_SWEEP_VERIFY", "C2GPU_SWEEP_FULL"])
def test_sweep_incremental_equals_full_rebuild(ctx, mode, monkeypatch):
    """The incremental fixpoint steps against full rebuilds: VERIFY rebuilds
    the converged story and checks it is a fixpoint with identical known==2
    times; FULL rebuilds every step.  Both must give the oracle's labels."""
    cell = CASES["triclinic-40x44x48-6atoms"]()
    print(f"[case] {mode} triclinic 40x44x48", flush=True)
    f = _grid(ctx, cell)
    ref, info, _, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat)
    monkeypatch.setenv(mode, "1")
    bas = gb.bader_integrate(ctx, f, cell.n, cell.x2c, cell.xat)
    assert int(np.count_nonzero(bas.idg != ref)) == 0


def test_bader_device_pointers_and_sums(ctx):
    cell = config(1)
    f = _grid(ctx, cell)
    ref, info, _, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat)
    fd = ctx.to_device(f)
    bas = gb.bader_integrate(ctx, fd, cell.n, cell.x2c, cell.xat)
    lab = bas.idg.numpy()
    assert np.count_nonzero(lab != ref) == 0
    # sums: GPU segmented reduction vs the oracle's sequential masked sums
    lapg = oracle.fft_deriv(f, cell.n, cell.x2c, "lap")
    c_ref, s_ref = oracle.basin_sums(ref, info["nattr"], [f, lapg])
    c_gpu, s_gpu = gb.basin_sums(ctx, bas.idg, bas.nattr, [fd, ctx.to_device(lapg)])
    np.testing.assert_array_equal(c_gpu, c_ref)
    np.testing.assert_allclose(s_gpu, s_ref, rtol=1e-10, atol=1e-10 * np.abs(s_ref).max())
    # identities: sum of volumes is Omega, sum of populations is the grid integral
    props = gb.intgrid_fields(ctx, bas, f, cell.x2c, lap=True)
    assert abs(props["volume"].sum() - cell.omega) <= 1e-12 * cell.omega
    assert abs(props["pop"].sum() - f.sum() * cell.omega / f.size) <= 1e-10 * f.sum() * cell.omega / f.size
    assert abs(props["pop"].sum() - cell.zat.sum()) < 1e-2 * cell.zat.sum()


def test_basin_sums_deterministic_and_ragged(ctx):
    """Reproducible sums, sizes not a multiple of the tile, labels 0 ignored."""
    rng = np.random.default_rng(7)
    for npts in (1, 63, 257, 100003):
        lab = rng.integers(0, 6, size=npts).astype(np.int32)
        lab = np.sort(lab)  # runs, like basin labels
        fld = rng.standard_normal(npts)
        c1, s1 = gb.basin_sums(ctx, lab, 5, [fld])
        c2, s2 = gb.basin_sums(ctx, lab, 5, [fld])
        assert np.array_equal(s1, s2)
        c_ref, s_ref = oracle.basin_sums(lab, 5, [fld])
        np.testing.assert_array_equal(c1, c_ref)
        np.testing.assert_allclose(s1, s_ref, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(fld).sum()))


@pytest.mark.parametrize("nattr,nf", [(1500, 1), (6000, 2), (40000, 0)])
def test_basin_sums_many_attractors(ctx, nattr, nf):
    """More bins than the LDS holds (attractors x (fields + 1) x 4 waves x 8
    bytes > 64 KiB): per-wavefront rows in global memory, same results;
    labels in runs (basins) and scattered (a ragged tail)."""
    rng = np.random.default_rng(nattr)
    npts = 300007
    lab = np.sort(rng.integers(0, nattr + 1, size=npts)).astype(np.int32)
    lab[-5000:] = rng.integers(0, nattr + 1, size=5000)
    flds = [rng.standard_normal(npts) for _ in range(nf)]
    c1, s1 = gb.basin_sums(ctx, lab, nattr, flds)
    c2, s2 = gb.basin_sums(ctx, lab, nattr, flds)
    assert np.array_equal(s1, s2) and np.array_equal(c1, c2)
    c_ref, s_ref = oracle.basin_sums(lab, nattr, flds)
    np.testing.assert_array_equal(c1, c_ref)
    if nf:
        np.testing.assert_allclose(s1, s_ref, rtol=1e-12, atol=1e-12 * max(1.0, float(np.abs(flds[0]).sum())))


def test_basin_sums_single_label_tiles(ctx):
    """Long runs of one label (whole 64-point tiles) take the plain
    reduction; the result still matches the oracle and repeats bit for bit."""
    rng = np.random.default_rng(11)
    npts = 64 * 5000 + 17
    lab = np.repeat(np.arange(1, 11, dtype=np.int32), npts // 10 + 1)[:npts]
    fld = rng.random(npts)
    c1, s1 = gb.basin_sums(ctx, lab, 10, [fld])
    _, s2 = gb.basin_sums(ctx, lab, 10, [fld])
    assert np.array_equal(s1, s2)
    c_ref, s_ref = oracle.basin_sums(lab, 10, [fld])
    np.testing.assert_array_equal(c1, c_ref)
    np.testing.assert_allclose(s1, s_ref, rtol=1e-12)


def test_step_pointers_match_oracle(ctx):
    cell = CASES["triclinic-40x44x48-6atoms"]()
    f = _grid(ctx, cell)
    on_ref, ne_ref, dr_ref = oracle.step_pointers(f, cell.n, cell.x2c)
    N = f.size
    on, ne, dr = np.zeros(N, np.int64), np.zeros(N, np.int64), np.zeros(3 * N)
    _, c2l, lid = _lib.bader_metrics(cell.x2c, cell.n)
    nn = np.asarray(cell.n, dtype=np.int32)
    rc = ctx.lib.c2gpu_bader_step_pointers(ctx.h, f.ctypes.data, nn.ctypes.data, c2l.ctypes.data,
                                           lid.ctypes.data, 0, on.ctypes.data, ne.ctypes.data, dr.ctypes.data)
    ctx.check(rc, "step_pointers")
    assert np.array_equal(on, on_ref)
    assert np.array_equal(ne, ne_ref)
    assert np.array_equal(dr, dr_ref)


def test_relabel(ctx):
    lab = np.array([1, 2, 3, 2, 1, 0, 4], dtype=np.int32)
    mp = np.array([3, 1, 2], dtype=np.int32)
    rc = ctx.lib.c2gpu_relabel(ctx.h, lab.size, mp.ctypes.data, 3, 0, lab.ctypes.data)
    ctx.check(rc, "relabel")
    assert lab.tolist() == [3, 1, 2, 1, 3, 0, 4]


def test_bad_arguments_return_status(ctx):
    cell = config(1)
    f = np.zeros(8)
    with pytest.raises(_lib.C2GpuError) as e:
        gb.bader_integrate(ctx, f, (2, 2, 2), cell.x2c, cell.xat)
    assert e.value.code == 1


def test_device_flag_rejects_host_pointers(ctx):
    """C2GPU_PTR_DEVICE with host memory is an argument error raised before
    any kernel is launched (regression: a host field with device labels)."""
    lab = np.ones(64, dtype=np.int32)
    fld = np.ones(64)
    fp = (_lib.C.c_void_p * 1)(fld.ctypes.data)
    sums = np.zeros(1)
    counts = np.zeros(1, dtype=np.int64)
    rc = ctx.lib.c2gpu_basin_sums(ctx.h, lab.ctypes.data, 64, 1, 1, fp, _lib.PTR_DEVICE,
                                  sums.ctypes.data, counts.ctypes.data)
    assert rc == 1
    dlab = ctx.to_device(lab)
    rc = ctx.lib.c2gpu_basin_sums(ctx.h, dlab.data_ptr(), 64, 1, 1, fp, _lib.PTR_DEVICE,
                                  sums.ctypes.data, counts.ctypes.data)
    assert rc == 1
    # the Python mirror moves host fields next to device labels
    c, s = gb.basin_sums(ctx, dlab, 1, [fld])
    assert c[0] == 64 and s[0, 0] == 64.0


def test_synthetic_grid_matches_oracle_generator(ctx):
    cell = random_cell(21, (7.0, 8.0, 9.0), (85, 95, 100), 5, (1, 6, 7, 8), (20, 22, 24))
    g = gb.synth_slater(ctx, cell)
    o = oracle.slater_grid(cell.x2c, cell.n, cell.xat, cell.zat, RIMG)
    np.testing.assert_allclose(g, o, rtol=1e-12, atol=1e-14 * o.max())


@pytest.mark.slow
def test_bader_256_orthorhombic_64_atoms(ctx):
    """The plain-FP64 GPU workload: 20x20x24 A, 256^3, 64 atoms."""
    cell = config(2)
    print(f"[case] config2 n={cell.n}", flush=True)
    f = _grid(ctx, cell)
    bas = gb.bader_integrate(ctx, f, cell.n, cell.x2c, cell.xat)
    print(f"  gpu stats {bas.stats}", flush=True)
    ref, info, _, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat)
    assert int(np.count_nonzero(bas.idg != ref)) == 0
    c_ref, s_ref = oracle.basin_sums(ref, info["nattr"], [f])
    c_gpu, s_gpu = gb.basin_sums(ctx, bas.idg, bas.nattr, [f])
    np.testing.assert_array_equal(c_gpu, c_ref)
    np.testing.assert_allclose(s_gpu, s_ref, rtol=1e-10)


EDGE = [  # name, seed, cell edges (A), angles, atoms, grid, quantisation digits (plateaus)
    ("tiny-3x5x7", 31, (4.0, 5.0, 6.0), (90, 90, 90), 2, (3, 5, 7), None),
    ("odd-17x3x11", 32, (7.0, 4.0, 6.0), (80, 100, 95), 3, (17, 3, 11), None),
    ("cube-5", 33, (5.0, 5.0, 5.0), (90, 90, 90), 1, (5, 5, 5), None),
    ("prime-13x19x23", 34, (7.0, 8.0, 9.0), (85, 92, 101), 4, (13, 19, 23), None),
    ("plateaus-3digits-24x26x28", 35, (8.0, 8.5, 9.0), (90, 90, 90), 5, (24, 26, 28), 3),
    ("plateaus-2digits-20x22x24", 36, (8.0, 8.5, 9.0), (90, 90, 90), 4, (20, 22, 24), 2),
    ("plateaus-4digits-30^3", 37, (9.0, 9.0, 9.0), (70, 80, 90), 6, (30, 30, 30), 4),
]


@pytest.mark.parametrize("name,seed,edge,ang,nat,n,q", EDGE, ids=[e[0] for e in EDGE])
def test_bader_edge_grids_bit_exact(ctx, name, seed, edge, ang, nat, n, q):
    """The smallest grids (3 points on an axis), odd and prime sizes, and
    grids rounded to q significant digits of their maximum, whose plateaus
    exercise the strict-> tie rules (first maximum of step_ongrid, plateau
    maxima of is_max, zeroed gradient components)."""
    cell = random_cell(seed, edge, ang, nat, (1, 6, 7, 8), n)
    print(f"[case] {name} n={cell.n}", flush=True)
    f = _grid(ctx, cell)
    if q:
        mx = f.max()
        f = np.round(f / mx * 10 ** q) / 10 ** q * mx
    ref, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat)
    bas = gb.bader_integrate(ctx, f, cell.n, cell.x2c, cell.xat)
    assert int(np.count_nonzero(bas.idg != ref)) == 0
    assert bas.nattr == info["nattr"]
    assert bas.stats["refine_iters"] == info["refine_iters"]
    np.testing.assert_array_equal(bas.xattr, xattr)
