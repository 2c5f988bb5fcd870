"""The drop-in boundary driven from Fortran: the amdflang-built stand-in host
(critic2_amd/fortran/c2gpu_host.f90) calls libc2gpu through the c2gpu
iso_c_binding module, as a bader@c2gpu submodule of critic2 would, and its
labels and integrated Volume/Pop must equal the oracle's."""
import os
import subprocess

import numpy as np
import pytest

from critic2_amd import bader as gb
from critic2_amd import build as cb
from critic2_amd.synth import random_cell
from oracle import oracle

pytestmark = pytest.mark.gpu


def test_fortran_host_bader(ctx, tmp_path):
    if not os.path.exists(cb.FORTRAN_HOST):
        pytest.skip("Fortran stand-in host not built (amdflang absent at build time)")
    cell = random_cell(51, (9.0, 9.5, 10.0), (85, 95, 100), 5, (1, 6, 7, 8), (36, 40, 44))
    print(f"[case] fortran host n={cell.n}", flush=True)
    f = gb.synth_slater(ctx, cell)
    case = tmp_path / "case.txt"
    with open(case, "w") as fh:
        fh.write(" ".join(str(v) for v in cell.n) + "\n")
        fh.write(" ".join(repr(float(v)) for v in np.asarray(cell.x2c).reshape(-1, order="F")) + "\n")
        fh.write(f"{len(cell.zat)}\n")
        for x in cell.xat:
            fh.write(" ".join(repr(float(v)) for v in x) + "\n")
    f.astype(np.float64).tofile(tmp_path / "grid.bin")
    r = subprocess.run([cb.FORTRAN_HOST, str(case), str(tmp_path / "grid.bin"), str(tmp_path / "lab.bin"),
                        str(tmp_path / "props.txt")], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       timeout=120)
    out = r.stdout.decode(errors="replace")
    assert r.returncode == 0, out
    lab = np.fromfile(tmp_path / "lab.bin", dtype=np.int32)
    ref, info, _, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat)
    assert int(np.count_nonzero(lab != ref)) == 0
    lines = open(tmp_path / "props.txt").read().splitlines()
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[1:]])
    counts, sums = oracle.basin_sums(ref, info["nattr"], [f])
    w = cell.omega / f.size
    np.testing.assert_allclose(rows[:, 1], counts * w, rtol=1e-14)
    np.testing.assert_allclose(rows[:, 2], sums[0] * w, rtol=1e-10)


def _write_case(path, cell):
    with open(path, "w") as fh:
        fh.write(" ".join(str(v) for v in cell.n) + "\n")
        fh.write(" ".join(repr(float(v)) for v in np.asarray(cell.x2c).reshape(-1, order="F")) + "\n")
        fh.write(f"{len(cell.zat)}\n")
        for x in cell.xat:
            fh.write(" ".join(repr(float(v)) for v in x) + "\n")


@pytest.mark.parametrize("mode,ratom,discard", [(1, 0.1, None), (2, 0.05, 0.5), (1, 0.1, 0.4)])
def test_fortran_host_attractor_rules_and_discard(ctx, tmp_path, mode, ratom, discard):
    """NNM with more attractors than atoms (the host's attractor array starts
    at nat entries and grows from the count the library reports, as the
    bader@c2gpu submodule does), and DISCARD through a bind(c) callback
    (new maxima with Cartesian x below a fraction of the cell rejected):
    labels, attractor count and positions equal the oracle's."""
    if not os.path.exists(cb.FORTRAN_HOST):
        pytest.skip("Fortran stand-in host not built (amdflang absent at build time)")
    cell = random_cell(53, (8.0, 9.0, 10.0), (85, 95, 100), 5, (1, 6, 7, 8), (32, 36, 40))
    print(f"[case] fortran host mode {mode} ratom {ratom} discard {discard}", flush=True)
    f = gb.synth_slater(ctx, cell)
    _write_case(tmp_path / "case.txt", cell)
    f.astype(np.float64).tofile(tmp_path / "grid.bin")
    args = [cb.FORTRAN_HOST, str(tmp_path / "case.txt"), str(tmp_path / "grid.bin"), str(tmp_path / "lab.bin"),
            str(tmp_path / "props.txt"), str(mode), repr(ratom)]
    xlim = None
    if discard is not None:
        xlim = discard * float(np.asarray(cell.x2c)[0].sum())
        args.append(repr(xlim))
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)
    out = r.stdout.decode(errors="replace")
    assert r.returncode == 0, out
    lab = np.fromfile(tmp_path / "lab.bin", dtype=np.int32)
    ref, info, xattr, _ = oracle.bader(f, cell.n, cell.x2c, cell.xat, mode=mode, ratom=ratom,
                                       discard=(lambda x: x[0] < xlim) if xlim is not None else None)
    assert int(np.count_nonzero(lab != ref)) == 0
    lines = open(tmp_path / "props.txt").read().splitlines()
    head = lines[0].split()
    assert int(head[1]) == info["nattr"]
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[1:]]).reshape(-1, 7)
    np.testing.assert_array_equal(rows[:, 4:7], xattr)
    if mode == 1 and discard is None:
        assert info["nattr"] > len(cell.zat) and int(head[-1]) == info["nattr"]  # the array was grown
    if discard is not None and mode == 2:
        assert np.count_nonzero(ref == 0) > 0


def test_fortran_host_reads_cube(ctx, tmp_path):
    """The Fortran host loading its grid through c2gpu_cube_read (the
    grid3%read_cube drop-in): the grid it labels is the cube's values, so
    its labels equal the oracle's on the oracle's own read of the file."""
    if not os.path.exists(cb.FORTRAN_HOST):
        pytest.skip("Fortran stand-in host not built (amdflang absent at build time)")
    from critic2_amd import cube
    cell = random_cell(52, (8.5, 9.0, 9.5), (90, 90, 90), 4, (1, 6, 8), (32, 36, 30))
    print(f"[case] fortran host cube n={cell.n}", flush=True)
    f = gb.synth_slater(ctx, cell)
    step = np.asarray(cell.x2c) / np.array(cell.n, dtype=np.float64)[None, :]
    path = tmp_path / "grid.cube"
    zat = np.asarray(cell.zat, dtype=np.int32)
    xc = (np.asarray(cell.x2c) @ np.asarray(cell.xat).T).T
    cube.write_cube(ctx, str(path), cell.n, (0.0, 0.0, 0.0), step, zat, xc, f=f, style=0)
    case = tmp_path / "case.txt"
    with open(case, "w") as fh:
        fh.write(" ".join(str(v) for v in cell.n) + "\n")
        fh.write(" ".join(repr(float(v)) for v in np.asarray(cell.x2c).reshape(-1, order="F")) + "\n")
        fh.write(f"{len(cell.zat)}\n")
        for x in cell.xat:
            fh.write(" ".join(repr(float(v)) for v in x) + "\n")
    r = subprocess.run([cb.FORTRAN_HOST, str(case), str(path), str(tmp_path / "lab.bin"),
                        str(tmp_path / "props.txt")], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       timeout=120)
    out = r.stdout.decode(errors="replace")
    assert r.returncode == 0, out
    lab = np.fromfile(tmp_path / "lab.bin", dtype=np.int32)
    _, g = oracle.read_cube(str(path))
    ref, info, _, _ = oracle.bader(g, cell.n, cell.x2c, cell.xat)
    assert int(np.count_nonzero(lab != ref)) == 0


def _nci_host_input(path, tables, x2c, xat, species, nodes, frags, opts):
    """The stream file c2gpu_nci_host.f90 reads (its header lists the layout)."""
    i4 = lambda *v: np.asarray(v, dtype=np.int32).tobytes()  # noqa: E731
    f8 = lambda *v: np.asarray(v, dtype=np.float64).reshape(-1).tobytes()  # noqa: E731
    b = [i4(len(tables))]
    for g in tables:
        b += [i4(g.ngrid), f8(g.a, g.b, g.rmax, g.cutrad), f8(g.r), f8(g.f), f8(g.fp), f8(g.fpp)]
    b += [f8(np.asarray(x2c).reshape(-1, order="F")), i4(len(species)), f8(np.asarray(xat).reshape(-1)),
          i4(*species)]
    b += [i4(*nodes.nstep), f8(*nodes.x0), f8(*nodes.xmat)]
    b += [i4(len(frags)), i4(*[len(s) for _, s in frags])]
    b += [f8(np.concatenate([x.reshape(-1) for x, _ in frags])), i4(*np.concatenate([s for _, s in frags]))]
    b += [f8(opts["rhocut"], opts["dimcut"], opts["rhoplot"], -1e30, 1e30), i4(int(opts["onlyneg"])),
          f8(opts.get("rho_void", -1.0), opts["rhoparam"], opts["rhoparam2"])]
    path.write_bytes(b"".join(b))


def test_fortran_nci_host_fragments(ctx, tmp_path):
    """The nciplot stand-in host (c2gpu_nci_host.f90) on the reference's
    006_nciplot_fragment case through the c2gpu module: the same cubes and
    masks as the reference, and bit for bit what the library gives the
    Python host."""
    if not os.path.exists(cb.FORTRAN_NCI_HOST):
        pytest.skip("Fortran stand-in host not built (amdflang absent at build time)")
    import json

    from critic2_amd import nci as gn
    from tests.helpers import GOLDEN, fmt6, urea
    d, x2c, tables, xat, species = urea()
    f = json.load(open(os.path.join(GOLDEN, "urea_frag.json")))
    nb = gn.box_nodes(x2c, *f["cube_cryst"], nstep=f["nstep"])
    frags = [gn.fragment_atoms(x2c, xat, species, np.array(fr)) for fr in f["fragments_angstrom"]]
    opts = dict(f["options"], onlyneg=f["onlyneg"])
    print("[case] fortran nci host 006_nciplot_fragment", flush=True)
    _nci_host_input(tmp_path / "in.bin", tables, x2c, xat, species, nb, frags, opts)
    r = subprocess.run([cb.FORTRAN_NCI_HOST, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)
    assert r.returncode == 0, r.stdout.decode(errors="replace")
    raw = (tmp_path / "out.bin").read_bytes()
    dlen = int(np.frombuffer(raw[:8], dtype=np.int64)[0])
    text = raw[8:8 + dlen]
    vals = np.frombuffer(raw[8 + dlen:], dtype=np.float64)
    npts = int(np.prod(f["nstep"]))
    crho, cg, rhoat = vals[:npts], vals[npts:2 * npts], vals[2 * npts:3 * npts]
    rf = vals[3 * npts:3 * npts + len(frags) * npts].reshape(len(frags), npts)
    assert [fmt6(v) for v in crho] == f["dens_cube"]
    assert [fmt6(v) for v in rhoat] == f["pdens_cube"]
    assert cg.tolist() == f["grad_cube"] and text == b""
    c1, g1, a1 = gn.promol_nodes(ctx, nb, x2c, tables, xat, species, precision="fp64")
    r1 = gn.fragments(ctx, nb, tables, frags, precision="fp64")
    assert np.array_equal(crho, c1) and np.array_equal(rhoat, a1) and np.array_equal(rf, r1)
