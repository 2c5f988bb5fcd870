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
