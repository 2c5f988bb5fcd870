"""Write the golden fixtures this repository's tests read.

Run in a checkout that has the reference mounted at /root/reference (the
machines that run the GPU tests need not have it; the fixtures are
committed).  Everything written is data
from the reference's own test and data files:

  urea_nci.json      urea cell and atoms (dat/lib/crystal.dat 'structure urea')
                     and the expected node values of the reference test
                     tests/015_grdplot/005_nciplot_basic (its ref/*-dens.cube,
                     ref/*-grad.cube and ref/*.dat), plus the test's options.
  urea_frag.json     tests/015_grdplot/006_nciplot_fragment: its CUBE box,
                     options and FRAGMENT coordinates (from the .cri input)
                     and the expected node values (ref/*-dens.cube,
                     ref/*-grad.cube, ref/*-pdens.cube, ref/*.dat) and the
                     cutoffs the run printed (ref/*.cro).
  urea_tables.npz    radial grids r and rr0 = sum(occ*psi^2) of the H, C, N, O
                     atomic wfc files (dat/wfc/*_pbe.wfc), truncated where
                     critic2's read_critic stops reading (rho < 1e-8).
"""
import json
import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from critic2_amd.radial import read_wfc  # noqa: E402


def urea_structure():
    lines = open(os.path.join(REF, "dat/lib/crystal.dat")).read().splitlines()
    i = lines.index("structure urea")
    cell, atoms = None, []
    for ln in lines[i + 1:]:
        t = ln.split()
        if not t:
            continue
        if t[0] == "cell":
            cell = [float(v) for v in t[1:7]]
        elif t[0] == "neq":
            atoms.append(([float(v) for v in t[1:4]], t[4]))
        elif t[0] == "endcrystal":
            break
    return cell, atoms


def cube_values(path):
    toks = open(path).read().splitlines()
    nat = int(toks[2].split()[0])
    vals = " ".join(toks[6 + nat:]).split()
    return [float(v.replace("D", "E")) for v in vals]


def main():
    cell, atoms = urea_structure()
    base = os.path.join(REF, "tests/015_grdplot/ref/005_nciplot_basic-gen")
    dens = cube_values(base + "-dens.cube")
    grad = cube_values(base + "-grad.cube")
    dat = [[float(v) for v in ln.split()] for ln in open(base + ".dat") if ln.strip()]
    out = {
        "source": "tests/015_grdplot/005_nciplot_basic.cri and dat/lib/crystal.dat (urea)",
        "cell_bohr_deg": cell,
        "atoms": [{"x": x, "sym": s} for x, s in atoms],
        "nstep": [2, 2, 2],
        "options": {"rhocut": 0.2, "dimcut": 2.0, "rhoplot": 0.5, "dimplot": 0.05,
                    "srhorange": [-5.0, 5.0]},
        "cube_order": "z fastest: value index = k + 2*(j + 2*i)",
        "dens_cube": dens,
        "grad_cube": grad,
        "dat": dat,
        "cube_atol": 1e-5,
    }
    with open(os.path.join(HERE, "urea_nci.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    fragment_case()
    zmap = {"h": 1, "c": 6, "n": 7, "o": 8}
    arrs = {}
    for sym, z in zmap.items():
        fn = os.path.join(REF, "dat/wfc", f"{sym}__pbe.wfc")
        r, rr0 = read_wfc(fn, z)
        arrs[f"r_{sym}"] = r
        arrs[f"rr0_{sym}"] = rr0
    np.savez_compressed(os.path.join(HERE, "urea_tables.npz"), **arrs)
    tricubic_vectors()


def fragment_case():
    """urea_frag.json from tests/015_grdplot/006_nciplot_fragment."""
    cri = open(os.path.join(REF, "tests/015_grdplot/006_nciplot_fragment.cri")).read().splitlines()
    frags, cur, box, nstep, onlyneg = [], None, None, None, False
    for ln in cri:
        t = ln.split()
        if not t or t[0].startswith("#"):
            continue
        w = t[0].lower()
        if w == "fragment":
            cur = []
        elif w == "endfragment":
            frags.append(cur)
            cur = None
        elif cur is not None:
            cur.append([float(v) for v in t[:3]])
        elif w == "cube":
            box = [float(v) for v in t[1:7]]
        elif w == "nstep":
            nstep = [int(v) for v in t[1:4]]
        elif w == "onlyneg":
            onlyneg = True
    base = os.path.join(REF, "tests/015_grdplot/ref/006_nciplot_fragment")
    cro = open(base + ".cro").read()

    def printed(label):
        i = cro.index(label)
        return float(cro[i + len(label):].split()[0])

    hdr = open(base + "-pdens.cube").read().splitlines()
    nat = int(hdr[2].split()[0])
    out = {
        "source": "tests/015_grdplot/006_nciplot_fragment.cri and its ref/ outputs",
        "cube_origin_bohr": [float(v) for v in hdr[2].split()[1:4]],
        "cube_steps_bohr": [[float(v) for v in hdr[3 + c].split()[1:4]] for c in range(3)],
        "cube_atoms_z": [int(hdr[6 + a].split()[0]) for a in range(nat)],
        "cube_cryst": [box[:3], box[3:]],
        "nstep": nstep,
        "onlyneg": onlyneg,
        "fragments_angstrom": frags,
        "options": {"rhocut": printed("Density cutoff (for dat file):"),
                    "dimcut": printed("RDG cutoff (for dat file):"),
                    "rhoplot": printed("Density plot cutoff:"),
                    "rhoparam": 0.95, "rhoparam2": 0.75},
        "cube_order": "z fastest: value index = k + 2*(j + 2*i)",
        "dens_cube": cube_values(base + "-dens.cube"),
        "grad_cube": cube_values(base + "-grad.cube"),
        "pdens_cube": cube_values(base + "-pdens.cube"),
        "dat": [[float(v) for v in ln.split()] for ln in open(base + ".dat") if ln.strip()],
        "cube_atol": 1e-5,
    }
    with open(os.path.join(HERE, "urea_frag.json"), "w") as fh:
        json.dump(out, fh, indent=1)


def tricubic_vectors():
    """tricubic_ab.json: eight seeded random 64-vectors b (the f, fx, fy, fz,
    fxy, fxz, fyz, fxyz corner values of grinterp_tricubic) and a = c b, with
    c the reference's tricubic matrix (src/grid3mod@proc.f90:76, read as
    data, column-major)."""
    import re

    src = open(os.path.join(REF, "src/grid3mod@proc.f90")).read()
    i = src.index("c(64,64) = reshape((/")
    nums = re.findall(r"-?\d+d0", src[i:src.index("/)", i + 30)])
    assert len(nums) == 4096
    c = np.array([float(v[:-2]) for v in nums]).reshape(64, 64, order="F")
    rng = np.random.default_rng(2649)
    b = rng.standard_normal((8, 64))
    a = b @ c.T
    with open(os.path.join(HERE, "tricubic_ab.json"), "w") as fh:
        json.dump({"source": "src/grid3mod@proc.f90:76 (c) applied as a = matmul(c,b), :2770",
                   "b": b.tolist(), "a": a.tolist()}, fh)


if __name__ == "__main__":
    main()
