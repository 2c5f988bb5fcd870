"""Build libc2gpu.so (HIP, gfx950) in-tree with hipcc.

Ascent/labelling sources are compiled with -ffp-contract=off: the Bader path
must reproduce the FP64 operation order of the reference bit for bit.  All HIP
code disables packed-FP32 instruction selection.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libc2gpu.so")
ARCH = os.environ.get("C2GPU_ARCH", "gfx950")

SOURCES = [
    ("bader.hip", ["-ffp-contract=off"]),
    ("sweep.hip", ["-ffp-contract=off"]),
    ("refine.hip", ["-ffp-contract=off"]),
    ("sums.hip", ["-ffp-contract=off"]),
    ("nci.hip", ["-ffp-contract=fast"]),
    ("ncigrid.hip", ["-ffp-contract=off"]),
    ("cube.hip", ["-ffp-contract=off"]),
    ("ncicut.hip", ["-ffp-contract=off"]),
    ("yt.hip", ["-ffp-contract=off"]),
    ("mpole.hip", ["-ffp-contract=off"]),
    ("wscell.cpp", ["-ffp-contract=off"]),
    ("qcksort.cpp", ["-ffp-contract=off"]),
    ("batch.cpp", ["-ffp-contract=off"]),
    ("capi.cpp", ["-ffp-contract=off"]),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every source to an object and link libc2gpu.so.  Always
    recompiles (this is the 'does it build' check)."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    common = [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument", "-Wno-unused-result",
        "-I", os.path.join(HERE, "..", "include"),
    ]
    procs = []
    objs = []
    for src, extra in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmd = [cc, *common, *extra, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{out.decode(errors='replace')}")
        elif verbose and out:
            print(out.decode(errors="replace"))
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp,
           "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    return LIB


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_HOST = os.path.join(FORTRAN_DIR, "c2gpu_host")
FORTRAN_NCI_HOST = os.path.join(FORTRAN_DIR, "c2gpu_nci_host")


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the iso_c_binding module and the Fortran stand-in hosts (Bader
    and nciplot) with amdflang (skipped, returning None, when amdflang is
    absent)."""
    fc = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        return None
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    steps = [
        [fc, "-O2", "-c", os.path.join(FORTRAN_DIR, "c2gpu_mod.f90"), "-o",
         os.path.join(objdir, "c2gpu_mod.o"), "-module-dir", objdir],
        [fc, "-O2", "-I", objdir, "-module-dir", objdir, os.path.join(FORTRAN_DIR, "c2gpu_host.f90"),
         os.path.join(objdir, "c2gpu_mod.o"), "-o", FORTRAN_HOST, "-L", HERE, "-lc2gpu",
         "-Wl,-rpath,$ORIGIN/.."],
        [fc, "-O2", "-I", objdir, "-module-dir", objdir, os.path.join(FORTRAN_DIR, "c2gpu_nci_host.f90"),
         os.path.join(objdir, "c2gpu_mod.o"), "-o", FORTRAN_NCI_HOST, "-L", HERE, "-lc2gpu",
         "-Wl,-rpath,$ORIGIN/.."],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            raise RuntimeError("amdflang failed:\n" + r.stdout.decode(errors="replace"))
    return FORTRAN_HOST


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
