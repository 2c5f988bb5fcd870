"""Throughput of the on-grid Bader + NCI path on one MI355X.

Workload (the 512^3 precision form): cubic cell a = 30 A, 512^3 FP64 grid of
the Slater-sum density of 216 seeded atoms (Z in {1,6,7,8}), resident in HBM
before the timed region.  One step is the whole path over that grid:
  1. near-grid Bader labels, exact (FP64 ascent),
  2. per-basin Volume and Pop (FP64 sums),
  3. NCI reduced density gradient and signed density from the promolecular
     radial tables in FP32 (the mixed-precision path).
The plain FP64 GPU path (same step with the FP64 promolecular evaluator) is
timed too and reported beside it.  value = grid points / step time.

--gpus above 1 exits: the path runs on one GPU (multi-GPU sharding is a
design, not part of this build).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from critic2_amd import _lib  # noqa: E402
from critic2_amd import bader as gb  # noqa: E402
from critic2_amd import nci as gn  # noqa: E402
from critic2_amd.radial import slater_grid1  # noqa: E402
from critic2_amd.synth import config  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


# Compulsory bytes per grid point of each kernel family (the roofline
# table of the path): the Bader kernels read the FP64 grid once and write an
# int32 label (12 B); the per-basin sums read a label and each field (4 + 8
# per field); NCI from the FP64 grid reads it and writes two FP32 values
# (16 B); the FP32 promolecular evaluator writes two FP32 values (8 B, but it
# is VALU-bound); the whole Bader + NCI step is 20 B a point.
STEP_BYTES_PER_POINT = 20
PROMOL_KERNELS = ("k_promol_f", "k_promol_d", "k_promol_f_queue")
SUMS_KERNELS = ("k_basin_partials", "k_basin_reduce")
NCI_GRID_KERNELS = ("k_nci_fft_point", "k_spectral", "k_real_to_complex", "k_accum_sq", "k_count_pos",
                    "k_real_part", "k_sqrt", "k_nci_grid")


def compulsory_bytes_per_point(name, nfields=1):
    if name in PROMOL_KERNELS:
        return 8
    if name in SUMS_KERNELS:
        return 4 + 8 * nfields
    if name in NCI_GRID_KERNELS:
        return 16
    return 12  # a Bader kernel


def kernel_bytes(name, N, ev_per_launch, traj_points, nfields=1):
    """The kernel's own algorithmic bytes per launch (each array element the
    kernel must read or write in its design, counted once) for the full-grid
    kernels whose launches the library does not report itself."""
    table = {
        # step record of each point (32 B), path length and pool position
        # written (4 + 8 B), every path point written (4 B)
        "k_traj": 44 * N + 4 * traj_points,
        "k_traj_reorder": 16 * N + 8 * traj_points,  # offsets; path points read and written
        "k_promol_f": 8 * N,                         # FP32 crho + cgrad
        "k_promol_d": 16 * N,
        "k_basin_partials": (4 + 8 * nfields) * N,
        "k_reassign": 8 * N,
    }
    return table.get(name)


def measured_flops(kernel):
    """FLOPs per launch of `kernel` from a committed rocprofv3 PMC summary
    (the `flops_per_launch` lines of profiles/*_valu.md), or None."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_valu.md"))):
        for line in open(path):
            t = line.split()
            if len(t) == 3 and t[0] == "flops_per_launch" and t[1] == kernel:
                return int(t[2])
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summaries (FETCH_SIZE and WRITE_SIZE runs of this workload, scaled by a
    streaming kernel of known traffic): the `traffic_bytes_per_launch` lines
    of profiles/*_traffic.md.  None when the kernel was not measured."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.md"))):
        for line in open(path):
            t = line.split()
            if len(t) == 3 and t[0] == "traffic_bytes_per_launch" and t[1] == kernel:
                return int(t[2])
    return None


def species_of(cell):
    zs = sorted(set(int(z) for z in cell.zat))
    return [slater_grid1(z) for z in zs], np.array([zs.index(int(z)) for z in cell.zat], dtype=np.int32)


def run_step(ctx, cell, rho, lab, outs, tables, species, precision):
    bas = gb.bader_integrate(ctx, rho, cell.n, cell.x2c, cell.xat, labels_out=lab)
    counts, sums = gb.basin_sums(ctx, lab, bas.nattr, [rho])
    gn.nciplot_promol(ctx, cell.n, cell.x2c, tables, cell.xat, species, precision=precision,
                      outs=outs, out_f32=(precision == "fp32"))
    ctx.sync()
    return bas


def run_host_buffers(ctx, args):
    """The same Bader + Volume/Pop + FP32 NCI step with every array in host
    memory, as a Fortran caller hands them over: the grid goes to the device
    and labels, sums and the NCI outputs come back inside each call (the
    PCIe-inclusive rate; the headline keeps the grid resident)."""
    cell = config(args.config, n_override=(args.n,) * 3 if args.config == 3 else None)
    N = cell.npts
    tables, species = species_of(cell)
    rho_d = ctx.empty((N,), np.float64)
    gb.synth_slater(ctx, cell, device_out=rho_d)
    rho = rho_d.numpy()
    rho_d.free()
    print(f"[bench] grid {cell.n} in host memory, {N} points", file=sys.stderr, flush=True)

    def step_():
        bas = gb.bader_integrate(ctx, rho, cell.n, cell.x2c, cell.xat)
        gb.basin_sums(ctx, bas.idg, bas.nattr, [rho])
        gn.nciplot_promol(ctx, cell.n, cell.x2c, tables, cell.xat, species, precision="fp32", out_f32=True)
        ctx.sync()
        return bas

    for w in range(args.warmup):
        step_()
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        step_()
        print(f"[bench] step {s_ + 1}/{args.steps} done", file=sys.stderr, flush=True)
    dt = (time.perf_counter() - t0) / args.steps
    out = {"metric": f"Gpoints/s Bader+NCI @{cell.n[0]}^3, host buffers (PCIe-inclusive)",
           "value": round(N / dt / 1e9, 6), "unit": "Gpoints/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
           "config": {"workload": "the Bader + Volume/Pop + FP32 promolecular NCI step with the grid, labels and "
                                  "NCI outputs in host memory", "grid": list(cell.n),
                      "parallelism": "single GPU"}}
    print(json.dumps(out), flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def host_threads():
    """Host threads the CPU leg may use: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def nci_cpu_time(cell, tables, species, c0, shape, threads):
    """The oracle's promolecular NCI point loop (OpenMP, schedule(dynamic)
    as nci@proc.f90:543) over a block of `shape` grid nodes from node c0 of
    `cell`, with every atom image that reaches the block: (nodes, seconds)."""
    from oracle import oracle
    n = cell.n
    I, J, K = np.meshgrid(*[np.arange(c0[d], c0[d] + shape[d]) for d in range(3)], indexing="ij")
    frac = np.stack([I / n[0], J / n[1], K / n[2]], -1).reshape(-1, 3)
    pts = frac @ np.asarray(cell.x2c).T
    otabs = [oracle.Table(g.r, g.f, g.fp, g.fpp, g.cutrad) for g in tables]
    rc = max(min(g.cutrad, g.rmax) for g in tables)
    centre = pts.mean(0)
    rad = np.sqrt(((pts - centre) ** 2).sum(1)).max() + rc
    pos, sp = [], []
    for a, x in enumerate(cell.xat):
        for t in np.ndindex(3, 3, 3):
            p = np.asarray(cell.x2c) @ (x + np.array(t) - 1)
            if np.linalg.norm(p - centre) <= rad:
                pos.append(p)
                sp.append(species[a])
    t0 = time.perf_counter()
    oracle.promol_nci_batch(pts, np.array(pos), np.array(sp, dtype=np.int32), otabs, threads=threads)
    return len(pts), time.perf_counter() - t0


def cpu_baseline_batch(ctx, cells, tables, species, bader_points=3_000_000, nci_cells=8):
    """The oracle on a bounded sample of the batch: Bader serially (as
    bader@proc.f90:151-224) on the first cells up to bader_points points,
    and the promolecular NCI point loop (OpenMP dynamic, every host thread)
    on a 32x32x16 node block at the centre of each of the first nci_cells
    cells; the per-point rates are applied to every point of the batch."""
    from oracle import oracle
    t_b, n_b, used = 0.0, 0, 0
    for c in cells:
        if n_b >= bader_points:
            break
        f = gb.synth_slater(ctx, c)
        t0 = time.perf_counter()
        oracle.bader(f, c.n, c.x2c, c.xat)
        t_b += time.perf_counter() - t0
        n_b += f.size
        used += 1
    threads = host_threads()
    t_n, n_n = 0.0, 0
    for q, c in enumerate(cells[:nci_cells]):
        shape = tuple(min(b, v) for b, v in zip((32, 32, 16), c.n))
        npts, dt_ = nci_cpu_time(c, tables, species[q], [v // 2 - b // 2 for v, b in zip(c.n, shape)], shape,
                                 threads)
        t_n += dt_
        n_n += npts
    tb, tn = t_b / n_b, t_n / n_n
    return {"value": 1e-9 / (tb + tn), "unit": "Gpoints/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": f"oracle Bader (serial sweep + refinement, 1 thread, as the reference) on the first {used} "
                      f"cells ({n_b} points): {tb * 1e6:.3f} us/pt; promolecular NCI (OpenMP dynamic over nodes, "
                      f"{threads} threads) on {n_n} nodes from the first {min(nci_cells, len(cells))} cells: "
                      f"{tn * 1e6:.3f} us/pt; both rates applied to every point of the batch",
            "bader_us_per_point": round(tb * 1e6, 4), "nci_us_per_point": round(tn * 1e6, 4)}


def cpu_baseline(ctx, cell, tables, species, n_bader=256):
    """The oracle (C restatement of critic2's algorithm) on a bounded sample
    of the same workload, with the reference's parallel structure: the
    Bader sweep and refinement serially (bader@proc.f90:151-224 has no
    OpenMP) on the same cell at n_bader^3, and the promolecular NCI point
    loop OpenMP-parallel with schedule(dynamic) (nci@proc.f90:543) on every
    host thread, over a 128x128x32 block of the 512^3 nodes.  The per-point
    rates are applied to the whole grid (Gpoints/s for Bader + NCI)."""
    from oracle import oracle

    small = config(3, n_override=(n_bader,) * 3)
    f = gb.synth_slater(ctx, small)
    t0 = time.perf_counter()
    oracle.bader(f, small.n, small.x2c, small.xat)
    t_bader = (time.perf_counter() - t0) / f.size
    del f
    n = cell.n
    threads = host_threads()
    pts_n, t_nci = nci_cpu_time(cell, tables, species, [v // 2 - 64 for v in n[:2]] + [n[2] // 2 - 16],
                                (128, 128, 32), threads)
    t_nci /= pts_n
    return {"value": 1e-9 / (t_bader + t_nci), "unit": "Gpoints/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"oracle Bader (serial sweep + refinement, 1 thread, as the reference) on the same "
                      f"30 A / 216-atom cell at {n_bader}^3: {t_bader * 1e6:.3f} us/pt; promolecular NCI "
                      f"(OpenMP dynamic over nodes, {threads} threads) on {pts_n} nodes of the 512^3 grid: "
                      f"{t_nci * 1e6:.3f} us/pt; both rates applied to the whole grid",
            "bader_us_per_point": round(t_bader * 1e6, 4), "nci_us_per_point": round(t_nci * 1e6, 4)}


def run_batch(ctx, args):
    """The batching workload: independent periodic dimer cells with seeded
    geometry and grid edges drawn from {48,64,96,128,160}, FP64 Slater grids
    resident in HBM; one step = exact Bader labels + Volume/Pop sums + FP32
    promolecular NCI of every cell through c2gpu_batch_run."""
    from critic2_amd import batch as gbt
    from critic2_amd.synth import batch_cells
    cells = batch_cells(args.batch)
    zs = (1, 6, 7, 8)
    tables = [slater_grid1(z) for z in zs]
    species = [np.array([zs.index(int(z)) for z in c.zat], dtype=np.int32) for c in cells]
    grids = []
    for c in cells:
        g = ctx.empty((c.npts,), np.float64)
        gb.synth_slater(ctx, c, device_out=g)
        grids.append(g)
    points = sum(c.npts for c in cells)
    print(f"[bench] batch of {len(cells)} cells, {points} points resident", file=sys.stderr, flush=True)

    def step(cs, gs, sp, chunk=128):
        # the cells go through c2gpu_batch_run in chunks, with a progress line each
        agg = {"nstreams": 0, "queue_workgroups": 0, "ms_nci": 0.0}
        for c0 in range(0, len(cs), chunk):
            res, st = gbt.batch_run(ctx, cs[c0:c0 + chunk], gs[c0:c0 + chunk], tables, sp[c0:c0 + chunk],
                                    nstreams=args.nstreams)
            for r in res:
                for key in ("labels", "crho", "cgrad"):
                    if key in r:
                        r[key].free()
            agg["nstreams"] = st["nstreams"]
            agg["queue_workgroups"] = st["queue_workgroups"]
            agg["ms_nci"] += st["ms_nci"]
            print(f"[bench]   cells {c0 + len(res)}/{len(cs)}", file=sys.stderr, flush=True)
        return agg

    for w in range(args.warmup):
        step(cells[:8], grids[:8], species[:8])
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    ctx.sync()
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        st = step(cells, grids, species)
        print(f"[bench] step {s_ + 1}/{args.steps} done", file=sys.stderr, flush=True)
    ctx.sync()
    dt = (time.perf_counter() - t0) / args.steps
    # no single kernel dominates the batch (the sweep's list kernels of every
    # cell): the step as a whole against HBM, with its compulsory bytes
    ach = STEP_BYTES_PER_POINT * points / dt / 1e9
    roof = {"bound": "hbm", "kernel": "whole step (every kernel of every cell)", "achieved": round(ach, 3),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 6), "traffic": None,
            "compulsory_bytes_per_point": STEP_BYTES_PER_POINT, "points_per_step": points}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_batch(ctx, cells, tables, species)
        print("[bench] cpu baseline done", file=sys.stderr, flush=True)
    out = {
        "metric": f"Gpoints/s Bader+NCI, batch of {len(cells)} cells (labels exact)",
        "value": round(points / dt / 1e9, 6),
        "unit": "Gpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"{len(cells)} seeded dimer cells (6-40 atoms, gap 2.5-4.0 A), edges from "
                               "{48,64,96,128,160}; Bader labels + Volume/Pop + FP32 promolecular NCI",
                   "cells": len(cells), "points": points, "parallelism": "single GPU"},
        "hbm_frac_compulsory": round(ach / HBM_PEAK_GBS, 6),
        "roofline": roof,
        "cpu_baseline": cpu,
        "batch": {"nstreams": st["nstreams"], "queue_workgroups": st["queue_workgroups"],
                  "ms_nci_queue": round(st["ms_nci"], 3)},
    }
    print(json.dumps(out), flush=True)
    for g in grids:
        g.free()


def run_cube(ctx, args):
    """Text cube input: the 512^3 Slater-sum grid written as critic2's
    default (precisecube) cube text by the GPU writer, resident in HBM; one
    step = c2gpu_cube_parse of that text into an x-fastest FP64 grid
    (tokenise, scan, correctly rounded conversion, transpose).  The
    cpu_baseline is the oracle's strtod restatement of grid3%read_cube on
    the 256^3 precisecube of the 64-atom Slater-sum cell."""
    from critic2_amd import cube as gc
    cell = config(3, n_override=(args.n,) * 3)
    N = cell.npts
    rho = ctx.empty((N,), np.float64)
    gb.synth_slater(ctx, cell, device_out=rho)
    n = cell.n
    step = np.asarray(cell.x2c) / np.array(n, dtype=np.float64)[None, :]
    import ctypes as C
    import tempfile
    with tempfile.TemporaryDirectory() as td:  # the header, as the writer prints it
        hp = os.path.join(td, "h.cube")
        gc.write_cube(ctx, hp, n, (0.0, 0.0, 0.0), step, f=None, style=0)
        head = open(hp, "rb").read()
    nn = np.asarray(n, dtype=np.int32)
    ln = C.c_uint64()
    ctx.check(ctx.lib.c2gpu_cube_format(ctx.h, None, nn.ctypes.data, 0, 0, 0, None, 0, C.byref(ln)), "size")
    nb = len(head) + ln.value
    text = ctx.empty((nb,), np.uint8)
    hb = np.frombuffer(head, dtype=np.uint8)
    ctx.check(ctx.lib.c2gpu_memcpy(ctx.h, text.data_ptr(), hb.ctypes.data, len(head), 1), "c2gpu_memcpy")
    ctx.check(ctx.lib.c2gpu_cube_format(ctx.h, rho.data_ptr(), nn.ctypes.data, 0, 0, _lib.PTR_DEVICE,
                                        text.data_ptr() + len(head), ln.value, C.byref(ln)), "c2gpu_cube_format")
    f = ctx.empty((N,), np.float64)
    h = _lib.CubeHeader()
    print(f"[bench] cube text of {nb / 1e9:.3f} GB resident ({N} values)", file=sys.stderr, flush=True)

    def step_():
        ctx.check(ctx.lib.c2gpu_cube_parse(ctx.h, text.data_ptr(), nb, _lib.PTR_DEVICE, C.byref(h), f.data_ptr()),
                  "c2gpu_cube_parse")

    for w in range(args.warmup):
        step_()
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    ctx.timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        step_()
        print(f"[bench] step {s_ + 1}/{args.steps} done", file=sys.stderr, flush=True)
    ctx.sync()
    dt = (time.perf_counter() - t0) / args.steps
    kstats = ctx.kernel_stats()
    ctx.timing(False)
    # the parse must return the grid the text was written from, to 14 digits
    g = f.numpy()
    src = rho.numpy()
    rel = float(np.max(np.abs(g - src) / np.maximum(np.abs(src), 1e-300)))
    name = max(kstats.items(), key=lambda kv: kv[1][0])[0]
    ms_tot, launches, nbytes = kstats[name]
    avg_ms = ms_tot / launches
    ach = nbytes / launches / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": name, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name),
            "avg_launch_ms": round(avg_ms, 4), "launches_per_step": launches // args.steps,
            "algorithmic_bytes_per_launch": int(nbytes / launches)}
    # the writer over the same grid (formatting into the resident text buffer)
    ctx.timing(True)
    ctx.sync()
    t2 = time.perf_counter()
    for s_ in range(args.steps):
        ctx.check(ctx.lib.c2gpu_cube_format(ctx.h, rho.data_ptr(), nn.ctypes.data, 0, 0, _lib.PTR_DEVICE,
                                            text.data_ptr() + len(head), ln.value, C.byref(ln)), "c2gpu_cube_format")
    ctx.sync()
    dtw = (time.perf_counter() - t2) / args.steps
    wst = ctx.kernel_stats()
    ctx.timing(False)
    wms, wl, wb = wst["k_cube_format"]
    wach = wb / wl / (wms / wl * 1e-3) / 1e9
    writer = {"value": round(N / dtw / 1e6, 3), "unit": "Mvalues/s", "ms_per_step": round(dtw * 1e3, 3),
              "roofline": {"bound": "hbm", "kernel": "k_cube_format", "achieved": round(wach, 2),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(wach / HBM_PEAK_GBS, 4),
                           "avg_launch_ms": round(wms / wl, 4), "algorithmic_bytes_per_launch": int(wb / wl)},
              "kernels_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in
                                      sorted(wst.items(), key=lambda kv: -kv[1][0])[:4]}}
    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        small = config(2)
        rs = gb.synth_slater(ctx, small)
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "s.cube")
            stp = np.asarray(small.x2c) / np.array(small.n, dtype=np.float64)[None, :]
            gc.write_cube(ctx, path, small.n, (0.0, 0.0, 0.0), stp, f=rs, style=0)
            t1 = time.perf_counter()
            _, og = oracle.read_cube(path)
            tc = time.perf_counter() - t1
        cpu = {"value": round(small.npts / tc / 1e6, 3), "unit": "Mvalues/s", "cores": 1, "kind": "port",
               "sample": f"oracle read_cube (strtod) of a {small.n[0]}^3 precise cube, {small.npts} values"}
    value = N / dt / 1e6
    out = {
        "metric": f"Mvalues/s text cube parse @{n[0]}^3 (correctly rounded, x-fastest FP64 out)",
        "value": round(value, 3), "unit": "Mvalues/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64+f64", "data": "synthetic",
        "config": {"workload": f"precisecube text (E22.14E3, 6 per record) of the {n[0]}^3 Slater-sum grid, "
                               f"{nb} bytes resident in HBM", "grid": list(n), "bytes": nb,
                   "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu,
        "speedup_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "max_rel_diff_vs_written_grid": rel,
        "writer": writer,
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in
                                sorted(kstats.items(), key=lambda kv: -kv[1][0])[:6]},
    }
    print(json.dumps(out), flush=True)


def run_yt(ctx, args):
    """Yu-Trinkle integration of the 512^3 workload's FP64 grid (resident in
    HBM): basins, surface weights, and Volume/Pop weight sums through c2gpu_yt.
    cpu_baseline: the oracle's restatement of yt_integrate/yt_weights on the
    same cell at 96^3."""
    from critic2_amd import yt as gy
    cell = config(3, n_override=(args.n,) * 3)
    N = cell.npts
    rho = ctx.empty((N,), np.float64)
    gb.synth_slater(ctx, cell, device_out=rho)
    print(f"[bench] grid {cell.n} resident, {N} points", file=sys.stderr, flush=True)
    res = None
    for w in range(args.warmup):
        res = gy.yt_integrate(ctx, rho, cell.n, cell.x2c, cell.xat, fields=[rho])
        res["labels"].free()
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    ctx.timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        res = gy.yt_integrate(ctx, rho, cell.n, cell.x2c, cell.xat, fields=[rho])
        res["labels"].free()
        print(f"[bench] step {s_ + 1}/{args.steps} done", file=sys.stderr, flush=True)
    ctx.sync()
    dt = (time.perf_counter() - t0) / args.steps
    kstats = ctx.kernel_stats()
    ctx.timing(False)
    reporting = {k: v for k, v in kstats.items() if v[2] > 0}
    name = max(reporting.items(), key=lambda kv: kv[1][0])[0]
    ms_tot, launches, nbytes = kstats[name]
    avg_ms = ms_tot / launches
    ach = nbytes / launches / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": name, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name),
            "avg_launch_ms": round(avg_ms, 4), "launches_per_step": launches // args.steps,
            "algorithmic_bytes_per_launch": int(nbytes / launches)}
    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        small = config(3, n_override=(96, 96, 96))
        fs = gb.synth_slater(ctx, small)
        vec, area = _lib.ws_facets(np.asarray(small.x2c) / np.array(small.n)[None, :])
        t1 = time.perf_counter()
        oracle.yt(fs, small.n, small.x2c, small.xat, vec, area, fields=[fs])
        tc = time.perf_counter() - t1
        cpu = {"value": small.npts / tc / 1e9, "unit": "Gpoints/s", "cores": 1, "kind": "port",
               "sample": f"oracle yt_integrate + yt_weights + Volume/Pop sums on the same 30 A / 216-atom "
                         f"cell at 96^3 ({small.npts} points), 1 thread"}
    value = N / dt / 1e9
    st = res["stats"]
    out = {
        "metric": f"Gpoints/s YT integration @{cell.n[0]}^3 (labels and weights exact)",
        "value": round(value, 6), "unit": "Gpoints/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"cubic cell a=30 A, 216 seeded atoms, Slater-sum FP64 grid "
                               f"{cell.n[0]}x{cell.n[1]}x{cell.n[2]}; YT basins + weights + Volume/Pop",
                   "grid": list(cell.n), "atoms": int(len(cell.zat)), "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu,
        "speedup_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "yt": {k: st[k] for k in ("nattr", "nmaxima", "nias", "label_passes", "weight_passes")},
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in
                                sorted(kstats.items(), key=lambda kv: -kv[1][0])[:8]},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--n", type=int, default=512, help="grid edge (default: the 512^3 workload)")
    ap.add_argument("--config", type=int, default=3,
                    help="workload: 3 = the 512^3 cubic cell (default), 5 = the 768^3 triclinic cell with 500 "
                         "atoms (Z up to 26)")
    ap.add_argument("--batch", type=int, default=0,
                    help="instead of the 512^3 step, time the batched path on this many dimer cells "
                         "(512 = the batching workload: edges from {48,64,96,128,160}, 6-40 atoms)")
    ap.add_argument("--nstreams", type=int, default=0, help="batched path: concurrent Bader workers (0: 8)")
    ap.add_argument("--yt", action="store_true",
                    help="instead of the Bader+NCI step, time Yu-Trinkle integration on the --n^3 grid")
    ap.add_argument("--cube", action="store_true",
                    help="instead of the Bader+NCI step, time the text cube parser on the --n^3 grid's "
                         "precisecube text")
    ap.add_argument("--host-buffers", action="store_true",
                    help="time the same step with the grid, labels and NCI outputs in host memory "
                         "(PCIe-inclusive rate)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a streaming kernel of known traffic (k_sqrt of the FFT |grad| grid: "
                         "8 B read + 8 B written per point, 256^3) to scale rocprof FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this path runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    ctx = _lib.Context(0)
    if args.batch:
        run_batch(ctx, args)
        return
    if args.cube:
        run_cube(ctx, args)
        return
    if args.yt:
        run_yt(ctx, args)
        return
    if args.host_buffers:
        run_host_buffers(ctx, args)
        return
    if args.copy_probe:
        probe = config(2)
        f = gb.synth_slater(ctx, probe)
        gb.fft_deriv(ctx, f, probe.n, probe.x2c, _lib.FT_GRAD)
        print(json.dumps({"copy_probe": "k_sqrt", "points": probe.npts,
                          "bytes_read": 8 * probe.npts, "bytes_written": 8 * probe.npts}), flush=True)
        return
    cell = config(args.config, n_override=(args.n,) * 3 if args.config == 3 else None)
    N = cell.npts
    tables, species = species_of(cell)
    rho = ctx.empty((N,), np.float64)
    gb.synth_slater(ctx, cell, device_out=rho)
    lab = ctx.empty((N,), np.int32)
    outs32 = (ctx.empty((N,), np.float32), ctx.empty((N,), np.float32))
    print(f"[bench] grid {cell.n} resident, {N} points", file=sys.stderr, flush=True)

    for w in range(args.warmup):
        run_step(ctx, cell, rho, lab, outs32, tables, species, "fp32")
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    ctx.timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for s in range(args.steps):
        bas = run_step(ctx, cell, rho, lab, outs32, tables, species, "fp32")
        print(f"[bench] step {s + 1}/{args.steps} done", file=sys.stderr, flush=True)
    ctx.sync()
    dt = (time.perf_counter() - t0) / args.steps
    kstats = ctx.kernel_stats()
    ktimes = {k: v[:2] for k, v in kstats.items()}
    ctx.timing(False)
    st = bas.stats

    # dominant kernel and its roofline: achieved = the kernel family's
    # compulsory bytes per point x N / the kernel's average launch time
    # (HIP events on the library's stream); the kernel's own count of the
    # bytes its design reads and writes is reported beside it
    dom = max(ktimes.items(), key=lambda kv: kv[1][0])
    name, (ms_tot, launches) = dom
    avg_ms = ms_tot / launches
    reported = kstats[name][2]
    if reported > 0:
        own_bytes = reported / launches
    else:
        ev_per_launch = st["events_total"] / max(st["sweep_iters"], 1)
        own_bytes = kernel_bytes(name, N, ev_per_launch, st["traj_points"])
    bpp = compulsory_bytes_per_point(name)
    ach = bpp * N / (avg_ms * 1e-3) / 1e9
    # the committed PMC summaries are of the 512^3 step: no traffic for other workloads
    measured = args.config == 3 and args.n == 512
    roof = {"bound": "hbm", "kernel": name, "achieved": round(ach, 3), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 6),
            "traffic": measured_traffic(name) if measured else None,
            "avg_launch_ms": round(avg_ms, 4), "launches_per_step": launches // args.steps,
            "compulsory_bytes_per_point": bpp, "points_per_launch": N,
            "algorithmic_bytes_per_launch": int(bpp * N),
            "kernel_design_bytes_per_launch": int(own_bytes) if own_bytes is not None else None,
            "achieved_kernel_design": round(own_bytes / (avg_ms * 1e-3) / 1e9, 3) if own_bytes else None}
    # the FP32 promolecular evaluator is VALU-bound: FLOP/s against the FP32 vector peak
    valu = None
    pk = next((k for k in ("k_promol_f", "k_promol_f_queue") if k in ktimes), None)
    if pk is not None:
        pms = ktimes[pk][0] / ktimes[pk][1]
        fl = measured_flops(pk) if measured else None
        valu = {"kernel": pk, "avg_launch_ms": round(pms, 4),
                "flops_per_launch": fl,
                "achieved_tflops": round(fl / (pms * 1e-3) / 1e12, 3) if fl else None,
                "peak_tflops_fp32": FP32_PEAK_TFLOPS,
                "frac": round(fl / (pms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4) if fl else None}

    fp64 = None
    precision = None
    if not args.no_fp64:
        outs64 = (ctx.empty((N,), np.float64), ctx.empty((N,), np.float64))
        ctx.sync()
        t1 = time.perf_counter()
        run_step(ctx, cell, rho, lab, outs64, tables, species, "fp64")
        t64 = time.perf_counter() - t1
        fp64 = {"value": N / t64 / 1e9, "ms_per_step": t64 * 1e3}
        print("[bench] fp64 step done", file=sys.stderr, flush=True)
        # FP32 outputs of the timed steps against the FP64 outputs, whole grid
        precision = gn.precision_report(outs32[0].numpy(), outs32[1].numpy(), outs64[0].numpy(),
                                        outs64[1].numpy(),
                                        zeta_max=2.0 * max(int(z) for z in cell.zat) ** (1.0 / 3.0))
        for o in outs64:
            o.free()
        print("[bench] precision report done", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, cell, tables, species)
        print("[bench] cpu baseline done", file=sys.stderr, flush=True)

    value = N / dt / 1e9
    top = sorted(ktimes.items(), key=lambda kv: -kv[1][0])[:8]
    out = {
        "metric": f"Gpoints/s Bader+NCI @{cell.n[0]}^3, mixed vs FP64 GPU (labels exact); %HBM peak"
                  if args.config == 3
                  else f"Gpoints/s Bader+NCI @{cell.n[0]}^3 triclinic, mixed (labels exact); %HBM peak",
        "value": round(value, 6),
        "unit": "Gpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": (("cubic cell a=30 A, 216 seeded atoms (Z in 1,6,7,8)" if args.config == 3 else
                                  "triclinic cell a=b=c=40 A, 500 seeded atoms (Z in 1,6,7,8,14,26)") +
                                 f", Slater-sum FP64 grid {cell.n[0]}x{cell.n[1]}x{cell.n[2]}; Bader labels + "
                                 "Volume/Pop + FP32 promolecular NCI"),
                   "grid": list(cell.n), "atoms": int(len(cell.zat)), "parallelism": "single GPU"},
        "hbm_frac_compulsory": round(STEP_BYTES_PER_POINT * N / dt / 1e9 / HBM_PEAK_GBS, 6),
        "roofline": roof,
        "promol_valu": valu,
        "fp32_vs_fp64": precision,
        "cpu_baseline": cpu,
        "fp64_gpu": fp64,
        "speedup_vs_fp64_gpu": round(fp64["ms_per_step"] / (dt * 1e3), 3) if fp64 else None,
        "speedup_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "bader": {k: st[k] for k in ("sweep_iters", "refine_iters", "traj_points", "events_total",
                                     "ms_traj", "ms_sweep", "ms_refine", "ms_total")},
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in top},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
