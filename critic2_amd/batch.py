"""Host mirror of c2gpu_batch_run: Bader labels with Volume/Pop sums and the
FP32 promolecular NCI block over many independent cells in one call.

critic2 itself analyses one system per run; the batch entry gives each cell
exactly what bader_integrate + intgrid_fields (integration@proc.f90:1170-1391)
and the promolecular branch of nciplot (nci@proc.f90:565-605) give it alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Context
from .radial import Grid1


class _Cell(C.Structure):
    _fields_ = [("rho", C.c_void_p), ("n", C.c_int * 3), ("x2c", C.c_double * 9), ("car2lat", C.c_double * 9),
                ("lat_i_dist", C.c_double * 27), ("nat", C.c_int), ("xat", C.c_void_p), ("species", C.c_void_p),
                ("label_out", C.c_void_p), ("nattr_out", C.c_void_p), ("xattr_out", C.c_void_p),
                ("xattr_cap", C.c_int), ("counts_out", C.c_void_p), ("sums_out", C.c_void_p),
                ("crho_out", C.c_void_p), ("cgrad_out", C.c_void_p)]


class BatchStats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_nci", C.c_double), ("points", C.c_int64),
                ("nstreams", C.c_int), ("queue_workgroups", C.c_int)]


def batch_run(ctx: Context, cells, grids, tables: list[Grid1] | None = None, species=None, bader=True,
              nci=True, nnm=False, ratom=None, noatoms=False, nstreams=0):
    """cells: synth.Cell-like objects (n, x2c, xat); grids: their FP64 grids
    (x fastest), all host numpy arrays or all DeviceArrays (then labels and
    NCI outputs stay on the device too); tables/species: radial tables and,
    per cell, the table index of each atom (needed when nci).  Returns
    (results, stats): per cell a dict with labels, nattr, xattr, counts, sums
    (Bader) and crho, cgrad (NCI, FP32, z-fastest)."""
    from .bader import _mode
    mode, ra = _mode(nnm, ratom, noatoms)
    ncell = len(cells)
    dev = ncell > 0 and hasattr(grids[0], "data_ptr")
    arr = (_Cell * ncell)()
    keep, results = [], []
    for q, (cell, g) in enumerate(zip(cells, grids)):
        n = tuple(int(v) for v in cell.n)
        npts = n[0] * n[1] * n[2]
        c = arr[q]
        if not dev:
            g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
        if int(np.prod(g.shape)) != npts:
            raise ValueError(f"cell {q}: grid has {int(np.prod(g.shape))} points, expected {npts}")
        c.rho = _lib.ptr(g)
        for a in range(3):
            c.n[a] = n[a]
        for a, v in enumerate(_lib.mat_f(cell.x2c)):
            c.x2c[a] = v
        _, c2l, lid = _lib.bader_metrics(cell.x2c, n)
        for a, v in enumerate(c2l):
            c.car2lat[a] = v
        for a, v in enumerate(lid):
            c.lat_i_dist[a] = v
        xat = np.ascontiguousarray(np.asarray(cell.xat, dtype=np.float64).reshape(-1))
        c.nat = xat.size // 3
        c.xat = xat.ctypes.data
        res = {}
        keep.extend([g, xat])
        if nci:
            sp = np.ascontiguousarray(species[q], dtype=np.int32)
            keep.append(sp)
            c.species = sp.ctypes.data
            alloc = (lambda: ctx.empty((npts,), np.float32)) if dev else (lambda: np.zeros(npts, dtype=np.float32))
            res["crho"], res["cgrad"] = alloc(), alloc()
            c.crho_out = _lib.ptr(res["crho"])
            c.cgrad_out = _lib.ptr(res["cgrad"])
        if bader:
            cap = 4 * c.nat + 4096
            res["labels"] = ctx.empty((npts,), np.int32) if dev else np.zeros(npts, dtype=np.int32)
            res["_nattr"] = np.zeros(1, dtype=np.int32)
            res["_xattr"] = np.zeros(3 * cap)
            res["_counts"] = np.zeros(cap, dtype=np.int64)
            res["_sums"] = np.zeros(cap)
            c.label_out = _lib.ptr(res["labels"])
            c.nattr_out = res["_nattr"].ctypes.data
            c.xattr_out = res["_xattr"].ctypes.data
            c.xattr_cap = cap
            c.counts_out = res["_counts"].ctypes.data
            c.sums_out = res["_sums"].ctypes.data
        results.append(res)
    ntab = 0
    tabs = None
    if nci:
        from .nci import _promol_struct
        pm, tkeep = _promol_struct(tables, np.zeros((1, 3)), np.zeros(1, dtype=np.int32))
        keep.append(tkeep)
        tabs, ntab = pm.tables, len(tables)
    st = BatchStats()
    lib = ctx.lib
    fn = lib.c2gpu_batch_run
    fn.argtypes = [C.c_void_p, C.c_int, C.POINTER(_Cell), C.c_int, C.c_double, C.c_int,
                   C.POINTER(_lib.RadialTable), C.c_int, C.c_int, C.POINTER(BatchStats)]
    fn.restype = C.c_int
    flags = _lib.PTR_DEVICE if dev else 0
    ctx.check(fn(ctx.h, ncell, arr, mode, ra, ntab, tabs, nstreams, flags, C.byref(st)), "c2gpu_batch_run")
    for res in results:
        if "labels" in res:
            na = int(res.pop("_nattr")[0])
            res["nattr"] = na
            res["xattr"] = res.pop("_xattr")[: 3 * na].reshape(-1, 3)
            res["counts"] = res.pop("_counts")[:na]
            res["sums"] = res.pop("_sums")[:na]
    del keep
    return results, {k: getattr(st, k) for k, _ in st._fields_}
