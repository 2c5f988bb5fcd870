"""Host-side mirror of critic2's NCIPLOT per-point block on periodic grids.

``nciplot_grid`` is the loaded-grid branch with Fourier derivatives
(nci@proc.f90:525-564, the default ``fourierint`` for grid fields) or, with
interpolation="grid", the tricubic grid interpolant (:565-571);
``nciplot_promol`` is the promolecular branch (:565-571 with
crystalmod@env.f90:622-753).  Both return crho = sign(rho, lambda2)*100 and
cgrad = s, z-fastest as nci@proc.f90:502-505 stores them, unless
zfast=False.  ``cutoffs`` is the cutoff pass of :615-643.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import Context
from .radial import Grid1


def _promol_struct(tables, xat, species):
    tabs = (_lib.RadialTable * len(tables))()
    keep = []
    for s, g in enumerate(tables):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (g.r, g.f, g.fp, g.fpp)]
        keep.extend(arrs)
        t = tabs[s]
        t.ngrid = g.ngrid
        t.a, t.b, t.rmax, t.cutrad = g.a, g.b, g.rmax, g.cutrad
        t.r, t.f, t.fp, t.fpp = (a.ctypes.data for a in arrs)
    xa = np.ascontiguousarray(np.asarray(xat, dtype=np.float64).reshape(-1))
    sp = np.ascontiguousarray(np.asarray(species, dtype=np.int32))
    keep.extend([xa, sp])
    pm = _lib.Promol()
    pm.nspecies = len(tables)
    pm.tables = tabs
    pm.nat = len(sp)
    pm.xat = xa.ctypes.data
    pm.species = sp.ctypes.data
    return pm, (tabs, keep)


def _run(ctx, rho, n, x2c, scheme, precision, pm, zfast, out_f32, outs=None):
    nn = np.asarray(n, dtype=np.int32)
    npts = int(nn.prod())
    x = _lib.mat_f(x2c)
    dev = outs is not None
    if dev:
        crho, cgrad = outs
    else:
        dt = np.float32 if out_f32 else np.float64
        crho = np.zeros(npts, dtype=dt)
        cgrad = np.zeros(npts, dtype=dt)
    if rho is not None and not hasattr(rho, "data_ptr"):
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
    flags = _lib.PTR_DEVICE if dev else 0
    fn = ctx.lib.c2gpu_nci_f32 if out_f32 else ctx.lib.c2gpu_nci
    rc = fn(ctx.h, _lib.ptr(rho) if rho is not None else None, nn.ctypes.data, x.ctypes.data, scheme,
            precision, _lib.C.byref(pm) if pm is not None else None, 1 if zfast else 0, flags,
            _lib.ptr(crho), _lib.ptr(cgrad))
    ctx.check(rc, "c2gpu_nci")
    return crho, cgrad


def nciplot_grid(ctx: Context, rho, n, x2c, zfast=True, outs=None, interpolation="fourier"):
    """NCI on a loaded FP64 grid.  interpolation="fourier" is the default
    fourierint branch (FFT derivatives, nci@proc.f90:546-564);
    interpolation="grid" is NCIPLOT's INTERPOLATION GRID keyword
    (nci@proc.f90:358-360): tricubic interpolant of the grid, :565-571."""
    if interpolation not in ("fourier", "grid"):
        raise ValueError("interpolation must be 'fourier' or 'grid'")
    scheme = _lib.NCI_FFT if interpolation == "fourier" else _lib.NCI_GRID
    return _run(ctx, rho, n, x2c, scheme, _lib.FP64, None, zfast, False, outs)


def nciplot_promol(ctx: Context, n, x2c, tables: list[Grid1], xat, species, precision="fp32",
                   zfast=True, outs=None, out_f32=None):
    """Promolecular NCI on the nodes of an n1 x n2 x n3 grid of the cell."""
    prec = _lib.FP32 if precision == "fp32" else _lib.FP64
    if out_f32 is None:
        out_f32 = prec == _lib.FP32
    pm, keep = _promol_struct(tables, xat, species)
    res = _run(ctx, None, n, x2c, _lib.NCI_PROMOL, prec, pm, zfast, out_f32, outs)
    del keep
    return res


def cutoffs(crho, cgrad, rhocut=0.2, dimcut=2.0, rhoplot=0.1, srhorange=(-1e30, 1e30),
            onlyneg=False):
    """Cutoff pass of nci@proc.f90:615-643 (no fragments): returns the
    (rho, s) rows of the .dat file and cgrad with 100 outside the plot range."""
    crho = np.asarray(crho, dtype=np.float64)
    cg = np.array(cgrad, dtype=np.float64, copy=True)
    sel = (np.abs(crho) < rhocut * 100.0) & (cg < dimcut)
    dat = np.stack([crho[sel] / 100.0, cg[sel]], axis=1)
    cg[np.abs(crho) > rhoplot * 100.0] = 100.0
    cg[(crho < srhorange[0] * 100.0) | (crho > srhorange[1] * 100.0)] = 100.0
    if onlyneg:
        cg[crho > 0] = 100.0
    return dat, cg
