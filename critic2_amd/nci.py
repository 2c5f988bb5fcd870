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


def cutoffs(ctx: Context, crho, cgrad, rhocut=0.2, dimcut=2.0, rhoplot=0.1, srhorange=(-1e30, 1e30),
            onlyneg=False, rhoat=None, rho_void=-1.0, dat=True):
    """nciplot's cutoff pass (nci@proc.f90:612-654, no fragments) on the GPU
    through c2gpu_nci_cutoff.  crho, cgrad (and rhoat with rho_void > 0):
    z-fastest arrays of one type (float64 or float32), host numpy arrays or
    DeviceArrays.  cgrad is masked in place (s <- 100).  Returns (dat_text,
    sums): the .dat rows '(1p,E15.7,E15.7)' as bytes (None if dat is False)
    and the unscaled void sums (charge, promolecular charge, node count)."""
    import ctypes as C
    dev = hasattr(crho, "data_ptr")
    dt = np.dtype(crho.dtype)
    prec = _lib.FP64 if dt == np.float64 else _lib.FP32
    if not dev:
        if not (isinstance(cgrad, np.ndarray) and cgrad.dtype == dt and cgrad.flags.c_contiguous):
            raise ValueError("cgrad must be a contiguous array of crho's type (it is masked in place)")
        crho = np.ascontiguousarray(crho)
        if rhoat is not None:
            rhoat = np.ascontiguousarray(rhoat, dtype=dt)
    npts = int(np.prod(crho.shape))
    n = np.array([npts, 1, 1], dtype=np.int32)
    o = _lib.NciCut(rhocut, dimcut, rhoplot, (C.c_double * 2)(*srhorange), int(bool(onlyneg)), rho_void)
    flags = _lib.PTR_DEVICE if dev else 0
    ln = C.c_uint64()
    sums = np.zeros(3)
    args = (ctx.h, _lib.ptr(crho), _lib.ptr(cgrad), _lib.ptr(rhoat), n.ctypes.data, prec, C.byref(o), flags)
    ctx.check(ctx.lib.c2gpu_nci_cutoff(*args, None, 0, C.byref(ln), None), "c2gpu_nci_cutoff")
    size = max(ln.value, 1)
    buf = ctx.empty((size,), np.uint8) if dev else np.empty(size, dtype=np.uint8)
    ctx.check(ctx.lib.c2gpu_nci_cutoff(*args, _lib.ptr(buf), size, C.byref(ln), sums.ctypes.data),
              "c2gpu_nci_cutoff")
    text = None
    if dat:
        text = (buf.numpy() if dev else buf)[:ln.value].tobytes()
    return text, sums


def dat_rows(text: bytes) -> np.ndarray:
    """The (rho, s) pairs of .dat text."""
    if not text:
        return np.zeros((0, 2))
    return np.array(text.split(), dtype=np.float64).reshape(-1, 2)
