"""Host mirror of critic2's Yu-Trinkle integration (INTEGRABLE + YT):
yt_integrate (yt@proc.f90:38-224) with the weights of yt_weights
(:399-531) summed as intgrid_fields does for imtype_yt
(integration@proc.f90:1208-1218, :1289-1300), over c2gpu_yt, and the YT
basin multipoles (:1314-1336) over c2gpu_yt_multipoles."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Context
from .bader import _attractor_opts, _mode


def yt_integrate(ctx: Context, rho, n, x2c, xat, fields=(), nnm=False, ratom=None, noatoms=False,
                 labels=True, cps=None) -> dict:
    """Basins and weights of the FP64 grid rho (x fastest; numpy array or
    DeviceArray, as are the fields).  Returns nattr, xattr (fractional),
    labels (basin per point in grid order, 0 on the interatomic surfaces),
    vol (sum of each basin's weights) and sums[q, a] (sum of weights times
    fields[q]); vol and sums are unscaled (critic2 multiplies by Omega/N).
    cps: known critical points other than the atoms, as for
    bader.bader_integrate."""
    n = tuple(int(v) for v in n)
    npts = n[0] * n[1] * n[2]
    mode, ra = _mode(nnm, ratom, noatoms)
    opts, xat_flat = _attractor_opts(mode, ra, x2c, xat, cps)
    dev = hasattr(rho, "data_ptr")
    flags = _lib.PTR_DEVICE if dev else 0
    if not dev:
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        fields = [np.ascontiguousarray(f, dtype=np.float64).reshape(-1) for f in fields]
    fptr = (C.c_void_p * max(len(fields), 1))(*[_lib.ptr(f) for f in fields])
    lab = None
    if labels:
        lab = ctx.empty((npts,), np.int32) if dev else np.zeros(npts, dtype=np.int32)
    cap = max(opts.nat, 1) * 4 + 4096
    xattr = np.zeros(3 * cap)
    vol = np.zeros(cap)
    sums = np.zeros(max(len(fields), 1) * cap)
    nattr = C.c_int(0)
    st = _lib.YtStats()
    nn = np.asarray(n, dtype=np.int32)
    m = np.asarray(x2c, dtype=np.float64)
    ctx.check(ctx.lib.c2gpu_yt(ctx.h, _lib.ptr(rho), nn.ctypes.data, _lib.mat_f(m).ctypes.data, C.byref(opts),
                               len(fields), fptr, flags, _lib.ptr(lab), C.byref(nattr), xattr.ctypes.data, cap,
                               vol.ctypes.data, sums.ctypes.data, C.byref(st)), "c2gpu_yt")
    na = nattr.value
    return {"nattr": na, "xattr": xattr[:3 * na].reshape(-1, 3).copy(), "labels": lab, "vol": vol[:na].copy(),
            "sums": sums[:len(fields) * na].reshape(len(fields), na).copy(), "stats": st.as_dict()}


def yt_multipoles(ctx: Context, labels, f, n, x2c, xattr, lmax: int = 5):
    """Basin multipoles of the field f over the YT partition of the last
    yt_integrate call on ctx (INTEGRABLE f MULTIPOLE lmax with YT,
    integration@proc.f90:1314-1336): array (nattr, (lmax+1)^2), each point
    weighted by its basin weight (1 inside, the surface weight on the
    interatomic surfaces), unscaled (critic2 multiplies by Omega/N).
    labels: that call's labels; f, xattr as for bader.multipoles."""
    n = tuple(int(v) for v in n)
    xattr = np.ascontiguousarray(xattr, dtype=np.float64).reshape(-1, 3)
    nattr = xattr.shape[0]
    nlm = (lmax + 1) ** 2
    dev = hasattr(labels, "data_ptr")
    if not dev:
        labels = np.ascontiguousarray(labels, dtype=np.int32).reshape(-1)
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    out = np.zeros(nattr * nlm)
    nn = np.asarray(n, dtype=np.int32)
    ctx.check(ctx.lib.c2gpu_yt_multipoles(ctx.h, _lib.ptr(labels), _lib.ptr(f), nn.ctypes.data,
                                          _lib.mat_f(x2c).ctypes.data, nattr, xattr.ctypes.data, lmax,
                                          _lib.PTR_DEVICE if dev else 0, out.ctypes.data), "c2gpu_yt_multipoles")
    return out.reshape(nattr, nlm)
