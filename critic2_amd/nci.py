"""Host-side mirror of critic2's NCIPLOT per-point block on periodic grids.

``nciplot_grid`` is the loaded-grid branch with Fourier derivatives
(nci@proc.f90:525-564, the default ``fourierint`` for grid fields) or, with
interpolation="grid", the tricubic grid interpolant (:565-571);
``nciplot_promol`` is the promolecular branch (:565-571 with
crystalmod@env.f90:622-753).  Both return crho = sign(rho, lambda2)*100 and
cgrad = s, z-fastest as nci@proc.f90:502-505 stores them, unless
zfast=False.  ``cutoffs`` is the cutoff pass of :615-643.

nciplot's node sets (``cell_nodes``, ``box_nodes``, ``fragment_box``,
nci@proc.f90:384-446), the promolecular block on any of them with the
promolecular density rhoat (``promol_nodes``), and the FRAGMENT densities
(``fragment_atoms``, ``fragments``, :580-596, read_fragment :949-1003) feed
the cutoff pass with fragments (:622-626).
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
    dt = np.float32 if out_f32 else np.float64
    if outs is None:
        outs = (np.zeros(npts, dtype=dt), np.zeros(npts, dtype=dt))
    crho, cgrad = outs
    flags = _out_flags(outs, dt, npts)
    if rho is not None:
        if hasattr(rho, "data_ptr") != (flags == _lib.PTR_DEVICE):
            raise ValueError("the grid and the outputs must both be device arrays or both host arrays")
        if flags == 0:
            rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
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
            onlyneg=False, rhoat=None, rho_void=-1.0, dat=True, rhofrag=None, rhoparam=0.95, rhoparam2=0.75):
    """nciplot's cutoff pass (nci@proc.f90:612-654) on the GPU through
    c2gpu_nci_cutoff, or c2gpu_nci_cutoff_frag with fragments.  crho, cgrad
    (and rhoat with rho_void > 0 or fragments): z-fastest arrays of one type
    (float64 or float32), host numpy arrays or DeviceArrays; rhofrag:
    (nfrag, npts) fragment densities as ``fragments`` returns them, with
    RHOPARAM and RHOPARAM2 (defaults of nci@proc.f90:124-125).  cgrad is
    masked in place (s <- 100).  Returns (dat_text, sums): the .dat rows
    '(1p,E15.7,E15.7)' as bytes (None if dat is False) and the unscaled void
    sums (charge, promolecular charge, node count)."""
    import ctypes as C
    dev = hasattr(crho, "data_ptr")
    dt = _np_dtype(crho.dtype)
    if dt not in (np.dtype(np.float64), np.dtype(np.float32)):
        raise ValueError("crho must be float64 or float32")
    prec = _lib.FP64 if dt == np.float64 else _lib.FP32
    npts = int(np.prod(crho.shape))
    nfrag = 0 if rhofrag is None else int(rhofrag.shape[0])
    # the library reads and writes npts values of every array (cgrad in place):
    # sizes, the memory space and the value type are checked before the call
    need = [("cgrad", cgrad, (npts,))]
    if rhoat is not None:
        need.append(("rhoat", rhoat, (npts,)))
    elif nfrag or rho_void > 0:
        raise ValueError("rhoat is required with fragments or rho_void > 0")
    if rhofrag is not None:
        if len(tuple(rhofrag.shape)) != 2:
            raise ValueError("rhofrag must have shape (nfrag, npts)")
        need.append(("rhofrag", rhofrag, (nfrag, npts)))
    for name, a, shape in need:
        if hasattr(a, "data_ptr") != dev:
            raise ValueError(f"{name} and crho must both be device arrays or both host arrays")
        if _np_dtype(a.dtype) != dt:
            raise ValueError(f"{name} must have crho's type {dt}")
        if int(np.prod(tuple(a.shape))) != int(np.prod(shape)) or (len(shape) == 2 and tuple(a.shape) != shape):
            raise ValueError(f"{name} must hold {shape} values (crho has {npts})")
    if not dev:
        if not (isinstance(cgrad, np.ndarray) and cgrad.flags.c_contiguous):
            raise ValueError("cgrad must be a contiguous array of crho's type (it is masked in place)")
        crho = np.ascontiguousarray(crho)
        if rhoat is not None:
            rhoat = np.ascontiguousarray(rhoat, dtype=dt)
        if rhofrag is not None:
            rhofrag = np.ascontiguousarray(rhofrag, dtype=dt)
    n = np.array([npts, 1, 1], dtype=np.int32)
    o = _lib.NciCut(rhocut, dimcut, rhoplot, (C.c_double * 2)(*srhorange), int(bool(onlyneg)), rho_void)
    flags = _lib.PTR_DEVICE if dev else 0
    ln = C.c_uint64()
    sums = np.zeros(3)
    if nfrag:
        fn, name = ctx.lib.c2gpu_nci_cutoff_frag, "c2gpu_nci_cutoff_frag"
        args = (ctx.h, _lib.ptr(crho), _lib.ptr(cgrad), _lib.ptr(rhoat), _lib.ptr(rhofrag), nfrag,
                float(rhoparam), float(rhoparam2), n.ctypes.data, prec, C.byref(o), flags)
    else:
        fn, name = ctx.lib.c2gpu_nci_cutoff, "c2gpu_nci_cutoff"
        args = (ctx.h, _lib.ptr(crho), _lib.ptr(cgrad), _lib.ptr(rhoat), n.ctypes.data, prec, C.byref(o), flags)
    ctx.check(fn(*args, None, 0, C.byref(ln), None), name)
    size = max(ln.value, 1)
    buf = ctx.empty((size,), np.uint8) if dev else np.empty(size, dtype=np.uint8)
    ctx.check(fn(*args, _lib.ptr(buf), size, C.byref(ln), sums.ctypes.data), name)
    text = None
    if dat:
        text = (buf.numpy() if dev else buf)[:ln.value].tobytes()
    return text, sums


def dat_rows(text: bytes) -> np.ndarray:
    """The (rho, s) pairs of .dat text."""
    if not text:
        return np.zeros((0, 2))
    return np.array(text.split(), dtype=np.float64).reshape(-1, 2)


# ------------------------------------------------------------------ node sets and fragments

BOHRTOA = 0.52917720859  # param.F90:42
ATOMEPS = 1e-2           # global.f90:33, identify_atom's default distance (bohr)


def _nodes(nstep, x0, xmat) -> _lib.NciNodes:
    nd = _lib.NciNodes()
    nd.nstep[:] = [int(v) for v in nstep]
    nd.x0[:] = [float(v) for v in x0]
    nd.xmat[:] = [float(v) for v in np.asarray(xmat, dtype=np.float64).reshape(3, 3).T.reshape(-1)]
    return nd


def cell_nodes(x2c, nstep=None, xinc=0.1) -> _lib.NciNodes:
    """The periodic cell's nodes (nci@proc.f90:413-426): x0 = 0,
    xmat(:,c) = a_c / nstep(c), with nstep(c) = ceiling(|a_c| / xinc(c))
    when not given (a grid field passes its own n).  x2c: lattice vectors as
    columns (bohr); xinc in bohr (critic2's default 0.1)."""
    a = np.asarray(x2c, dtype=np.float64)
    if nstep is None:
        xi = np.broadcast_to(np.asarray(xinc, dtype=np.float64), (3,))
        nstep = [int(np.ceil(np.sqrt(a[:, c] @ a[:, c]) / xi[c])) for c in range(3)]
    xm = np.stack([a[:, c] / float(nstep[c]) for c in range(3)], axis=1)
    return _nodes(nstep, (0.0, 0.0, 0.0), xm)


def _box(x0, x1, nstep, xinc):
    x0, x1 = np.asarray(x0, dtype=np.float64), np.asarray(x1, dtype=np.float64)
    lo, hi = np.minimum(x0, x1), np.maximum(x0, x1)
    if nstep is None:  # nci@proc.f90:436-437
        xinc = np.broadcast_to(np.asarray(xinc, dtype=np.float64), (3,))
        nstep = np.ceil(np.abs(hi - lo) / xinc).astype(int)
    else:              # :438-439
        nstep = np.asarray(nstep, dtype=int)
        xinc = (hi - lo) / nstep.astype(np.float64)
    return _nodes(nstep, lo, np.diag(xinc))


def box_nodes(x2c, x0_cryst, x1_cryst, nstep=None, xinc=0.1) -> _lib.NciNodes:
    """The nodes of NCIPLOT's CUBE x0 x1 box (crystallographic corners,
    nci@proc.f90:427-443): lower Cartesian corner, step xinc per axis
    (bohr), nstep = ceiling(|x1 - x0| / xinc), or xinc = (x1 - x0)/nstep
    with NSTEP."""
    a = np.asarray(x2c, dtype=np.float64)
    return _box(a @ np.asarray(x0_cryst, dtype=np.float64), a @ np.asarray(x1_cryst, dtype=np.float64), nstep, xinc)


def fragment_box(frags, rthres=2.0, nstep=None, xinc=0.1) -> _lib.NciNodes:
    """The box NCIPLOT builds around its fragments when no CUBE is given
    (findlimits, nci@proc.f90:385-403): the atoms' bounding box widened by
    rthres (bohr, the default 2) on every side.  frags: list of (xyz, species)."""
    xyz = np.concatenate([np.asarray(f[0], dtype=np.float64).reshape(-1, 3) for f in frags])
    return _box(xyz.min(0) - rthres, xyz.max(0) + rthres, nstep, xinc)


def fragment_atoms(x2c, xat, species, coords_ang):
    """A FRAGMENT block as read_fragment reads it (nci@proc.f90:949-1003):
    Cartesian angstrom -> bohr, and only the positions that coincide with an
    atom of the crystal (some lattice translate within ATOMEPS, identify_atom)
    are kept, with that atom's species.  Returns (xyz_bohr (m, 3), species (m,))."""
    a = np.asarray(x2c, dtype=np.float64)
    c2x = np.linalg.inv(a)
    xa = np.asarray(xat, dtype=np.float64).reshape(-1, 3)
    sp = np.asarray(species).reshape(-1)
    shifts = np.indices((3, 3, 3)).reshape(3, -1).T - 1  # the 27 neighbouring translates
    keep, spc = [], []
    for x in np.asarray(coords_ang, dtype=np.float64).reshape(-1, 3) / BOHRTOA:
        d = c2x @ x - xa
        d -= np.round(d)
        dist = np.sqrt((((d[:, None, :] + shifts[None, :, :]) @ a.T) ** 2).sum(-1)).min(1)
        i = int(np.argmin(dist))
        if dist[i] < ATOMEPS:
            keep.append(x)
            spc.append(int(sp[i]))
    return np.array(keep).reshape(-1, 3), np.array(spc, dtype=np.int32)


def _np_dtype(dt):
    """numpy dtype of a numpy, DeviceArray or torch array's dtype."""
    try:
        return np.dtype(dt)
    except TypeError:  # a torch dtype
        return np.dtype(str(dt).replace("torch.", ""))


def _out_flags(outs, dtype, size):
    """C2GPU_PTR_DEVICE for device outputs, 0 for host arrays (written in
    place); a mix of the two is refused.  Every output must hold `size`
    values of `dtype`."""
    dev = [hasattr(o, "data_ptr") for o in outs if o is not None]
    for o in outs:
        if o is None:
            continue
        try:
            odt = np.dtype(o.dtype)
        except TypeError:  # a torch dtype
            odt = np.dtype(str(o.dtype).replace("torch.", ""))
        if odt != np.dtype(dtype) or int(np.prod(tuple(o.shape))) != size:
            raise ValueError(f"outputs must hold {size} values of {np.dtype(dtype).name}")
        if not hasattr(o, "data_ptr") and not (isinstance(o, np.ndarray) and o.flags.c_contiguous
                                               and o.flags.writeable):
            raise ValueError("host outputs must be writeable C-contiguous numpy arrays")
    if all(dev):
        return _lib.PTR_DEVICE
    if not any(dev):
        return 0
    raise ValueError("outputs must be all device arrays or all host arrays")


def promol_nodes(ctx: Context, nodes, x2c, tables: list[Grid1], xat, species, precision="fp64", outs=None):
    """The promolecular NCI block on an nciplot node set (c2gpu_nci_promol_nodes):
    (crho, cgrad, rhoat), z-fastest, of the precision's type."""
    prec = _lib.FP32 if precision == "fp32" else _lib.FP64
    npts = int(np.prod(list(nodes.nstep)))
    pm, keep = _promol_struct(tables, xat, species)
    x = _lib.mat_f(x2c)
    dt = np.float32 if prec == _lib.FP32 else np.float64
    if outs is None:
        outs = tuple(np.zeros(npts, dtype=dt) for _ in range(3))
    flags = _out_flags(outs, dt, npts)
    rc = ctx.lib.c2gpu_nci_promol_nodes(ctx.h, _lib.C.byref(nodes), x.ctypes.data, _lib.C.byref(pm), prec, flags,
                                        *(_lib.ptr(o) for o in outs))
    ctx.check(rc, "c2gpu_nci_promol_nodes")
    del keep
    return outs


def fragments(ctx: Context, nodes, tables: list[Grid1], frags, precision="fp64", out=None):
    """Fragment densities on an nciplot node set (c2gpu_nci_fragments):
    frags is a list of (xyz_bohr (m, 3), species (m,)); returns (nfrag, npts)."""
    prec = _lib.FP32 if precision == "fp32" else _lib.FP64
    npts = int(np.prod(list(nodes.nstep)))
    pm, keep = _promol_struct(tables, np.zeros((1, 3)), np.zeros(1, dtype=np.int32))
    nat = np.array([len(np.asarray(f[1]).reshape(-1)) for f in frags], dtype=np.int32)
    xyz = np.ascontiguousarray(np.concatenate([np.asarray(f[0], dtype=np.float64).reshape(-1, 3) for f in frags])
                               .reshape(-1)) if nat.sum() else np.zeros(3)
    spc = np.ascontiguousarray(np.concatenate([np.asarray(f[1], dtype=np.int32).reshape(-1) for f in frags])) \
        if nat.sum() else np.zeros(1, dtype=np.int32)
    dt = np.float32 if prec == _lib.FP32 else np.float64
    if out is None:
        out = np.zeros((len(frags), npts), dtype=dt)
    flags = _out_flags((out,), dt, len(frags) * npts)
    rc = ctx.lib.c2gpu_nci_fragments(ctx.h, _lib.C.byref(nodes), pm.nspecies, pm.tables, len(frags),
                                     nat.ctypes.data, xyz.ctypes.data, spc.ctypes.data, prec, flags, _lib.ptr(out))
    ctx.check(rc, "c2gpu_nci_fragments")
    del keep
    return out


# ------------------------------------------------------------------ mixed-precision error report

RDG_CONST = 2.0 * (3.0 * np.pi ** 2) ** (1.0 / 3.0)  # nci@proc.f90:91
RHO_BANDS = ((1e-6, 1e-4), (1e-4, 1e-2), (1e-2, np.inf))


def cancellation_bound(s64, rho64, zeta_max, eps=2.0 ** -24):
    """Relative FP32 error expected of s at each node: every atom's gradient
    term carries a relative error of about eps, and the terms (of total
    magnitude up to zeta_max * rho for exponential atomic densities) cancel
    down to |grad rho| = s * 2(3 pi^2)^(1/3) * rho^(4/3), so the error of the
    sum relative to it is eps * zeta_max / (s * 2(3 pi^2)^(1/3) * rho^(1/3)),
    plus eps for rho itself."""
    return eps * (1.0 + zeta_max / (np.maximum(s64, 1e-300) * RDG_CONST * np.cbrt(rho64)))


def precision_report(c32, s32, c64, s64, s_floor=0.1, rho_min=1e-6, zeta_max=None):
    """Error of the FP32 promolecular NCI outputs against the FP64 ones in
    nciplot's region s < 2 (nodes with rho > rho_min): max and RMS relative
    error of rho (|crho|/100), |grad rho| (s * 2(3 pi^2)^(1/3) rho^(4/3), from
    each path's own rho and s) and s.  Near a critical point of rho the
    gradient is a sum of atom terms that cancel, so the relative error of s
    grows as 1/s for any fixed precision; s is therefore reported relative
    for s_floor <= s < 2 and absolute for s < s_floor; "by_rho" repeats the
    relative errors per density band (RHO_BANDS).  Host arrays; returns a
    dict of floats and counts."""
    c32 = np.asarray(c32, dtype=np.float64).reshape(-1)
    c64 = np.asarray(c64, dtype=np.float64).reshape(-1)
    s32 = np.asarray(s32, dtype=np.float64).reshape(-1)
    s64 = np.asarray(s64, dtype=np.float64).reshape(-1)
    r32, r64 = np.abs(c32) / 100.0, np.abs(c64) / 100.0
    reg = (s64 < 2.0) & (r64 > rho_min)
    out = {"points": int(reg.sum()), "s_floor": s_floor, "rho_min": rho_min}
    if not out["points"]:
        return out

    def stats(e):
        return (float(e.max()), float(np.sqrt(np.mean(e * e)))) if e.size else (0.0, 0.0)

    e_rho = np.abs(r32[reg] - r64[reg]) / r64[reg]
    g32 = s32[reg] * RDG_CONST * r32[reg] ** (4.0 / 3.0)
    g64 = s64[reg] * RDG_CONST * r64[reg] ** (4.0 / 3.0)
    hi = s64[reg] >= s_floor
    e_g = np.abs(g32 - g64)[hi] / g64[hi]
    e_s = np.abs(s32[reg] - s64[reg])[hi] / s64[reg][hi]
    a_s = np.abs(s32[reg] - s64[reg])[~hi]
    out["rho_max_rel"], out["rho_rms_rel"] = stats(e_rho)
    out["gradmod_max_rel"], out["gradmod_rms_rel"] = stats(e_g)
    out["s_max_rel"], out["s_rms_rel"] = stats(e_s)
    out["s_below_floor_points"] = int(a_s.size)
    out["s_below_floor_max_abs"] = float(a_s.max()) if a_s.size else 0.0
    out["sign_agree"] = float(np.mean(np.sign(c32[reg]) == np.sign(c64[reg])))
    # the same by density band: the gradient's atom terms cancel most where
    # the density is lowest, so the FP32 error of |grad rho| and s grows there
    r_reg = r64[reg]
    bands = [(b_lo, b_hi, (r_reg >= b_lo) & (r_reg < b_hi)) for b_lo, b_hi in RHO_BANDS]
    out["by_rho"] = []
    for b_lo, b_hi, sel in bands:
        shi = sel & (s64[reg] >= s_floor)
        row = {"rho_lo": b_lo, "rho_hi": b_hi, "points": int(sel.sum())}
        row["rho_max_rel"], row["rho_rms_rel"] = stats(e_rho[sel])
        eg = np.abs(g32[shi] - g64[shi]) / g64[shi]
        es = np.abs(s32[reg][shi] - s64[reg][shi]) / s64[reg][shi]
        row["gradmod_max_rel"], row["gradmod_rms_rel"] = stats(eg)
        row["s_max_rel"], row["s_rms_rel"] = stats(es)
        out["by_rho"].append(row)
    if zeta_max is not None:
        # measured error of s over the cancellation bound (s >= s_floor), and
        # the largest error where that bound is below 1e-6
        b = cancellation_bound(s64[reg][hi], r64[reg][hi], zeta_max)
        ratio = e_s / b
        out["s_over_bound_max"] = float(ratio.max()) if ratio.size else 0.0
        well = b < 1e-6
        out["s_max_rel_well_conditioned"] = float(e_s[well].max()) if well.any() else 0.0
        out["well_conditioned_points"] = int(well.sum())
    return out
