"""Host-side mirror of critic2's BADER keyword on a loaded grid.

``bader_integrate`` follows bader_integrate(s, bas, iref)
(bader.f90:48-53, bader@proc.f90:80-234) and ``intgrid_fields`` the
per-basin Volume / Pop / Lap integration of intgrid_fields
(integration@proc.f90:1170-1391).  ``bader`` mirrors the BADER keyword as
run by intgrid_driver (integration@proc.f90:60-379): options, labelling,
attractor reordering (identity in the default mode, :1027-1166) and the
integrated properties.  The arithmetic runs in libc2gpu on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import Context, C2GpuError


@dataclass
class BasinData:
    """The parts of critic2's basindat (types.f90:361-383) this path fills."""

    n: tuple
    atexist: bool = True
    ratom: float = 1e40
    nattr: int = 0
    xattr: np.ndarray = None       # (nattr, 3) fractional
    idg: np.ndarray = None         # labels, Fortran order (n1,n2,n3) flattened x-fastest
    icp: np.ndarray = None
    stats: dict = field(default_factory=dict)


def _mode(nnm: bool, ratom, noatoms: bool):
    """Option resolution of intgrid_driver (integration@proc.f90:236-248)."""
    if noatoms:
        return _lib.ATTR_NOATOMS, (1.0 if ratom is None else ratom)
    if nnm or ratom is not None:
        return _lib.ATTR_RATOM, (1.0 if ratom is None else ratom)
    return _lib.ATTR_NEAREST_ATOM, 1e40


def _attractor_opts(mode, ratom, x2c, xat, cps=None):
    """c2gpu_attractor_opts for atoms xat and other known critical points
    cps (both fractional); returns the struct and the array it points to."""
    xat = np.ascontiguousarray(xat, dtype=np.float64).reshape(-1, 3)
    opts = _lib.AttractorOpts()
    opts.mode = mode
    opts.ratom = ratom
    for q, v in enumerate(_lib.mat_f(x2c)):
        opts.x2c[q] = v
    opts.nat_atoms = xat.shape[0]
    if cps is not None and len(cps):
        xat = np.concatenate([xat, np.asarray(cps, dtype=np.float64).reshape(-1, 3)])
    opts.nat = xat.shape[0]
    xat_flat = np.ascontiguousarray(xat.reshape(-1))
    opts.xat = xat_flat.ctypes.data
    return opts, xat_flat


def bader_integrate(ctx: Context, rho, n, x2c, xat, nnm=False, ratom=None, noatoms=False,
                    car2lat=None, lat_i_dist=None, labels_out=None, discard=None, cps=None) -> BasinData:
    """Near-grid Bader labelling of the FP64 grid rho (x fastest, n1*n2*n3).

    discard (BADER's DISCARD keyword, bader@proc.f90:185-191): None, or a
    function of the Cartesian position (bohr) of a maximum that would become
    a new attractor, true to discard it; its basin keeps label 0.

    cps: the field's critical points other than the atoms (fractional,
    (m, 3), cpcel order after the atoms), known attractors before the sweep
    as in bader@proc.f90:109-117; atoms keep labels 1..nat, cps nat+1...

    rho may be a numpy array (host) or a DeviceArray resident on the GPU;
    labels_out (optional) may likewise be a device int32 DeviceArray that
    receives the labels without a host copy.  car2lat / lat_i_dist default to the reference's
    metrics (bader@proc.f90:123-145)."""
    n = tuple(int(v) for v in n)
    npts = n[0] * n[1] * n[2]
    mode, ra = _mode(nnm, ratom, noatoms)
    if car2lat is None or lat_i_dist is None:
        _, car2lat, lat_i_dist = _lib.bader_metrics(x2c, n)
    opts, xat_flat = _attractor_opts(mode, ra, x2c, xat, cps)
    cb = None
    if discard is not None:
        cb = _lib.DISCARD_FN(lambda x, user: 1 if discard(np.array([x[0], x[1], x[2]])) else 0)
        opts.discard = cb
    nn = np.asarray(n, dtype=np.int32)
    c2l = np.ascontiguousarray(car2lat, dtype=np.float64)
    lid = np.ascontiguousarray(lat_i_dist, dtype=np.float64)
    on_device = hasattr(rho, "data_ptr")
    flags = _lib.PTR_DEVICE if on_device else 0
    if on_device:
        if labels_out is None:
            labels_out = ctx.empty((npts,), np.int32)
        lab = labels_out
    else:
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        lab = np.zeros(npts, dtype=np.int32)
    cap = max(opts.nat, 1) * 4 + 4096  # non-nuclear maxima of the NNM/NOATOMS rules append
    xattr = np.zeros(3 * cap)
    nattr = _lib.C.c_int(0)
    st = _lib.BaderStats()
    rc = ctx.lib.c2gpu_bader(ctx.h, _lib.ptr(rho), nn.ctypes.data, c2l.ctypes.data, lid.ctypes.data,
                             _lib.C.byref(opts), flags, _lib.ptr(lab), _lib.C.byref(nattr),
                             xattr.ctypes.data, cap, _lib.C.byref(st))
    ctx.check(rc, "c2gpu_bader")
    if nattr.value > cap:
        raise RuntimeError(f"c2gpu_bader: {nattr.value} attractors exceed the {cap} attractor slots")
    bas = BasinData(n=n, atexist=(mode != _lib.ATTR_NOATOMS), ratom=ra, nattr=nattr.value,
                    xattr=xattr[:3 * nattr.value].reshape(-1, 3).copy(), idg=lab,
                    stats=st.as_dict())
    bas.icp = np.arange(1, bas.nattr + 1, dtype=np.int32)  # default mode: atoms in cell order
    return bas


def basin_sums(ctx: Context, labels, nattr: int, fields=()):
    """count(idg == i) and sum(field, idg == i) for i = 1..nattr."""
    on_device = hasattr(labels, "data_ptr")
    npts = int(labels.numel() if on_device else labels.size)
    flags = _lib.PTR_DEVICE if on_device else 0
    if on_device:
        # one memory space per call: host fields go to the device first
        fields = [f if hasattr(f, "data_ptr") else ctx.to_device(np.ascontiguousarray(f, dtype=np.float64))
                  for f in fields]
    else:
        labels = np.ascontiguousarray(labels, dtype=np.int32).reshape(-1)
        fields = [f.numpy() if hasattr(f, "data_ptr") else f for f in fields]
        fields = [np.ascontiguousarray(f, dtype=np.float64).reshape(-1) for f in fields]
    nf = len(fields)
    fp = (_lib.C.c_void_p * max(nf, 1))(*[_lib.ptr(f) for f in fields])
    sums = np.zeros(max(nf, 1) * nattr)
    counts = np.zeros(nattr, dtype=np.int64)
    rc = ctx.lib.c2gpu_basin_sums(ctx.h, _lib.ptr(labels), npts, nattr, nf, fp, flags,
                                  sums.ctypes.data, counts.ctypes.data)
    ctx.check(rc, "c2gpu_basin_sums")
    return counts, sums[: nf * nattr].reshape(nf, nattr)


def fft_deriv(ctx: Context, rho, n, x2c, kind):
    """grid3%fft (grid3mod@proc.f90:1757-1872) on a host grid."""
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
    out = np.zeros_like(rho)
    nn = np.asarray(n, dtype=np.int32)
    x = _lib.mat_f(x2c)
    rc = ctx.lib.c2gpu_fft_deriv(ctx.h, rho.ctypes.data, nn.ctypes.data, x.ctypes.data, kind, 0,
                                 out.ctypes.data)
    ctx.check(rc, "c2gpu_fft_deriv")
    return out


def intgrid_fields(ctx: Context, bas: BasinData, rho, x2c, lap=True):
    """Volume, Pop and (optionally) Lap per basin, scaled by Omega/N as in
    integration@proc.f90:1215 and :1296.  Lap integrates the FFT Laplacian
    of the grid (:1249)."""
    n = bas.n
    npts = n[0] * n[1] * n[2]
    omega = abs(float(np.linalg.det(np.asarray(x2c))))
    fields = [rho]
    if lap:
        host_rho = rho.numpy() if hasattr(rho, "data_ptr") else rho
        lapg = fft_deriv(ctx, host_rho, n, x2c, _lib.FT_LAP)
        if hasattr(rho, "data_ptr"):
            lapg = ctx.to_device(lapg)
        fields.append(lapg)
    counts, sums = basin_sums(ctx, bas.idg, bas.nattr, fields)
    w = omega / npts
    out = {"volume": counts * w, "pop": sums[0] * w}
    if lap:
        out["lap"] = sums[1] * w
    return out


TYP_NUC = -3  # critical-point type of maxima (atoms and non-nuclear maxima)


def _shortest_dist(x2c, xd):
    """Length of the shortest lattice translate of the fractional vector xd
    (crystal%shortest, here by search over translations -2..2 after the
    nint fold; exact for the reduced cells of this path)."""
    m = np.asarray(x2c, dtype=np.float64)
    x0 = np.asarray(xd, dtype=np.float64) - np.round(xd)
    t = np.array([(a, b, c) for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)], dtype=np.float64)
    v = (x0[None, :] + t) @ m.T
    return float(np.sqrt(np.min(np.sum(v * v, axis=1))))


def reorder_gridout(ctx: Context, labels, xattr, x2c, cpcel, cptyp, nat_atoms, ratom, atexist=True, yt=False):
    """int_reorder_gridout (integration@proc.f90:1027-1166) for BADER and YT:
    the map of reorder_map applied to the labels (numpy array or
    DeviceArray, in place) through c2gpu_relabel (the `where` loop,
    :1113-1122) and with yt=True to the surface weights kept by the last
    c2gpu_yt call through c2gpu_yt_remap (:1125-1158).  Returns
    reorder_map's dict."""
    r = reorder_map(xattr, x2c, cpcel, cptyp, nat_atoms, ratom, atexist)
    assigned = r["assigned"]
    nattr0 = assigned.size
    if np.any(assigned != np.arange(1, nattr0 + 1)):
        dev = hasattr(labels, "data_ptr")
        lab = labels if dev else labels.reshape(-1)
        if not dev and not (lab.dtype == np.int32 and lab.flags.c_contiguous):
            raise ValueError("reorder_gridout: host labels must be a contiguous int32 array")
        npts = lab.numel() if dev else lab.size
        ctx.check(ctx.lib.c2gpu_relabel(ctx.h, npts, assigned.ctypes.data, nattr0,
                                        _lib.PTR_DEVICE if dev else 0, _lib.ptr(lab)), "c2gpu_relabel")
    if yt:
        ctx.check(ctx.lib.c2gpu_yt_remap(ctx.h, assigned.ctypes.data, nattr0, r["nattr"]), "c2gpu_yt_remap")
    return r


def reorder_map(xattr, x2c, cpcel, cptyp, nat_atoms, ratom, atexist=True):
    """The attractor map of int_reorder_gridout (integration@proc.f90:
    1027-1110), host logic.

    The attractors found (xattr, fractional, (nattr0, 3)) are matched to the
    field's critical points cpcel ((ncpcel, 3), types cptyp, the first
    nat_atoms being the atoms): an atom within ratom (identify_atom), else
    the nearest maximum-type point if nearer than ratom (nearest_cp with
    type typnuc); the critical points so matched become attractors 1.. in
    cpcel order, the others new non-nuclear maxima after them, merged within
    ratom (are_lclose).  As in the reference, attractors matched to a
    critical point keep that point's cpcel index as their new label.
    Returns dict(nattr, xattr, icp, assigned, new_cps): the new attractor
    count and positions, icp (cpcel index, 0 for a new maximum), assigned
    (new label of each old attractor) and the positions that ff%addcp would
    add to the critical points."""
    xattr = np.asarray(xattr, dtype=np.float64).reshape(-1, 3)
    cpcel = np.asarray(cpcel, dtype=np.float64).reshape(-1, 3)
    cptyp = np.asarray(cptyp, dtype=np.int32).reshape(-1)
    nattr0 = xattr.shape[0]
    assigned = np.zeros(nattr0, dtype=np.int32)
    if atexist:
        for i in range(nattr0):
            # identify_atom(x, icrd_crys, distmax=ratom): the nearest atom
            # (first on ties) if at most ratom away
            d = [_shortest_dist(x2c, xattr[i] - cpcel[a]) for a in range(nat_atoms)]
            nid = 0
            if d:
                a = int(np.argmin(d))
                if d[a] <= ratom:
                    nid = a + 1
            if nid == 0:
                # nearest_cp(x, nid, dist, type=typnuc): strict < keeps the first
                best, bd = 0, 1e30
                for j in range(cpcel.shape[0]):
                    if cptyp[j] != TYP_NUC:
                        continue
                    dj = _shortest_dist(x2c, cpcel[j] - xattr[i])
                    if dj < bd:
                        best, bd = j + 1, dj
                if bd < ratom:
                    nid = best
            assigned[i] = nid
    nattr, icp, xnew, new_cps = 0, [], [], []
    for i in range(1, cpcel.shape[0] + 1):
        if np.any(assigned == i):
            nattr += 1
            icp.append(i)
            xnew.append(cpcel[i - 1])
    for i in range(nattr0):
        if assigned[i] > 0:
            continue
        nattr += 1
        assigned[i] = nattr
        icp.append(0)
        xnew.append(xattr[i])
        for j in range(i + 1, nattr0):
            if assigned[j] > 0:
                continue
            if _shortest_dist(x2c, xattr[i] - xattr[j]) < ratom:
                assigned[j] = nattr
        new_cps.append(xattr[i])
    return {"nattr": nattr, "xattr": np.array(xnew).reshape(-1, 3), "icp": np.array(icp, dtype=np.int32),
            "assigned": assigned, "new_cps": np.array(new_cps).reshape(-1, 3)}


def bader(ctx: Context, rho, n, x2c, xat, nnm=False, ratom=None, noatoms=False, lap=True):
    """BADER keyword: labels + attractors + integrated Volume/Pop/Lap."""
    bas = bader_integrate(ctx, rho, n, x2c, xat, nnm=nnm, ratom=ratom, noatoms=noatoms)
    props = intgrid_fields(ctx, bas, rho, x2c, lap=lap)
    return bas, props


def synth_slater(ctx: Context, cell, rimg_ang=8.0, device_out=None):
    """FP64 Slater-sum grid of a synth.Cell, generated on the GPU."""
    n = np.asarray(cell.n, dtype=np.int32)
    npts = int(n.prod())
    x = _lib.mat_f(cell.x2c)
    xat = np.ascontiguousarray(cell.xat.reshape(-1), dtype=np.float64)
    zat = np.ascontiguousarray(cell.zat, dtype=np.int32)
    rimg = rimg_ang / 0.52917720859
    if device_out is not None:
        rc = ctx.lib.c2gpu_synth_slater(ctx.h, x.ctypes.data, n.ctypes.data, len(zat), xat.ctypes.data,
                                        zat.ctypes.data, rimg, _lib.PTR_DEVICE, _lib.ptr(device_out))
        ctx.check(rc, "c2gpu_synth_slater")
        return device_out
    out = np.zeros(npts)
    rc = ctx.lib.c2gpu_synth_slater(ctx.h, x.ctypes.data, n.ctypes.data, len(zat), xat.ctypes.data,
                                    zat.ctypes.data, rimg, 0, out.ctypes.data)
    ctx.check(rc, "c2gpu_synth_slater")
    return out


__all__ = ["BasinData", "bader_integrate", "basin_sums", "intgrid_fields", "bader", "fft_deriv",
           "synth_slater", "Context", "C2GpuError"]


def multipoles(ctx: Context, labels, f, n, x2c, xattr, lmax: int = 5):
    """Basin multipoles of the field f (INTEGRABLE f MULTIPOLE lmax,
    integration@proc.f90:1340-1362): array (nattr, (lmax+1)^2) of sums of
    the real solid harmonics (genrlm_real order) times f over each basin,
    unscaled (critic2 multiplies by Omega/N).  labels and f: host arrays or
    DeviceArrays; xattr fractional (nattr, 3)."""
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
    ctx.check(ctx.lib.c2gpu_multipoles(ctx.h, _lib.ptr(labels), _lib.ptr(f), nn.ctypes.data,
                                       _lib.mat_f(x2c).ctypes.data, nattr, xattr.ctypes.data, lmax,
                                       _lib.PTR_DEVICE if dev else 0, out.ctypes.data), "c2gpu_multipoles")
    return out.reshape(nattr, nlm)
