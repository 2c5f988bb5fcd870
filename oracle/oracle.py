"""Python access to the oracle (CPU restatement of critic2's path).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Never used by critic2_amd.

C functions (liboracle.so, built by oracle/Makefile) restate
bader@proc.f90, intgrid_fields, grid1%interp / read_critic,
promolecular_atom and the NCI point formulas; see the file headers of
bader.c and nci.c for the line-by-line references.  The FFT derivative grids
of grid3%fft (grid3mod@proc.f90:1757-1872) are restated here with numpy's
FFT: same transform conventions (forward exp(-i), scaled 1/N; backward
exp(+i), unscaled), same reciprocal vectors and Nyquist handling; only the
round-off of the FFT itself differs from fftpack5.

What pins each part (the checks live in tests/):
- FFT grids: single Fourier modes in a triclinic cell, even and odd grids,
  Nyquist-plane modes.
- Promolecular NCI and the cutoff pass: the reference's urea outputs
  (005_nciplot_basic cubes and .dat, 006_nciplot_fragment cubes).
- Fragments: the partition identity, table nodes and hand-worked cutoff
  cases.
- Tricubic interpolation: golden a = c.b vectors from the reference's table,
  and exact quadratics.
- Cube readers: the reference's cube files, and Python's float().
- YT: the partition of unity and the one-maximum limit.
- Facets: closed-form Voronoi cells.
- Multipoles: the closed-form harmonics, the sum rule and the addition
  theorem.
- Bader: see bader.c.  Its order-dependent labels are parity unpinned
  against the reference itself.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


DISCARD_FN = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_double), C.c_void_p)


class OrcGeom(C.Structure):
    _fields_ = [("x2c", C.c_double * 9), ("nat", C.c_int), ("xat", C.c_void_p), ("mode", C.c_int),
                ("ratom", C.c_double), ("discard", DISCARD_FN), ("discard_user", C.c_void_p),
                ("nat_atoms", C.c_int)]


class OrcInfo(C.Structure):
    _fields_ = [("nattr", C.c_int), ("nmax_found", C.c_int), ("refine_iters", C.c_int),
                ("nstarts", C.c_int64), ("nsteps", C.c_int64), ("status", C.c_int)]


class OrcRgrid(C.Structure):
    _fields_ = [("ngrid", C.c_int), ("a", C.c_double), ("b", C.c_double), ("rmax", C.c_double),
                ("r", C.c_void_p), ("f", C.c_void_p), ("fp", C.c_void_p), ("fpp", C.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        lib.oracle_bader.argtypes = [vp, vp, vp, vp, C.POINTER(OrcGeom), vp, vp, C.c_int, C.POINTER(OrcInfo)]
        lib.oracle_bader_sweep_only.argtypes = [vp, vp, vp, vp, C.POINTER(OrcGeom), vp, vp, vp, C.c_int,
                                                C.POINTER(OrcInfo)]
        lib.oracle_bader_metrics.argtypes = [vp, vp, vp, vp, vp]
        lib.oracle_basin_sums.argtypes = [vp, C.c_int64, C.c_int, C.c_int, vp, vp, vp]
        lib.oracle_grid1_build.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp]
        lib.oracle_grid1_interp.argtypes = [C.POINTER(OrcRgrid), C.c_double, vp, vp, vp]
        lib.oracle_promol.argtypes = [vp, C.c_int, vp, vp, C.POINTER(OrcRgrid), vp, C.c_int, vp, vp, vp]
        lib.oracle_frag_density.argtypes = [vp, C.c_int, vp, vp, vp, C.POINTER(OrcRgrid), vp]
        lib.oracle_nci_point.argtypes = [C.c_double, C.c_double, vp, C.c_int, vp, vp]
        lib.oracle_slater_grid.argtypes = [vp, vp, C.c_int, vp, vp, C.c_double, vp]
        lib.oracle_set_mutation.argtypes = [C.c_int]
        lib.oracle_refine_counts.argtypes = [C.c_void_p, C.c_int]
        _lib = lib
    return _lib


# Mutation switches of the Bader restatement (c2oracle.h): each replaces one
# rule of bader@proc.f90 by a plausible mistake.  Tests only.
MUTATIONS = {
    "dr_reset": 1,        # dr reset at every step (:459-465 resets it only when pnum == 1)
    "no_known2_stop": 2,  # max_neargrid ignores known == 2 (:447)
    "rint": 4,            # nint rounds half to even (Fortran nint: half away from zero)
    "no_pt_quirk": 8,     # refine pass >= 2 reads known(p), not known(pt) after the loop (:359)
    "no_refine": 16,      # no refine_edge passes (:218-224)
    "ongrid_ge": 32,      # step_ongrid takes the last of equal maxima (:518 is a strict >)
    "no_loop_check": 64,  # no on-grid fallback when the next point is on the path (:487-491)
    "keep_label": 128,    # a committed path keeps earlier labels (:203-206 overwrites them)
}


class mutation:
    """Context manager: run the Bader restatement with one mutation switched on."""

    def __init__(self, name):
        self.mask = MUTATIONS[name]

    def __enter__(self):
        load().oracle_set_mutation(self.mask)
        return self

    def __exit__(self, *exc):
        load().oracle_set_mutation(0)
        return False


def refine_counts():
    """Points re-ascended by each refine_edge pass of the last bader() run."""
    out = np.zeros(64, dtype=np.int64)
    n = load().oracle_refine_counts(out.ctypes.data, 64)
    return [int(v) for v in out[:min(n, 64)]]


def _matf(m):
    return np.asfortranarray(np.asarray(m, dtype=np.float64)).reshape(-1, order="F").copy()


def metrics(x2c, n):
    lib = load()
    x = _matf(x2c)
    nn = np.asarray(n, dtype=np.int32)
    l2c, c2l, lid = np.zeros(9), np.zeros(9), np.zeros(27)
    lib.oracle_bader_metrics(x.ctypes.data, nn.ctypes.data, l2c.ctypes.data, c2l.ctypes.data, lid.ctypes.data)
    return l2c, c2l, lid


def bader(f, n, x2c, xat, mode=0, ratom=1e40, sweep_only=False, car2lat=None, lat_i_dist=None, discard=None,
          cps=None):
    """Serial bader_integrate; returns (labels int32 x-fastest, info dict,
    xattr (nattr,3), known or None).  discard: None or a function of the
    Cartesian position (bohr) of a new maximum, true to discard it (BADER's
    DISCARD expression, bader@proc.f90:185-191).  cps: the field's other
    known critical points (fractional, after the atoms in cpcel order)."""
    lib = load()
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    nn = np.asarray(n, dtype=np.int32)
    if car2lat is None:
        _, car2lat, lat_i_dist = metrics(x2c, n)
    c2l = np.ascontiguousarray(car2lat, dtype=np.float64)
    lid = np.ascontiguousarray(lat_i_dist, dtype=np.float64)
    xa = np.ascontiguousarray(np.asarray(xat, dtype=np.float64).reshape(-1))
    g = OrcGeom()
    g.nat_atoms = len(xa) // 3
    if cps is not None and len(cps):
        xa = np.ascontiguousarray(np.concatenate([xa, np.asarray(cps, dtype=np.float64).reshape(-1)]))
    for q, v in enumerate(_matf(x2c)):
        g.x2c[q] = v
    g.nat = len(xa) // 3
    g.xat = xa.ctypes.data
    g.mode = mode
    g.ratom = ratom
    cb = None
    if discard is not None:
        cb = DISCARD_FN(lambda x, user: 1 if discard(np.array([x[0], x[1], x[2]])) else 0)
        g.discard = cb
    lab = np.zeros(f.size, dtype=np.int32)
    cap = 4 * g.nat + 256
    xattr = np.zeros(3 * cap)
    info = OrcInfo()
    known = None
    if sweep_only:
        known = np.zeros(f.size, dtype=np.int32)
        lib.oracle_bader_sweep_only(f.ctypes.data, nn.ctypes.data, c2l.ctypes.data, lid.ctypes.data,
                                    C.byref(g), lab.ctypes.data, known.ctypes.data, xattr.ctypes.data,
                                    cap, C.byref(info))
    else:
        lib.oracle_bader(f.ctypes.data, nn.ctypes.data, c2l.ctypes.data, lid.ctypes.data, C.byref(g),
                         lab.ctypes.data, xattr.ctypes.data, cap, C.byref(info))
    d = {k: getattr(info, k) for k, _ in info._fields_}
    return lab, d, xattr[: 3 * info.nattr].reshape(-1, 3), known


def basin_sums(labels, nattr, fields=()):
    lib = load()
    labels = np.ascontiguousarray(labels, dtype=np.int32).reshape(-1)
    fields = [np.ascontiguousarray(f, dtype=np.float64).reshape(-1) for f in fields]
    fp = (C.c_void_p * max(len(fields), 1))(*[f.ctypes.data for f in fields])
    counts = np.zeros(nattr, dtype=np.int64)
    sums = np.zeros(max(len(fields), 1) * nattr)
    lib.oracle_basin_sums(labels.ctypes.data, labels.size, nattr, len(fields), fp, counts.ctypes.data,
                          sums.ctypes.data)
    return counts, sums[: len(fields) * nattr].reshape(len(fields), nattr)


def grid1_build(r, rr0):
    lib = load()
    r = np.ascontiguousarray(r, dtype=np.float64)
    rr0 = np.ascontiguousarray(rr0, dtype=np.float64)
    f, fp, fpp = np.zeros_like(r), np.zeros_like(r), np.zeros_like(r)
    a, b = C.c_double(), C.c_double()
    lib.oracle_grid1_build(len(r), r.ctypes.data, rr0.ctypes.data, f.ctypes.data, fp.ctypes.data,
                           fpp.ctypes.data, C.byref(a), C.byref(b))
    return f, fp, fpp, a.value, b.value


class Table:
    """Oracle-side radial table (keeps the arrays alive)."""

    def __init__(self, r, f, fp, fpp, cutrad):
        self.r, self.f, self.fp, self.fpp = (np.ascontiguousarray(v, dtype=np.float64) for v in (r, f, fp, fpp))
        self.s = OrcRgrid()
        self.s.ngrid = len(self.r)
        self.s.a = float(self.r[0])
        self.s.b = math.log(self.r[1] / self.r[0])
        self.s.rmax = float(self.r[-1])
        self.s.r, self.s.f, self.s.fp, self.s.fpp = (v.ctypes.data for v in (self.r, self.f, self.fp, self.fpp))
        self.cutrad = cutrad


def promol_nci(points_cart, images, img_species, tables, full_hessian=True):
    """rho, |grad|, s, crho at Cartesian points; images: (m,3) Cartesian
    positions of every atom image to include; img_species: table index."""
    lib = load()
    tabs = (OrcRgrid * len(tables))(*[t.s for t in tables])
    cut = np.array([min(t.cutrad, t.s.rmax) for t in tables], dtype=np.float64)
    ximg = np.ascontiguousarray(np.asarray(images, dtype=np.float64).reshape(-1))
    zi = np.ascontiguousarray(np.asarray(img_species, dtype=np.int32))
    pts = np.asarray(points_cart, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((len(pts), 4))
    g = np.zeros(3)
    h = np.zeros(9)
    f = C.c_double()
    cr, s = C.c_double(), C.c_double()
    for q, x in enumerate(pts):
        xc = np.ascontiguousarray(x)
        lib.oracle_promol(xc.ctypes.data, len(zi), ximg.ctypes.data, zi.ctypes.data, tabs, cut.ctypes.data, 2,
                          C.byref(f), g.ctypes.data, h.ctypes.data)
        # nucleus: gradient nullified within 1e-5 bohr (fieldmod@proc.f90:1146-1155)
        d = np.sqrt(((np.asarray(images).reshape(-1, 3) - x) ** 2).sum(1)).min() if len(zi) else 1.0
        gm = 0.0 if d <= 1e-5 else float(np.sqrt((g * g).sum()))
        lib.oracle_nci_point(f.value, gm, h.ctypes.data, 1 if full_hessian else 0, C.byref(cr), C.byref(s))
        out[q] = (f.value, gm, s.value, cr.value)
    return out


def promol_nci_batch(points_cart, images, img_species, tables, threads=None):
    """promol_nci over many nodes in C, OpenMP-parallel with schedule(dynamic)
    as nciplot's loop (nci@proc.f90:543); threads: OMP threads (None = the
    OpenMP default).  Returns (npts, 4): rho, |grad|, s, crho."""
    lib = load()
    lib.oracle_promol_nci_batch.argtypes = [C.c_int64, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                            C.POINTER(OrcRgrid), C.c_void_p, C.c_int, C.c_void_p]
    tabs = (OrcRgrid * len(tables))(*[t.s for t in tables])
    cut = np.array([min(t.cutrad, t.s.rmax) for t in tables], dtype=np.float64)
    ximg = np.ascontiguousarray(np.asarray(images, dtype=np.float64).reshape(-1))
    zi = np.ascontiguousarray(np.asarray(img_species, dtype=np.int32))
    pts = np.ascontiguousarray(np.asarray(points_cart, dtype=np.float64).reshape(-1, 3))
    out = np.zeros((len(pts), 4))
    if threads is not None:
        omp = C.CDLL("libgomp.so.1")
        omp.omp_set_num_threads(int(threads))
    lib.oracle_promol_nci_batch(len(pts), pts.ctypes.data, len(zi), ximg.ctypes.data, zi.ctypes.data, tabs,
                                cut.ctypes.data, 1, out.ctypes.data)
    return out


def frag_density(points_cart, frags, tables):
    """Fragment densities (nci@proc.f90:580-596) at Cartesian points; frags:
    list of (xyz (m, 3) Cartesian, table index (m,)).  Returns (nfrag, npts)."""
    lib = load()
    tabs = (OrcRgrid * len(tables))(*[t.s for t in tables])
    nat = [len(np.asarray(f[1]).reshape(-1)) for f in frags]
    fstart = np.ascontiguousarray(np.concatenate([[0], np.cumsum(nat)]).astype(np.int32))
    xyz = np.ascontiguousarray(np.concatenate([np.asarray(f[0], dtype=np.float64).reshape(-1)
                                               for f in frags] + [np.zeros(3)]))
    zi = np.ascontiguousarray(np.concatenate([np.asarray(f[1], dtype=np.int32).reshape(-1)
                                              for f in frags] + [np.zeros(1, dtype=np.int32)]))
    pts = np.asarray(points_cart, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((len(pts), len(frags)))
    row = np.zeros(len(frags))
    for q, x in enumerate(pts):
        xc = np.ascontiguousarray(x)
        lib.oracle_frag_density(xc.ctypes.data, len(frags), fstart.ctypes.data, xyz.ctypes.data, zi.ctypes.data,
                                tabs, row.ctypes.data)
        out[q] = row
    return out.T.copy()


def box_points(nstep, x0, xmat):
    """Cartesian nodes x0 + i xmat(:,1) + j xmat(:,2) + k xmat(:,3) in
    nciplot's z-fastest order (nci@proc.f90:545-548); xmat columns are the
    steps."""
    xm = np.asarray(xmat, dtype=np.float64).reshape(3, 3)
    out = []
    for i in range(nstep[0]):
        for j in range(nstep[1]):
            for k in range(nstep[2]):
                out.append(np.asarray(x0, dtype=np.float64) + i * xm[:, 0] + j * xm[:, 1] + k * xm[:, 2])
    return np.array(out)


def nci_point(rho, gmod, hdiag):
    lib = load()
    e = np.zeros(9)
    e[:3] = hdiag
    cr, s = C.c_double(), C.c_double()
    lib.oracle_nci_point(float(rho), float(gmod), e.ctypes.data, 0, C.byref(cr), C.byref(s))
    return cr.value, s.value


def slater_grid(x2c, n, xat, zat, rimg_bohr):
    lib = load()
    out = np.zeros(int(np.prod(n)))
    x = _matf(x2c)
    nn = np.asarray(n, dtype=np.int32)
    xa = np.ascontiguousarray(np.asarray(xat, dtype=np.float64).reshape(-1))
    za = np.ascontiguousarray(np.asarray(zat, dtype=np.int32))
    lib.oracle_slater_grid(x.ctypes.data, nn.ctypes.data, len(za), xa.ctypes.data, za.ctypes.data,
                           float(rimg_bohr), out.ctypes.data)
    return out


# ---- grid3%fft restated with numpy (grid3mod@proc.f90:1757-1872)

def _vgc(x2c, n):
    x2c = np.asarray(x2c, dtype=np.float64)
    b = np.zeros((3, 3))
    b[:, 0] = np.cross(x2c[:, 2], x2c[:, 1])
    b[:, 1] = np.cross(x2c[:, 0], x2c[:, 2])
    b[:, 2] = np.cross(x2c[:, 1], x2c[:, 0])
    b = 2.0 * math.pi / abs(np.linalg.det(x2c)) * b

    def idx(m):
        i = np.arange(m)
        return np.where(i <= m // 2, i, i - m).astype(np.float64)

    i1, i2, i3 = np.meshgrid(idx(n[0]), idx(n[1]), idx(n[2]), indexing="ij")
    return [i1 * b[c, 0] + i2 * b[c, 1] + i3 * b[c, 2] for c in range(3)]


def fft_deriv(f, n, x2c, kind):
    """kind: 'grad', 'xx', 'yy', 'zz', 'lap'.  f is x-fastest flat."""
    g = np.asarray(f, dtype=np.float64).reshape(n[2], n[1], n[0]).transpose(2, 1, 0)  # (n1,n2,n3)
    N = g.size
    z = np.fft.fftn(g) / N
    v = _vgc(x2c, n)
    if kind == "grad":
        acc = np.zeros(g.shape)
        for c in range(3):
            w = v[c] * (1j * z)
            acc += np.real(np.fft.ifftn(w) * N) ** 2
        out = np.sqrt(acc)
    elif kind in ("xx", "yy", "zz"):
        c = "xyz".index(kind[0])
        out = np.real(np.fft.ifftn(-v[c] * v[c] * z) * N)
    elif kind == "lap":
        out = np.real(np.fft.ifftn(-(v[0] ** 2 + v[1] ** 2 + v[2] ** 2) * z) * N)
    else:
        raise ValueError(kind)
    return out.transpose(2, 1, 0).reshape(-1)


def nci_grid_fft(f, n, x2c):
    """NCI node quantities of the fourierint branch (nci@proc.f90:546-564),
    x-fastest flat arrays (crho, s)."""
    f = np.asarray(f, dtype=np.float64).reshape(-1)
    gm = fft_deriv(f, n, x2c, "grad")
    hs = [fft_deriv(f, n, x2c, k) for k in ("xx", "yy", "zz")]
    npos = sum((h > 0).astype(int) for h in hs)
    cst = 2.0 * (3.0 * math.pi ** 2) ** (1.0 / 3.0)
    s = gm / (cst * np.maximum(f, 1e-80) ** (4.0 / 3.0))
    crho = np.where(npos >= 2, np.abs(f), -np.abs(f)) * 100.0
    return crho, s


def step_pointers(f, n, x2c):
    lib = load()
    lib.oracle_step_pointers.argtypes = [C.c_void_p] * 7
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    _, c2l, lid = metrics(x2c, n)
    nn = np.asarray(n, dtype=np.int32)
    on = np.zeros(f.size, dtype=np.int64)
    ne = np.zeros(f.size, dtype=np.int64)
    dr = np.zeros(3 * f.size)
    lib.oracle_step_pointers(f.ctypes.data, nn.ctypes.data, c2l.ctypes.data, lid.ctypes.data,
                             on.ctypes.data, ne.ctypes.data, dr.ctypes.data)
    return on, ne, dr


# ---- NCIPLOT grid-interpolation scheme (nci@proc.f90:565-571)

def matinv3(m):
    """critic2's matinv of a 3x3 matrix (LAPACK operation order)."""
    lib = load()
    lib.oracle_matinv3.argtypes = [C.c_void_p]
    a = _matf(m)
    lib.oracle_matinv3(a.ctypes.data)
    return a.reshape(3, 3, order="F")


def tricubic_matrix():
    """The 64x64 tricubic coefficient matrix c(l, m) (a = c b)."""
    lib = load()
    lib.oracle_tricubic_matrix.argtypes = [C.c_void_p]
    c = np.zeros(64 * 64)
    lib.oracle_tricubic_matrix(c.ctypes.data)
    return c.reshape(64, 64, order="F")


def tricubic(f, n, xi):
    """grinterp_tricubic at fractional xi: (y, yp[3], ypp[3,3]) with the
    fractional-coordinate derivatives scaled by n."""
    lib = load()
    lib.oracle_tricubic.argtypes = [C.c_void_p] * 6
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    nn = np.asarray(n, dtype=np.int32)
    x = np.asarray(xi, dtype=np.float64).copy()
    y = np.zeros(1)
    yp = np.zeros(3)
    ypp = np.zeros(9)
    lib.oracle_tricubic(f.ctypes.data, nn.ctypes.data, x.ctypes.data, y.ctypes.data, yp.ctypes.data,
                        ypp.ctypes.data)
    return y[0], yp, ypp.reshape(3, 3, order="F")


def nci_grid_interp(f, n, x2c):
    """NCI node quantities of the grid-interpolation branch, z-fastest flat
    arrays (crho, s) as nciplot stores them."""
    lib = load()
    lib.oracle_nci_grid.argtypes = [C.c_void_p] * 6
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    nn = np.asarray(n, dtype=np.int32)
    x = _matf(x2c)
    c2x = _matf(matinv3(x2c))
    crho = np.zeros(f.size)
    s = np.zeros(f.size)
    lib.oracle_nci_grid(f.ctypes.data, nn.ctypes.data, x.ctypes.data, c2x.ctypes.data, crho.ctypes.data,
                        s.ctypes.data)
    return crho, s


def _cube_call(fn_name, path):
    lib = load()
    fn = getattr(lib, fn_name)
    fn.argtypes = [C.c_char_p] + [C.c_void_p] * 6
    fn.restype = C.c_int
    nat, ismo = C.c_int(), C.c_int()
    n = np.zeros(3, dtype=np.int32)
    x0 = np.zeros(3)
    step = np.zeros(9)
    p = os.fsencode(path)
    if fn(p, C.byref(nat), C.byref(ismo), n.ctypes.data, x0.ctypes.data, step.ctypes.data, None):
        raise ValueError(f"{fn_name}: error reading {path}")
    f = np.empty(int(np.prod(n)), dtype=np.float64)
    if fn(p, C.byref(nat), C.byref(ismo), n.ctypes.data, x0.ctypes.data, step.ctypes.data, f.ctypes.data):
        raise ValueError(f"{fn_name}: error reading {path}")
    hdr = {"nat": nat.value, "ismo": bool(ismo.value), "n": tuple(int(v) for v in n), "x0_cart": x0,
           "step": step.reshape(3, 3).T}
    return hdr, f


def read_cube(path):
    """grid3%read_cube restated (strtod conversion): (header, f x-fastest)."""
    return _cube_call("oracle_read_cube", path)


def read_bincube(path):
    """grid3%read_bincube restated: (header, f x-fastest)."""
    return _cube_call("oracle_read_bincube", path)


def nci_cutoffs(crho, cgrad, rhocut=0.2, dimcut=2.0, rhoplot=0.1, srhorange=(-1e30, 1e30), onlyneg=False,
                rhoat=None, rho_void=-1.0, rhofrag=None, rhoparam=0.95, rhoparam2=0.75):
    """Cutoff pass of nci@proc.f90:612-643, restated with numpy: the (rho, s)
    rows of the .dat file in node order, cgrad with 100 outside the plot
    range, and the unscaled void sums.  rhofrag (nfrag, npts): the fragment
    test of :622-626, its sum taken fragment by fragment."""
    crho = np.asarray(crho, dtype=np.float64).reshape(-1)
    cg = np.array(cgrad, dtype=np.float64, copy=True).reshape(-1)
    inter = np.ones(crho.shape, dtype=bool)
    sums = np.zeros(3)
    if rhofrag is not None and len(rhofrag):
        rf = np.asarray(rhofrag, dtype=np.float64)
        ra = np.asarray(rhoat, dtype=np.float64).reshape(-1)
        tot = np.zeros(crho.shape)
        for f in range(rf.shape[0]):
            tot = tot + rf[f]
        inter = (tot >= rhoparam2 * ra) & np.all(rf <= tot * rhoparam, axis=0)
    if rho_void > 0:
        ra = np.asarray(rhoat, dtype=np.float64).reshape(-1)
        void = ra < rho_void
        inter = inter & void
        sums = np.array([np.sum(np.abs(crho[void]) / 100.0), np.sum(ra[void]), float(np.count_nonzero(void))])
    sel = (np.abs(crho) < rhocut * 100.0) & (cg < dimcut) & inter
    dat = np.stack([crho[sel] / 100.0, cg[sel]], axis=1)
    cg[(np.abs(crho) > rhoplot * 100.0) | ~inter] = 100.0
    cg[(crho < srhorange[0] * 100.0) | (crho > srhorange[1] * 100.0)] = 100.0
    if onlyneg:
        cg[crho > 0] = 100.0
    return dat, cg, sums


def yt(f, n, x2c, xat, vec, area, fields=(), mode=0, ratom=1e40, weights=False, cps=None):
    """yt_integrate + yt_weights + the imtype_yt sums (oracle_yt).  Returns
    (labels in grid order, nattr, xattr, vol, sums[q, a]), and with
    weights=True also w[a, p], every basin's yt_weights in grid order.
    Equal values are ranked in the reference quicksort's order
    (qcksort_order)."""
    lib = load()
    lib.oracle_yt.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(OrcGeom), C.c_int, C.c_void_p, C.c_void_p,
                              C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.c_int]
    lib.oracle_yt.restype = C.c_int
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    nn = np.asarray(n, dtype=np.int32)
    xat = np.ascontiguousarray(xat, dtype=np.float64).reshape(-1)
    g = OrcGeom()
    g.nat_atoms = xat.size // 3
    if cps is not None and len(cps):
        xat = np.ascontiguousarray(np.concatenate([xat, np.asarray(cps, dtype=np.float64).reshape(-1)]))
    for q, v in enumerate(_matf(x2c)):
        g.x2c[q] = v
    g.nat = xat.size // 3
    g.xat = xat.ctypes.data
    g.mode = mode
    g.ratom = ratom
    vec = np.ascontiguousarray(vec, dtype=np.int32).reshape(-1)
    area = np.ascontiguousarray(area, dtype=np.float64)
    fl = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in fields]
    fp = (C.c_void_p * max(len(fl), 1))(*[x.ctypes.data for x in fl])
    idg = np.zeros(f.size, dtype=np.int32)
    cap = 4 * g.nat + 4096
    xattr = np.zeros(3 * cap)
    vol = np.zeros(cap)
    sums = np.zeros(max(len(fl), 1) * cap)
    na = C.c_int(0)
    wcap = g.nat + 64 if weights else 0
    w = np.zeros(wcap * f.size) if weights else None
    rc = lib.oracle_yt(f.ctypes.data, nn.ctypes.data, C.byref(g), len(area), vec.ctypes.data, area.ctypes.data,
                       len(fl), fp, idg.ctypes.data, C.byref(na), xattr.ctypes.data, cap, vol.ctypes.data,
                       sums.ctypes.data, w.ctypes.data if weights else None, wcap)
    if rc:
        raise ValueError(f"oracle_yt: status {rc}")
    k = na.value
    out = (idg, k, xattr[:3 * k].reshape(-1, 3), vol[:k], sums[:len(fl) * k].reshape(len(fl), k))
    if weights:
        if k > wcap:
            raise ValueError("oracle_yt: more basins than the weight array holds")
        out = out + (w[:k * f.size].reshape(k, f.size),)
    return out


def yt_multipoles(w, f, n, x2c, xattr, lmax=5):
    """YT multipoles from yt(..., weights=True)'s per-basin weights."""
    lib = load()
    lib.oracle_yt_multipoles.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                         C.c_int, C.c_void_p]
    w = np.ascontiguousarray(w, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    nn = np.asarray(n, dtype=np.int32)
    x = _matf(x2c)
    xa = np.ascontiguousarray(np.asarray(xattr, dtype=np.float64).reshape(-1))
    na = xa.size // 3
    nlm = (lmax + 1) ** 2
    out = np.zeros(na * nlm)
    lib.oracle_yt_multipoles(w.ctypes.data, f.ctypes.data, nn.ctypes.data, x.ctypes.data, na, xa.ctypes.data,
                             lmax, out.ctypes.data)
    return out.reshape(na, nlm)


def rrlm(v, lmax):
    """genrlm_real at the Cartesian vector v through tosphere (the
    reference's route through spherical angles)."""
    lib = load()
    lib.oracle_tosphere.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.oracle_genrlm_real.argtypes = [C.c_int, C.c_double, C.c_void_p, C.c_void_p]
    v = np.ascontiguousarray(v, dtype=np.float64)
    r = np.zeros(1)
    tp = np.zeros(2)
    lib.oracle_tosphere(v.ctypes.data, r.ctypes.data, tp.ctypes.data)
    out = np.zeros((lmax + 1) ** 2)
    lib.oracle_genrlm_real(lmax, float(r[0]), tp.ctypes.data, out.ctypes.data)
    return out


def multipoles(idg, f, n, x2c, xattr, lmax=5):
    """Basin multipoles (oracle_multipoles): (nattr, (lmax+1)^2), unscaled."""
    lib = load()
    lib.oracle_multipoles.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                      C.c_void_p]
    idg = np.ascontiguousarray(idg, dtype=np.int32).reshape(-1)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    nn = np.asarray(n, dtype=np.int32)
    xattr = np.ascontiguousarray(xattr, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(xattr.shape[0] * (lmax + 1) ** 2)
    lib.oracle_multipoles(idg.ctypes.data, f.ctypes.data, nn.ctypes.data, _matf(x2c).ctypes.data, xattr.shape[0],
                          xattr.ctypes.data, lmax, out.ctypes.data)
    return out.reshape(xattr.shape[0], -1)


# ---------------------------------------------------------------- attractor reordering

def _shortest(x2c, xd):
    """crystal%shortest length of a fractional difference (search over
    translations -2..2 after the nint fold)."""
    x0 = np.asarray(xd, dtype=np.float64) - np.round(xd)
    best = np.inf
    for a in range(-2, 3):
        for b in range(-2, 3):
            for c in range(-2, 3):
                v = np.asarray(x2c, dtype=np.float64) @ (x0 + np.array([a, b, c], dtype=np.float64))
                best = min(best, float(np.sqrt(v @ v)))
    return best


def reorder_gridout(xattr, x2c, cpcel, cptyp, nat_atoms, ratom, atexist=True):
    """int_reorder_gridout (integration@proc.f90:1027-1122), Bader/YT branch,
    line by line: returns (assigned, nattr, icp, xattr_new).  assigned[i] is
    the new label of attractor i+1; an attractor matched to critical point k
    keeps label k (the reference does not compact those ids)."""
    xattr = np.asarray(xattr, dtype=np.float64).reshape(-1, 3)
    cpcel = np.asarray(cpcel, dtype=np.float64).reshape(-1, 3)
    nattr0 = len(xattr)
    assigned = [0] * nattr0
    if atexist:                                              # :1069-1082
        for i in range(nattr0):
            nid, bd = 0, np.inf
            for a in range(nat_atoms):                       # identify_atom, distmax = ratom
                d = _shortest(x2c, xattr[i] - cpcel[a])
                if d < bd:
                    nid, bd = a + 1, d
            if nid and bd <= ratom:
                assigned[i] = nid
            else:
                nid, dist = 0, 1e30                          # nearest_cp(..., type=typnuc)
                for j in range(len(cpcel)):
                    if cptyp[j] != -3:
                        continue
                    d = _shortest(x2c, cpcel[j] - xattr[i])
                    if d < dist:
                        nid, dist = j + 1, d
                if dist < ratom:
                    assigned[i] = nid
    nattr, icp, xnew = 0, [], []
    for i in range(1, len(cpcel) + 1):                       # :1084-1091
        if i in assigned:
            nattr += 1
            icp.append(i)
            xnew.append(cpcel[i - 1])
    for i in range(nattr0):                                  # :1094-1110
        if assigned[i] > 0:
            continue
        nattr += 1
        assigned[i] = nattr
        icp.append(0)
        xnew.append(xattr[i])
        for j in range(i + 1, nattr0):
            if assigned[j] > 0:
                continue
            if _shortest(x2c, xattr[i] - xattr[j]) < ratom:
                assigned[j] = nattr
    return np.array(assigned, dtype=np.int32), nattr, np.array(icp, dtype=np.int32), np.array(xnew).reshape(-1, 3)


def relabel(idg, assigned):
    """The `where` loop of int_reorder_gridout (:1113-1122): every i with
    assigned(i) /= i maps the labels of the ORIGINAL idg."""
    idg = np.asarray(idg)
    out = idg.copy()
    for i, a in enumerate(assigned, start=1):
        if a != i:
            out[idg == i] = a
    return out


def qcksort_order(arr):
    """qcksort_r8 over the whole array (tools@proc.f90:83-168): the
    permutation (0-based element indices) that sorts arr ascending, equal
    keys in that quicksort's arrangement."""
    lib = load()
    lib.oracle_qcksort_r8.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64]
    arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
    io = np.arange(arr.size, dtype=np.int64)
    if arr.size:
        lib.oracle_qcksort_r8(arr.ctypes.data, io.ctypes.data, 0, arr.size - 1)
    return io
