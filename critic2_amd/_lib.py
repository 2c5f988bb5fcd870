"""ctypes bindings of libc2gpu.so (the C ABI declared in include/c2gpu.h).

The library is the product: there is no CPU implementation behind these
functions.  If libc2gpu.so is missing, or no HIP device is present, calls
raise instead of computing anything.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libc2gpu.so")

OK = 0
ERR_NAMES = {
    1: "C2GPU_ERR_ARG", 2: "C2GPU_ERR_HIP", 3: "C2GPU_ERR_ALLOC", 4: "C2GPU_ERR_UNSUPPORTED",
    5: "C2GPU_ERR_NOCONVERGE", 6: "C2GPU_ERR_NEWMAX", 7: "C2GPU_ERR_PATHLEN",
    8: "C2GPU_ERR_DEGENERATE", 9: "C2GPU_ERR_FORMAT",
}
PTR_DEVICE = 1
ATTR_NEAREST_ATOM, ATTR_RATOM, ATTR_NOATOMS = 0, 1, 2
FT_GRAD, FT_XX, FT_YY, FT_ZZ, FT_LAP = 0, 1, 2, 3, 4
NCI_FFT, NCI_GRID, NCI_PROMOL = 0, 1, 2
CUBE_E22_14, CUBE_1P_E12_5, CUBE_1P_E13_5 = 0, 1, 2
FP64, FP32 = 0, 1

EXPORTS = (
    "c2gpu_ctx_create", "c2gpu_ctx_destroy", "c2gpu_ctx_error", "c2gpu_ctx_set_stream",
    "c2gpu_ctx_sync", "c2gpu_dev_alloc", "c2gpu_dev_free", "c2gpu_memcpy", "c2gpu_ctx_timing",
    "c2gpu_ctx_kernel_time", "c2gpu_ctx_kernel_names", "c2gpu_ctx_kernel_bytes",
    "c2gpu_bader_metrics", "c2gpu_bader_step_pointers", "c2gpu_bader", "c2gpu_relabel",
    "c2gpu_basin_sums", "c2gpu_basin_sums_noalloc", "c2gpu_fft_deriv", "c2gpu_nci", "c2gpu_nci_f32", "c2gpu_batch_run",
    "c2gpu_synth_slater", "c2gpu_cube_read_header", "c2gpu_cube_read", "c2gpu_cube_parse",
    "c2gpu_bincube_read", "c2gpu_cube_format", "c2gpu_cube_write", "c2gpu_nci_cutoff", "c2gpu_ws_facets",
    "c2gpu_yt", "c2gpu_multipoles", "c2gpu_yt_multipoles", "c2gpu_yt_remap", "c2gpu_nci_promol_nodes", "c2gpu_nci_fragments", "c2gpu_nci_cutoff_frag",
)


class C2GpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


# c2gpu_discard_fn: int (*)(const double x[3], void *user)
DISCARD_FN = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_double), C.c_void_p)


class AttractorOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("ratom", C.c_double), ("x2c", C.c_double * 9),
                ("nat", C.c_int), ("xat", C.c_void_p), ("discard", DISCARD_FN), ("discard_user", C.c_void_p),
                ("nat_atoms", C.c_int)]


class BaderStats(C.Structure):
    _fields_ = [("nattr", C.c_int), ("nmaxima", C.c_int), ("sweep_iters", C.c_int),
                ("refine_iters", C.c_int), ("nstarts", C.c_int64), ("traj_points", C.c_int64),
                ("edge_points", C.c_int64), ("events_total", C.c_int64),
                ("ms_traj", C.c_double), ("ms_sweep", C.c_double),
                ("ms_refine", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CubeHeader(C.Structure):
    _fields_ = [("nat", C.c_int), ("ismo", C.c_int), ("n", C.c_int * 3), ("x0", C.c_double * 3),
                ("step", C.c_double * 9), ("data_offset", C.c_uint64)]


class YtStats(C.Structure):
    _fields_ = [("nattr", C.c_int), ("nmaxima", C.c_int), ("nias", C.c_int64), ("label_passes", C.c_int),
                ("weight_passes", C.c_int), ("wide", C.c_int64), ("ms_total", C.c_double), ("ranked", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class NciCut(C.Structure):
    _fields_ = [("rhocut", C.c_double), ("dimcut", C.c_double), ("rhoplot", C.c_double),
                ("srhorange", C.c_double * 2), ("onlyneg", C.c_int), ("rho_void", C.c_double)]


class NciNodes(C.Structure):
    _fields_ = [("nstep", C.c_int * 3), ("x0", C.c_double * 3), ("xmat", C.c_double * 9)]


class RadialTable(C.Structure):
    _fields_ = [("ngrid", C.c_int), ("a", C.c_double), ("b", C.c_double), ("rmax", C.c_double),
                ("cutrad", C.c_double), ("r", C.c_void_p), ("f", C.c_void_p), ("fp", C.c_void_p),
                ("fpp", C.c_void_p)]


class Promol(C.Structure):
    _fields_ = [("nspecies", C.c_int), ("tables", C.POINTER(RadialTable)), ("nat", C.c_int),
                ("xat", C.c_void_p), ("species", C.c_void_p)]


_lib = None


def load():
    """Load libc2gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python critic2_amd/build.py`")
    lib = C.CDLL(LIB_PATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    sig = {
        "c2gpu_ctx_create": [C.POINTER(vp), i],
        "c2gpu_ctx_destroy": [vp],
        "c2gpu_ctx_set_stream": [vp, vp],
        "c2gpu_ctx_sync": [vp],
        "c2gpu_dev_alloc": [vp, C.c_uint64, C.POINTER(vp)],
        "c2gpu_dev_free": [vp, vp],
        "c2gpu_memcpy": [vp, vp, vp, C.c_uint64, i],
        "c2gpu_ctx_timing": [vp, i],
        "c2gpu_ctx_kernel_time": [vp, C.c_char_p, C.POINTER(d), C.POINTER(C.c_int64)],
        "c2gpu_ctx_kernel_names": [vp, C.c_char_p, C.c_uint64],
        "c2gpu_ctx_kernel_bytes": [vp, C.c_char_p, C.POINTER(d)],
        "c2gpu_bader_metrics": [vp, vp, vp, vp, vp],
        "c2gpu_bader_step_pointers": [vp, vp, vp, vp, vp, i, vp, vp, vp],
        "c2gpu_bader": [vp, vp, vp, vp, vp, C.POINTER(AttractorOpts), i, vp, C.POINTER(i), vp, i,
                        C.POINTER(BaderStats)],
        "c2gpu_relabel": [vp, C.c_int64, vp, i, i, vp],
        "c2gpu_basin_sums": [vp, vp, C.c_int64, i, i, vp, i, vp, vp],
        "c2gpu_fft_deriv": [vp, vp, vp, vp, i, i, vp],
        "c2gpu_nci": [vp, vp, vp, vp, i, i, C.POINTER(Promol), i, i, vp, vp],
        "c2gpu_nci_f32": [vp, vp, vp, vp, i, i, C.POINTER(Promol), i, i, vp, vp],
        "c2gpu_synth_slater": [vp, vp, vp, i, vp, vp, d, i, vp],
        "c2gpu_cube_read_header": [vp, C.c_char_p, i, C.POINTER(CubeHeader)],
        "c2gpu_cube_read": [vp, C.c_char_p, i, C.POINTER(CubeHeader), vp],
        "c2gpu_cube_parse": [vp, vp, C.c_uint64, i, C.POINTER(CubeHeader), vp],
        "c2gpu_bincube_read": [vp, C.c_char_p, i, C.POINTER(CubeHeader), vp],
        "c2gpu_cube_format": [vp, vp, vp, i, i, i, vp, C.c_uint64, C.POINTER(C.c_uint64)],
        "c2gpu_cube_write": [vp, C.c_char_p, C.POINTER(CubeHeader), vp, vp, vp, i, i],
        "c2gpu_nci_cutoff": [vp, vp, vp, vp, vp, i, C.POINTER(NciCut), i, vp, C.c_uint64, C.POINTER(C.c_uint64),
                             vp],
        "c2gpu_nci_cutoff_frag": [vp, vp, vp, vp, vp, i, d, d, vp, i, C.POINTER(NciCut), i, vp, C.c_uint64,
                                  C.POINTER(C.c_uint64), vp],
        "c2gpu_nci_promol_nodes": [vp, C.POINTER(NciNodes), vp, C.POINTER(Promol), i, i, vp, vp, vp],
        "c2gpu_nci_fragments": [vp, C.POINTER(NciNodes), i, C.POINTER(RadialTable), i, vp, vp, vp, i, i, vp],
        "c2gpu_ws_facets": [vp, C.POINTER(i), vp, vp],
        "c2gpu_multipoles": [vp, vp, vp, vp, vp, i, vp, i, i, vp],
        "c2gpu_yt_multipoles": [vp, vp, vp, vp, vp, i, vp, i, i, vp],
        "c2gpu_yt_remap": [vp, vp, i, i],
        "c2gpu_yt": [vp, vp, vp, vp, C.POINTER(AttractorOpts), i, vp, i, vp, C.POINTER(i), vp, i, vp, vp,
                     C.POINTER(YtStats)],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.c2gpu_ctx_error.argtypes = [vp]
    lib.c2gpu_ctx_error.restype = C.c_char_p
    _lib = lib
    return lib


def ptr(a) -> int:
    """Address of a numpy array or a torch tensor (device or host)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mat_f(m):
    """3x3 numpy matrix -> 9 doubles column-major (Fortran order)."""
    return np.asfortranarray(np.asarray(m, dtype=np.float64)).reshape(-1, order="F").copy()


class Context:
    """A c2gpu_ctx on one HIP device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.c2gpu_ctx_create(C.byref(h), device)
        if rc != OK:
            raise C2GpuError(rc, f"c2gpu_ctx_create(device={device}) failed (no HIP device?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.c2gpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc, what):
        if rc != OK:
            msg = self.lib.c2gpu_ctx_error(self.h).decode(errors="replace")
            raise C2GpuError(rc, f"{what}: {msg}")

    def set_stream(self, stream_handle):
        self.check(self.lib.c2gpu_ctx_set_stream(self.h, stream_handle), "c2gpu_ctx_set_stream")

    def sync(self):
        self.check(self.lib.c2gpu_ctx_sync(self.h), "c2gpu_ctx_sync")

    def timing(self, enable=True):
        self.check(self.lib.c2gpu_ctx_timing(self.h, 1 if enable else 0), "c2gpu_ctx_timing")

    def kernel_times(self):
        """{kernel name: (ms, launches)} accumulated since timing(True)."""
        return {k: v[:2] for k, v in self.kernel_stats().items()}

    def kernel_stats(self):
        """{kernel name: (ms, launches, algorithmic bytes or 0)} accumulated
        since timing(True); read before timing() is called again, which
        resets the table."""
        buf = C.create_string_buffer(1 << 16)
        self.check(self.lib.c2gpu_ctx_kernel_names(self.h, buf, len(buf)), "c2gpu_ctx_kernel_names")
        out = {}
        for name in filter(None, buf.value.decode().split(",")):
            ms, n, b = C.c_double(), C.c_int64(), C.c_double()
            self.check(self.lib.c2gpu_ctx_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)),
                       "c2gpu_ctx_kernel_time")
            self.check(self.lib.c2gpu_ctx_kernel_bytes(self.h, name.encode(), C.byref(b)),
                       "c2gpu_ctx_kernel_bytes")
            out[name] = (ms.value, n.value, b.value)
        return out

    def empty(self, shape, dtype):
        return DeviceArray(self, shape, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.copy_from(a)
        return d


class DeviceArray:
    """A device allocation owned by a Context (grids kept resident in HBM)."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx.check(ctx.lib.c2gpu_dev_alloc(ctx.h, self.nbytes, C.byref(p)), "c2gpu_dev_alloc")
        self.ptr = p.value

    def data_ptr(self):
        return self.ptr

    def numel(self):
        return int(np.prod(self.shape))

    @property
    def size(self):
        return self.numel()

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        self.ctx.check(self.ctx.lib.c2gpu_memcpy(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes, 1),
                       "c2gpu_memcpy")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        self.ctx.check(self.ctx.lib.c2gpu_memcpy(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes, 2),
                       "c2gpu_memcpy")
        return out

    def free(self):
        if self.ptr:
            self.ctx.lib.c2gpu_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


def ws_facets(m):
    """Voronoi-relevant lattice vectors (lattice coordinates) and facet areas
    of the lattice with vectors m[:, i] (host code, no GPU needed)."""
    lib = load()
    mm = mat_f(m)
    nv = C.c_int()
    vec = np.zeros(42, dtype=np.int32)
    area = np.zeros(14)
    rc = lib.c2gpu_ws_facets(mm.ctypes.data, C.byref(nv), vec.ctypes.data, area.ctypes.data)
    if rc != OK:
        raise C2GpuError(rc, "c2gpu_ws_facets failed")
    return vec[:3 * nv.value].reshape(-1, 3).copy(), area[:nv.value].copy()


def bader_metrics(x2c, n):
    lib = load()
    x = mat_f(x2c)
    nn = np.asarray(n, dtype=np.int32)
    l2c, c2l, lid = np.zeros(9), np.zeros(9), np.zeros(27)
    rc = lib.c2gpu_bader_metrics(x.ctypes.data, nn.ctypes.data, l2c.ctypes.data, c2l.ctypes.data,
                                 lid.ctypes.data)
    if rc != OK:
        raise C2GpuError(rc, "c2gpu_bader_metrics")
    return l2c, c2l, lid
