"""c2gpu_basin_sums_noalloc (the per-basin sums with a fixed device scratch,
for callers that cannot allocate) against c2gpu_basin_sums through the C
ABI: two calls in a row, the second with a smaller input; results equal bit
for bit."""
import ctypes as C

import numpy as np
import pytest

from critic2_amd import bader as gb

pytestmark = pytest.mark.gpu


def _noalloc(ctx, lab_d, npts, nattr, fields_d, want_counts=True):
    lib = ctx.lib
    lib.c2gpu_basin_sums_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p]
    nf = len(fields_d)
    ptrs = (C.c_void_p * max(nf, 1))(*[f.data_ptr() for f in fields_d])
    sums_d = ctx.empty((max(nf * nattr, 1),), np.float64)
    counts_d = ctx.empty((nattr,), np.int64)
    rc = lib.c2gpu_basin_sums_noalloc(ctx.h, lab_d.data_ptr(), npts, nattr, nf, ptrs, sums_d.data_ptr(),
                                      counts_d.data_ptr() if want_counts else None)
    ctx.check(rc, "c2gpu_basin_sums_noalloc")
    ctx.sync()
    sums = sums_d.numpy()[:nf * nattr].reshape(nf, nattr) if nf else np.zeros((0, nattr))
    counts = counts_d.numpy()
    sums_d.free()
    counts_d.free()
    return counts, sums


def _case(rng, npts, nattr, nf):
    # spatially coherent labels (runs), a few out-of-range labels, random fields
    runs = rng.integers(1, 400, size=npts // 50 + 2)
    lab = np.repeat(rng.integers(0, nattr + 2, size=runs.size), runs)[:npts].astype(np.int32)
    if lab.size < npts:
        lab = np.concatenate([lab, np.ones(npts - lab.size, dtype=np.int32)])
    fields = [rng.standard_normal(npts) for _ in range(nf)]
    return lab, fields


def test_noalloc_sums_equal_basin_sums(ctx):
    rng = np.random.default_rng(5)
    for npts, nattr, nf in ((3_000_001, 300, 2), (40_000, 17, 1)):  # the second call is smaller
        print(f"[case] npts {npts} nattr {nattr} nfields {nf}", flush=True)
        lab, fields = _case(rng, npts, nattr, nf)
        lab_d = ctx.to_device(lab)
        fields_d = [ctx.to_device(f) for f in fields]
        c_ref, s_ref = gb.basin_sums(ctx, lab, nattr, fields)
        c_new, s_new = _noalloc(ctx, lab_d, npts, nattr, fields_d)
        assert np.array_equal(c_new, c_ref)
        assert np.array_equal(np.asarray(s_new).view(np.uint64), np.asarray(s_ref).reshape(nf, nattr).view(np.uint64))
        # counts in plain arithmetic: the scratch holds no stale partials
        want = np.bincount(lab[(lab >= 1) & (lab <= nattr)], minlength=nattr + 1)[1:]
        assert np.array_equal(c_new, want)
        lab_d.free()
        for f in fields_d:
            f.free()


def test_noalloc_sums_refuse_sizes_beyond_the_scratch(ctx):
    lab_d = ctx.to_device(np.ones(1000, dtype=np.int32))
    f_d = ctx.to_device(np.ones(1000))
    lib = ctx.lib
    lib.c2gpu_basin_sums_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p]
    ptrs = (C.c_void_p * 1)(f_d.data_ptr())
    out = ctx.empty((4096,), np.float64)
    rc = lib.c2gpu_basin_sums_noalloc(ctx.h, lab_d.data_ptr(), 1000, 1025, 1, ptrs, out.data_ptr(), None)
    assert rc != 0  # 1025 * 2 bins > 2048
    # a host pointer is refused before any launch
    host = np.zeros(16)
    rc = lib.c2gpu_basin_sums_noalloc(ctx.h, lab_d.data_ptr(), 1000, 4, 1, ptrs, host.ctypes.data, None)
    assert rc != 0
    for a in (lab_d, f_d, out):
        a.free()
