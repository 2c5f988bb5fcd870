"""Host mirror of critic2's cube readers, grid3%read_cube and
grid3%read_bincube (grid3mod@proc.f90:512-572, :575-635), over the GPU entry
points c2gpu_cube_read / c2gpu_cube_parse / c2gpu_bincube_read.

The grid comes back as critic2 stores it, f(n1,n2,n3) with x fastest (a flat
array of n1*n2*n3 doubles), and the header as critic2 uses it: x2cl(:,i) is
the file's axis vector times n(i) (:546-547) and, given the cell, x0 in
crystallographic coordinates c2x (x0 - molx0) (:540-541).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import Context, CubeHeader


def _header_dict(h: CubeHeader, x2c=None, molx0=None) -> dict:
    n = tuple(int(v) for v in h.n)
    step = np.array(list(h.step), dtype=np.float64).reshape(3, 3).T  # columns = axes
    out = {
        "nat": int(h.nat), "ismo": bool(h.ismo), "n": n,
        "x0_cart": np.array(list(h.x0), dtype=np.float64), "step": step,
        "x2cl": step * np.array(n, dtype=np.float64)[None, :],
        "data_offset": int(h.data_offset),
    }
    if x2c is not None:
        c2x = np.linalg.inv(np.asarray(x2c, dtype=np.float64))
        m0 = np.zeros(3) if molx0 is None else np.asarray(molx0, dtype=np.float64)
        out["x0"] = c2x @ (out["x0_cart"] - m0)
    return out


def read_header(ctx: Context, path: str, binary: bool = False, x2c=None, molx0=None) -> dict:
    h = CubeHeader()
    ctx.check(ctx.lib.c2gpu_cube_read_header(ctx.h, os.fsencode(path), 1 if binary else 0, C.byref(h)),
              "c2gpu_cube_read_header")
    return _header_dict(h, x2c, molx0)


def _out(ctx, hdr, device):
    npts = int(np.prod(hdr["n"]))
    return ctx.empty((npts,), np.float64) if device else np.empty(npts, dtype=np.float64)


def read_cube(ctx: Context, path: str, x2c=None, molx0=None, device: bool = False):
    """grid3%read_cube: (header dict, f).  device=True leaves f in HBM (a
    DeviceArray)."""
    hdr = read_header(ctx, path, False)
    f = _out(ctx, hdr, device)
    h = CubeHeader()
    flags = _lib.PTR_DEVICE if device else 0
    ctx.check(ctx.lib.c2gpu_cube_read(ctx.h, os.fsencode(path), flags, C.byref(h), _lib.ptr(f)),
              "c2gpu_cube_read")
    return _header_dict(h, x2c, molx0), f


def parse_cube(ctx: Context, text, x2c=None, molx0=None):
    """The text of a cube already in memory: bytes (host; f returned as a
    numpy array) or a (DeviceArray of uint8, length) pair (device; f returned
    as a DeviceArray)."""
    h = CubeHeader()
    if isinstance(text, (bytes, bytearray)):
        buf = np.frombuffer(bytes(text), dtype=np.uint8)
        ptr, length, flags = buf.ctypes.data, len(buf), 0
    else:
        dev, length = text
        ptr, flags = dev.data_ptr(), _lib.PTR_DEVICE
    # header first (f NULL), then the values into an array of its size
    ctx.check(ctx.lib.c2gpu_cube_parse(ctx.h, ptr, int(length), flags, C.byref(h), None), "c2gpu_cube_parse")
    f = _out(ctx, _header_dict(h), bool(flags))
    ctx.check(ctx.lib.c2gpu_cube_parse(ctx.h, ptr, int(length), flags, C.byref(h), _lib.ptr(f)),
              "c2gpu_cube_parse")
    return _header_dict(h, x2c, molx0), f


def read_bincube(ctx: Context, path: str, x2c=None, molx0=None, device: bool = False):
    """grid3%read_bincube: (header dict, f)."""
    hdr = read_header(ctx, path, True)
    f = _out(ctx, hdr, device)
    h = CubeHeader()
    flags = _lib.PTR_DEVICE if device else 0
    ctx.check(ctx.lib.c2gpu_bincube_read(ctx.h, os.fsencode(path), flags, C.byref(h), _lib.ptr(f)),
              "c2gpu_bincube_read")
    return _header_dict(h, x2c, molx0), f


def format_body(ctx: Context, f, n, style: int = _lib.CUBE_E22_14, zfast: bool = False) -> bytes:
    """The body text of a cube for f (host numpy array, x fastest unless
    zfast): critic2's writegrid_cube (styles 0, 1) or nciplot's
    write_cube_body (style 2) records."""
    nn = np.asarray(n, dtype=np.int32)
    ln = C.c_uint64()
    ctx.check(ctx.lib.c2gpu_cube_format(ctx.h, None, nn.ctypes.data, style, int(zfast), 0, None, 0, C.byref(ln)),
              "c2gpu_cube_format")
    out = np.empty(ln.value, dtype=np.uint8)
    if hasattr(f, "data_ptr"):
        d = ctx.empty((ln.value,), np.uint8)
        ctx.check(ctx.lib.c2gpu_cube_format(ctx.h, f.data_ptr(), nn.ctypes.data, style, int(zfast),
                                            _lib.PTR_DEVICE, d.data_ptr(), ln.value, C.byref(ln)),
                  "c2gpu_cube_format")
        return d.numpy().tobytes()
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    ctx.check(ctx.lib.c2gpu_cube_format(ctx.h, f.ctypes.data, nn.ctypes.data, style, int(zfast), 0,
                                        out.ctypes.data, ln.value, C.byref(ln)), "c2gpu_cube_format")
    return out.tobytes()


def write_cube(ctx: Context, path: str, n, x0, step, zat=(), xat=(), f=None, style: int = _lib.CUBE_E22_14):
    """writegrid_cube, text form: step 3x3 with column i = axis-i vector,
    xat Cartesian (nat, 3); f x fastest (host or DeviceArray) or None for
    the header alone."""
    h = CubeHeader()
    zat = np.ascontiguousarray(zat, dtype=np.int32).reshape(-1)
    xat = np.ascontiguousarray(xat, dtype=np.float64).reshape(-1)
    h.nat = len(zat)
    for a in range(3):
        h.n[a] = int(n[a])
        h.x0[a] = float(x0[a])
    st = np.asarray(step, dtype=np.float64)
    for i in range(3):
        for a in range(3):
            h.step[a + 3 * i] = float(st[a, i])
    flags = _lib.PTR_DEVICE if hasattr(f, "data_ptr") else 0
    fp = None if f is None else (f if flags else np.ascontiguousarray(f, dtype=np.float64).reshape(-1))
    ctx.check(ctx.lib.c2gpu_cube_write(ctx.h, os.fsencode(path), C.byref(h), zat.ctypes.data if len(zat) else None,
                                       xat.ctypes.data if len(xat) else None, _lib.ptr(fp), style, flags),
              "c2gpu_cube_write")
