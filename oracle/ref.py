"""Calls into oracle/_ref/libc2ref.so: the reference's own vendored LAPACK
(src/oldlibs) and FFT (src/cfftnd.f90), built unchanged by oracle/ref.mk.

TEST INFRASTRUCTURE ONLY: tests/test_oracle_ref.py uses it to check the
oracle's restatements (the 3x3 inverse of matinv, the middle Hessian
eigenvalue of eigsym, the FFT derivative grids of grid3%fft) against the
reference's arithmetic.  ``available()`` is False where the library was not
built (no reference checkout, or no amdflang).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libc2ref.so")
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def load():
    global _lib
    if _lib is None:
        lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        lib.dgetrf_.argtypes = [vp] * 6
        lib.dgetri_.argtypes = [vp] * 7
        lib.dsyev_.argtypes = [vp] * 9 + [C.c_size_t, C.c_size_t]
        lib.cfftnd_.argtypes = [vp] * 4
        _lib = lib
    return _lib


def _i(v):
    return C.byref(C.c_int(v))


def matinv(m):
    """matinv(m, 3) of tools_math@proc.f90:1203-1235: dgetrf, a workspace
    query of dgetri, then dgetri.  m: 3x3 (row, column)."""
    lib = load()
    a = np.asfortranarray(np.asarray(m, dtype=np.float64)).copy(order="F")
    n = a.shape[0]
    ipiv = np.zeros(n, dtype=np.int32)
    info = C.c_int(0)
    lib.dgetrf_(_i(n), _i(n), a.ctypes.data, _i(n), ipiv.ctypes.data, C.byref(info))
    if info.value:
        raise ArithmeticError(f"dgetrf info {info.value}")
    one = np.zeros(1)
    lib.dgetri_(_i(n), a.ctypes.data, _i(n), ipiv.ctypes.data, one.ctypes.data, _i(-1), C.byref(info))
    lwork = int(round(one[0]))
    work = np.zeros(max(lwork, 1))
    lib.dgetri_(_i(n), a.ctypes.data, _i(n), ipiv.ctypes.data, work.ctypes.data, _i(lwork), C.byref(info))
    if info.value:
        raise ArithmeticError(f"dgetri info {info.value}")
    return np.asarray(a)


def eigsym_values(h):
    """Eigenvalues (ascending) of eigsym, tools_math@proc.f90:1005-1039:
    dsyev('V','U') with a workspace query."""
    lib = load()
    a = np.asfortranarray(np.asarray(h, dtype=np.float64)).copy(order="F")
    n = a.shape[0]
    w = np.zeros(n)
    info = C.c_int(0)
    one = np.zeros(1)
    jobz, uplo = C.c_char_p(b"V"), C.c_char_p(b"U")
    lib.dsyev_(jobz, uplo, _i(n), a.ctypes.data, _i(n), w.ctypes.data, one.ctypes.data, _i(-1),
               C.byref(info), 1, 1)
    lwork = int(round(one[0]))
    work = np.zeros(max(lwork, 1))
    lib.dsyev_(jobz, uplo, _i(n), a.ctypes.data, _i(n), w.ctypes.data, work.ctypes.data, _i(lwork),
               C.byref(info), 1, 1)
    if info.value:
        raise ArithmeticError(f"dsyev info {info.value}")
    return w


def cfftnd(z, n, sgn):
    """In-place 3-D cfftnd(3, n, sgn, z) (cfftnd.f90:30): z complex128,
    x fastest; the forward transform (sgn = -1) is scaled by 1/N."""
    lib = load()
    nn = np.asarray(n, dtype=np.int32)
    assert z.dtype == np.complex128 and z.flags.c_contiguous and z.size == int(nn.prod())
    lib.cfftnd_(_i(3), nn.ctypes.data, _i(int(sgn)), z.ctypes.data)


def _det3(m):
    # tools_math@proc.f90 det3, term for term
    return (m[0, 0] * (m[1, 1] * m[2, 2] - m[1, 2] * m[2, 1])
            + m[0, 1] * (m[1, 2] * m[2, 0] - m[1, 0] * m[2, 2])
            + m[0, 2] * (m[1, 0] * m[2, 1] - m[1, 1] * m[2, 0]))


def fft_deriv(f, n, x2c, kind):
    """grid3%fft (grid3mod@proc.f90:1757-1872) with the reference's own
    cfftnd: kind 'grad', 'xx', 'yy', 'zz' or 'lap'; f x-fastest flat."""
    x2c = np.asarray(x2c, dtype=np.float64)
    b = np.zeros((3, 3))
    b[:, 0] = np.cross(x2c[:, 2], x2c[:, 1])
    b[:, 1] = np.cross(x2c[:, 0], x2c[:, 2])
    b[:, 2] = np.cross(x2c[:, 1], x2c[:, 0])
    b = 2.0 * math.pi / abs(_det3(x2c)) * b
    n1, n2, n3 = (int(v) for v in n)

    def idx(m):
        i = np.arange(m)
        return np.where(i <= m // 2, i, i - m).astype(np.float64)

    # vgc(:, igfft), igfft x fastest: arrays indexed [i3, i2, i1]
    I3, I2, I1 = np.meshgrid(idx(n3), idx(n2), idx(n1), indexing="ij")
    vgc = [(I1 * b[c, 0] + I2 * b[c, 1] + I3 * b[c, 2]).reshape(-1) for c in range(3)]
    f = np.asarray(f, dtype=np.float64).reshape(-1)
    if kind == "grad":
        acc = np.zeros(f.size)
        for c in range(3):
            z = f.astype(np.complex128)
            cfftnd(z, n, -1)
            z = vgc[c] * (-z.imag + 1j * z.real)   # vgc * cmplx(-aimag(z), dble(z))
            z = np.ascontiguousarray(z)
            cfftnd(z, n, +1)
            acc = acc + z.real ** 2
        return np.sqrt(acc)
    z = f.astype(np.complex128)
    cfftnd(z, n, -1)
    if kind in ("xx", "yy", "zz"):
        c = "xyz".index(kind[0])
        z = -vgc[c] * vgc[c] * z
    elif kind == "lap":
        z = -(vgc[0] * vgc[0] + vgc[1] * vgc[1] + vgc[2] * vgc[2]) * z
    else:
        raise ValueError(kind)
    z = np.ascontiguousarray(z)
    cfftnd(z, n, +1)
    return z.real.copy()
