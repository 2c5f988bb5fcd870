"""The oracle against the reference's own arithmetic: critic2's vendored
LAPACK (src/oldlibs) and FFT (src/cfftnd.f90), built unchanged into
oracle/_ref by oracle/ref.mk.  Skipped where that library was not built.

- matinv (tools_math@proc.f90:1203-1235; bader@proc.f90:121-122 inverts
  lat2car with it): the oracle's 3x3 restatement of dgetrf2/dgetri equals
  the reference's car2lat bit for bit (the library restates the same
  operation order and is compared with the oracle on every GPU Bader test).
- eigsym (tools_math@proc.f90:1005-1039, nci@proc.f90:568): the sign of the
  middle eigenvalue the oracle uses equals dsyev's.
- grid3%fft (grid3mod@proc.f90:1757-1872): the oracle's numpy FFT
  derivative grids equal those computed with cfftnd to round-off.
"""
import math

import numpy as np
import pytest

from critic2_amd.synth import config, lattice_from_params
from oracle import oracle, ref

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (needs the reference and amdflang)")


def _cells(rng, count):
    for _ in range(count):
        a = rng.uniform(3.0, 80.0, 3)
        ang = rng.uniform(55.0, 125.0, 3)
        try:
            x2c = lattice_from_params(*a, *ang)
        except (ValueError, FloatingPointError):
            continue
        if not np.all(np.isfinite(x2c)) or abs(np.linalg.det(x2c)) < 1e-3 * np.prod(a):
            continue
        yield x2c, rng.integers(3, 800, 3)


def test_car2lat_bits_equal_reference_lapack():
    rng = np.random.default_rng(12)
    cells = list(_cells(rng, 3000))
    for c in (1, 2, 3, 5):  # the benchmark cells at their grid sizes
        cl = config(c) if c != 3 else config(3, n_override=(512, 512, 512))
        cells.append((np.asarray(cl.x2c), np.asarray(cl.n)))
    assert len(cells) > 2000
    for x2c, n in cells:
        lat2car = np.asarray(x2c) / np.asarray(n, dtype=np.float64)[None, :]
        want = ref.matinv(lat2car)
        got = oracle.matinv3(lat2car)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (x2c, n)
        _, c2l, _ = oracle.metrics(x2c, n)
        assert np.array_equal(c2l.reshape(3, 3, order="F").view(np.uint64), want.view(np.uint64))


def test_matinv_restatement_would_catch_a_reordered_inverse():
    """A plausible mistake (the textbook cofactor inverse) differs from the
    reference's bits on triclinic metrics, so the bitwise test above has
    power."""
    rng = np.random.default_rng(3)
    differ = 0
    for x2c, n in _cells(rng, 200):
        lat2car = np.asarray(x2c) / np.asarray(n, dtype=np.float64)[None, :]
        cof = np.linalg.inv(lat2car)
        differ += not np.array_equal(cof, ref.matinv(lat2car))
    assert differ > 50


def _sym(rng, scale_mid):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    lam = np.array([rng.uniform(-2, -0.1), scale_mid * rng.choice([-1.0, 1.0]), rng.uniform(0.1, 2)])
    rng.shuffle(lam)
    return (q * lam) @ q.T


def test_lambda2_sign_equals_dsyev():
    """sign(rho, lambda2) of nci@proc.f90:599 with lambda2 from the oracle's
    eigenvalues equals the sign from the reference's dsyev on random
    symmetric matrices, including middle eigenvalues down to 1e-9 of the
    others (well above the eigensolvers' round-off)."""
    rng = np.random.default_rng(7)
    for t in range(4000):
        mid = 10.0 ** rng.uniform(-9, 0)
        h = _sym(rng, mid)
        w = ref.eigsym_values(h)
        cr, _ = oracle_nci_full(h)
        assert (cr < 0) == (w[1] < 0), (h, w)


def oracle_nci_full(h):
    import ctypes as C
    lib = oracle.load()
    e = np.asfortranarray(h).reshape(-1, order="F").copy()
    cr, s = C.c_double(), C.c_double()
    lib.oracle_nci_point(1.0, 0.5, e.ctypes.data, 1, C.byref(cr), C.byref(s))
    return cr.value, s.value


@pytest.mark.parametrize("n", [(8, 8, 8), (9, 7, 5), (12, 10, 6), (16, 15, 9)])
def test_fft_derivatives_equal_cfftnd(n):
    """The oracle's grid3%fft (numpy FFT) against the same formulas with the
    reference's cfftnd, triclinic cell, random grid: every kind the path
    uses agrees to 1e-13 of the largest value (fftpack and numpy round-off
    differ at the 1e-15 level)."""
    rng = np.random.default_rng(sum(n))
    x2c = lattice_from_params(6.0, 7.0, 8.0, 80.0, 95.0, 105.0)
    f = rng.uniform(0.0, 1.0, int(np.prod(n)))
    for kind in ("grad", "xx", "yy", "zz", "lap"):
        want = ref.fft_deriv(f, n, x2c, kind)
        got = oracle.fft_deriv(f, n, x2c, kind)
        scale = np.max(np.abs(want))
        assert np.max(np.abs(got - want)) <= 1e-13 * scale, kind


def test_cfftnd_forward_scaling_and_sign():
    """cfftnd's forward transform of a single mode exp(+2 pi i m.x/n) puts
    1 (scaled by 1/N) at index m: the convention the restatement assumes."""
    n = (6, 5, 4)
    m = (2, 1, 3)
    i, j, k = np.meshgrid(*[np.arange(v) for v in n], indexing="ij")
    ph = 2 * math.pi * (m[0] * i / n[0] + m[1] * j / n[1] + m[2] * k / n[2])
    z = np.exp(1j * ph).transpose(2, 1, 0).reshape(-1).copy()  # x fastest
    ref.cfftnd(z, n, -1)
    zz = z.reshape(n[2], n[1], n[0]).transpose(2, 1, 0)
    assert abs(zz[m] - 1.0) < 1e-13
    zz[m] = 0.0
    assert np.max(np.abs(zz)) < 1e-13
