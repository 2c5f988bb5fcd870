"""CPU checks of the multipole oracle: genrlm_real (through tosphere and
genylm, the reference's route) against the documented closed forms, the sum
rule and the addition theorem of the Racah-normalised solid harmonics, and
the monopole of the basin multipoles."""
import numpy as np
import pytest
from numpy.polynomial import legendre

from oracle import oracle


def test_rrlm_closed_forms_l2():
    rng = np.random.default_rng(1)
    for _ in range(20):
        x, y, z = rng.normal(size=3)
        r2 = x * x + y * y + z * z
        want = [1, x, z, y, np.sqrt(3) / 2 * (x * x - y * y), np.sqrt(3) * x * z, 0.5 * (3 * z * z - r2),
                np.sqrt(3) * y * z, np.sqrt(3) * x * y]
        np.testing.assert_allclose(oracle.rrlm([x, y, z], 2), want, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("lmax", [5, 8])
def test_rrlm_sum_rule_and_addition_theorem(lmax):
    rng = np.random.default_rng(lmax)
    for _ in range(10):
        a, b = rng.normal(size=3), rng.normal(size=3)
        ra, rb = oracle.rrlm(a, lmax), oracle.rrlm(b, lmax)
        na, nb = np.linalg.norm(a), np.linalg.norm(b)
        cg = a @ b / (na * nb)
        for l in range(lmax + 1):
            s = slice(l * l, (l + 1) * (l + 1))
            assert np.sum(ra[s] ** 2) == pytest.approx(na ** (2 * l), rel=1e-12)
            pl = legendre.legval(cg, [0] * l + [1])
            assert np.sum(ra[s] * rb[s]) == pytest.approx(na ** l * nb ** l * pl, rel=1e-10, abs=1e-12)


def test_multipoles_monopole_is_the_basin_sum():
    rng = np.random.default_rng(3)
    n = (6, 7, 8)
    f = rng.random(np.prod(n))
    idg = rng.integers(1, 4, np.prod(n)).astype(np.int32)
    xattr = rng.random((3, 3))
    mp = oracle.multipoles(idg, f, n, np.diag([5.0, 6.0, 7.0]), xattr, lmax=3)
    for a in range(3):
        assert mp[a, 0] == pytest.approx(f[idg == a + 1].sum(), rel=1e-14)
