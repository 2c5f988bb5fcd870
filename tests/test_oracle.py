"""CPU checks that pin the oracle (the checker the GPU path is compared with)."""
import math

import numpy as np
import pytest

from oracle import oracle
from critic2_amd import radial
from tests.helpers import urea, fmt6, images, node_points


def test_urea_promolecular_nci_matches_reference_golden():
    """tests/015_grdplot/005_nciplot_basic: 8 nodes of the urea cell, real PBE
    atomic tables; the reference's cube files print 6 significant digits."""
    d, x2c, tables, xat, species = urea()
    otabs = [oracle.Table(g.r, g.f, g.fp, g.fpp, g.cutrad) for g in tables]
    pos, idx = images(x2c, xat, lim=3)
    pts = node_points(x2c, d["nstep"])
    out = oracle.promol_nci(pts, pos, species[idx], otabs)
    crho, s = out[:, 3], out[:, 2]
    o = d["options"]
    dat, cg, _ = oracle.nci_cutoffs(crho, s, rhocut=o["rhocut"], dimcut=o["dimcut"], rhoplot=o["rhoplot"],
                                    srhorange=o["srhorange"])
    assert [fmt6(v) for v in crho] == d["dens_cube"]
    # RDG: exact digits where the value is not round-off noise at a symmetric node
    for v, ref in zip(cg, d["grad_cube"]):
        if abs(ref) > 1e-10:
            assert fmt6(v) == ref
        else:
            assert abs(v - ref) < d["cube_atol"]
    ref_dat = np.array(d["dat"])
    assert dat.shape == ref_dat.shape
    np.testing.assert_allclose(dat[:, 0], ref_dat[:, 0], rtol=1e-7)
    np.testing.assert_allclose(dat[:, 1], ref_dat[:, 1], atol=1e-5)


def test_radial_table_builder_matches_oracle_bitwise():
    import os
    T = np.load(os.path.join(os.path.dirname(__file__), "golden", "urea_tables.npz"))
    for s in ("h", "c", "n", "o"):
        g = radial.build_grid1(T["r_" + s], T["rr0_" + s])
        f, fp, fpp, a, b = oracle.grid1_build(T["r_" + s], T["rr0_" + s])
        assert np.array_equal(f, g.f) and np.array_equal(fp, g.fp) and np.array_equal(fpp, g.fpp)
        assert a == g.a and b == g.b


def test_slater_table_reproduces_analytic_density():
    """Table semantics against the closed form of the Slater model: the
    6-point log-grid derivatives must match rho', rho'' analytically."""
    for z in (1, 6, 8):
        g = radial.slater_grid1(z)
        zeta = 2 * z ** (1 / 3)
        # the log-grid derivative combination cancels as r -> 0 (the reference's
        # construction), so derivatives are compared away from the origin
        sel = (g.r > 0.1) & (np.arange(g.ngrid) < g.ngrid - 3)
        r = g.r[sel]
        rho = z * zeta ** 3 / (8 * math.pi) * np.exp(-zeta * r)
        np.testing.assert_allclose(g.f[sel], rho, rtol=1e-12)
        np.testing.assert_allclose(g.fp[sel], -zeta * rho, rtol=1e-6)
        np.testing.assert_allclose(g.fpp[sel], zeta ** 2 * rho, rtol=1e-5)
        # interpolation between nodes (grid1%interp) against the closed form
        ot = oracle.Table(g.r, g.f, g.fp, g.fpp, g.cutrad)
        import ctypes as C
        lib = oracle.load()
        for x in (0.05, 0.3, 1.7, 3.3):
            f, fp, fpp = C.c_double(), C.c_double(), C.c_double()
            lib.oracle_grid1_interp(C.byref(ot.s), x, C.byref(f), C.byref(fp), C.byref(fpp))
            ex = z * zeta ** 3 / (8 * math.pi) * math.exp(-zeta * x)
            assert abs(f.value - ex) <= 1e-5 * ex
            assert abs(fp.value + zeta * ex) <= 1e-4 * zeta * ex


def test_metrics_inverse_and_distances():
    """car2lat is the inverse of lat2car and lat_i_dist the inverse lengths."""
    from critic2_amd.synth import lattice_from_params
    for x2c in (np.diag([10.0, 12.0, 13.0]), lattice_from_params(20, 22, 25, 75, 100, 110)):
        n = (40, 48, 56)
        l2c, c2l, lid = oracle.metrics(x2c, n)
        L = l2c.reshape(3, 3, order="F")
        Ci = c2l.reshape(3, 3, order="F")
        np.testing.assert_allclose(L @ Ci, np.eye(3), atol=1e-13)
        for d1 in (-1, 0, 1):
            for d2 in (-1, 0, 1):
                for d3 in (-1, 0, 1):
                    v = L @ np.array([d1, d2, d3], dtype=float)
                    e = 0.0 if (d1, d2, d3) == (0, 0, 0) else 1 / np.linalg.norm(v)
                    assert abs(lid[(d1 + 1) + 3 * (d2 + 1) + 9 * (d3 + 1)] - e) <= 1e-15 * max(e, 1)


def _gauss_cell(n, centers, widths, x2c):
    f = np.zeros(n)
    ii = [np.arange(m) / m for m in n]
    I, J, K = np.meshgrid(*ii, indexing="ij")
    for c, w in zip(centers, widths):
        d = np.stack([I - c[0], J - c[1], K - c[2]], -1)
        d -= np.round(d)
        dc = d @ x2c.T
        f += np.exp(-(dc ** 2).sum(-1) / w)
    return f.transpose(2, 1, 0).reshape(-1)  # x fastest


def test_bader_single_maximum_is_one_basin():
    """One atom per cell: every point belongs to basin 1, whatever the path."""
    x2c = np.diag([9.0, 10.0, 11.0])
    n = (20, 22, 24)
    f = _gauss_cell(n, [(0.31, 0.42, 0.57)], [3.0], x2c)
    lab, info, xattr, _ = oracle.bader(f, n, x2c, [[0.31, 0.42, 0.57]])
    assert info["nattr"] == 1 and np.all(lab == 1)


def test_bader_sums_identities_and_symmetric_pair():
    """Two identical atoms related by inversion through a grid point: the grid
    is symmetric, so the basin volumes are equal up to the boundary points
    the visiting order assigns, and the counts add up to N."""
    x2c = np.diag([12.0, 12.0, 12.0])
    n = (24, 24, 24)
    a, b = (0.25, 0.5, 0.5), (0.75, 0.5, 0.5)
    f = _gauss_cell(n, [a, b], [2.0, 2.0], x2c)
    lab, info, _, _ = oracle.bader(f, n, x2c, [a, b])
    counts, sums = oracle.basin_sums(lab, 2, [f])
    N = np.prod(n)
    assert counts.sum() == N
    assert abs(sums.sum() - f.sum()) <= 1e-12 * f.sum()
    # boundary planes x = 0 and x = 1/2 hold 2*24*24 points; the rest splits evenly
    assert abs(int(counts[0]) - int(counts[1])) <= 2 * 24 * 24
    assert counts[0] > 0.4 * N and counts[1] > 0.4 * N


def test_bader_labels_follow_steepest_ascent_in_the_interior():
    """Points whose 27-neighbourhood ascends on-grid monotonically to one
    maximum carry that maximum's atom (three separated atoms, triclinic)."""
    from critic2_amd.synth import lattice_from_params
    x2c = lattice_from_params(11, 12, 13, 80, 95, 105)
    n = (22, 24, 26)
    cen = [(0.2, 0.2, 0.2), (0.7, 0.3, 0.6), (0.4, 0.8, 0.7)]
    f = _gauss_cell(n, cen, [1.5, 1.5, 1.5], x2c)
    lab, info, _, _ = oracle.bader(f, n, x2c, cen)
    assert info["status"] == 0 and set(np.unique(lab)) == {1, 2, 3}
    # the grid point nearest each centre belongs to that centre's basin
    for a, c in enumerate(cen):
        p = [int(round(c[q] * n[q])) % n[q] for q in range(3)]
        assert lab[p[0] + n[0] * (p[1] + n[1] * p[2])] == a + 1


@pytest.mark.parametrize("n,mode", [((12, 14, 16), (1, -2, 3)), ((11, 13, 15), (1, -2, 3)),
                                    ((12, 14, 16), (6, 0, 0)), ((12, 13, 16), (0, 0, 8))])
def test_fft_derivatives_closed_form(n, mode):
    """grid3%fft restatement against a single Fourier mode in a triclinic cell:
    rho = 2 + cos(G.r) -> |grad| = |G sin|, d2/dx2 = -Gx^2 cos, lap = -|G|^2 cos.
    Even and odd grids; modes on the Nyquist plane (|m| = n/2) keep their
    second derivatives (a dropped Nyquist bin would zero them) and have no
    gradient at the nodes."""
    from critic2_amd.synth import lattice_from_params
    x2c = lattice_from_params(7, 8, 9, 80, 100, 110)
    B = 2 * math.pi * np.linalg.inv(x2c)  # rows: reciprocal vectors
    m = np.array(mode)
    G = m @ B
    ii = [np.arange(q) / q for q in n]
    I, J, K = np.meshgrid(*ii, indexing="ij")
    ph = 2 * math.pi * (m[0] * I + m[1] * J + m[2] * K)
    rho = (2 + np.cos(ph)).transpose(2, 1, 0).reshape(-1)
    grad = np.abs(np.linalg.norm(G) * np.sin(ph)).transpose(2, 1, 0).reshape(-1)
    xx = (-G[0] ** 2 * np.cos(ph)).transpose(2, 1, 0).reshape(-1)
    lap = (-(G @ G) * np.cos(ph)).transpose(2, 1, 0).reshape(-1)
    np.testing.assert_allclose(oracle.fft_deriv(rho, n, x2c, "grad"), grad, atol=1e-11)
    np.testing.assert_allclose(oracle.fft_deriv(rho, n, x2c, "xx"), xx, atol=1e-11)
    np.testing.assert_allclose(oracle.fft_deriv(rho, n, x2c, "lap"), lap, atol=1e-10)


def test_nci_point_formula_constants():
    """s = |grad rho| / (2 (3 pi^2)^(1/3) rho^(4/3)); uniform-gas check:
    for rho = 1, |grad| = 2 (3 pi^2)^(1/3) the RDG is exactly 1; the
    lambda2 sign follows count(diag > 0) >= 2."""
    cst = 2 * (3 * math.pi ** 2) ** (1 / 3)
    cr, s = oracle.nci_point(1.0, cst, [1.0, 1.0, -1.0])
    assert abs(s - 1.0) < 1e-15 and cr == 100.0
    cr, s = oracle.nci_point(0.5, 0.0, [1.0, -1.0, -1.0])
    assert cr == -50.0 and s == 0.0


def test_tricubic_matrix_matches_reference_vectors():
    """The tricubic coefficient matrix (Hermite tensor product) applied to
    the golden b vectors gives the a = c b the reference's own table gives
    (src/grid3mod@proc.f90:76)."""
    import json
    import os
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "tricubic_ab.json")))
    c = oracle.tricubic_matrix()
    b = np.array(d["b"])
    a = np.array(d["a"])
    np.testing.assert_allclose(b @ c.T, a, rtol=0, atol=1e-12)


def _quadratic_grid(n, coef):
    """f(i,j,k) = quadratic polynomial of the node indices, x-fastest."""
    I, J, K = np.meshgrid(*[np.arange(q, dtype=float) for q in n], indexing="ij")
    c0, cx, cy, cz, cxx, cyy, czz, cxy, cxz, cyz = coef
    f = (c0 + cx * I + cy * J + cz * K + cxx * I * I + cyy * J * J + czz * K * K + cxy * I * J + cxz * I * K
         + cyz * J * K)
    return f.transpose(2, 1, 0).reshape(-1)


def test_tricubic_reproduces_quadratics():
    """Cubic Hermite with central-difference derivatives is exact for
    quadratics: value, gradient and Hessian (fractional derivatives, scaled
    by n) of a quadratic in the node indices, away from the periodic seam."""
    rng = np.random.default_rng(5)
    n = (10, 11, 12)
    coef = rng.standard_normal(10)
    f = _quadratic_grid(n, coef)
    c0, cx, cy, cz, cxx, cyy, czz, cxy, cxz, cyz = coef
    for _ in range(20):
        u = np.array([rng.uniform(1.0, q - 3.0) for q in n])
        y, yp, ypp = oracle.tricubic(f, n, u / np.array(n))
        val = (c0 + cx * u[0] + cy * u[1] + cz * u[2] + cxx * u[0] ** 2 + cyy * u[1] ** 2 + czz * u[2] ** 2
               + cxy * u[0] * u[1] + cxz * u[0] * u[2] + cyz * u[1] * u[2])
        g = np.array([cx + 2 * cxx * u[0] + cxy * u[1] + cxz * u[2],
                      cy + 2 * cyy * u[1] + cxy * u[0] + cyz * u[2],
                      cz + 2 * czz * u[2] + cxz * u[0] + cyz * u[1]])
        h = np.array([[2 * cxx, cxy, cxz], [cxy, 2 * cyy, cyz], [cxz, cyz, 2 * czz]])
        nn = np.array(n, dtype=float)
        assert abs(y - val) < 1e-10
        np.testing.assert_allclose(yp, g * nn, atol=1e-9)
        np.testing.assert_allclose(ypp, h * np.outer(nn, nn), atol=1e-7)


def test_nci_grid_scheme_at_nodes():
    """Grid-interpolation NCI on a triclinic cell: at a node the tricubic
    value is the node value, its gradient the central difference, and its
    second derivative along an axis the one-sided Hermite stencil
    2 f(-1) - 5 f(0) + 4 f(1) - f(2); the Cartesian gradient and Hessian are
    c2x^T g and c2x^T H c2x."""
    from critic2_amd.synth import lattice_from_params
    x2c = lattice_from_params(7.0, 7.5, 8.0, 95, 100, 80)
    n = (9, 10, 11)
    rng = np.random.default_rng(7)
    f3 = 1.0 + rng.random(n)  # (n1,n2,n3)
    f = f3.transpose(2, 1, 0).reshape(-1)
    crho, s = oracle.nci_grid_interp(f, n, x2c)
    c2x = oracle.matinv3(x2c)
    nn = np.array(n, dtype=float)
    cst = 2 * (3 * math.pi ** 2) ** (1 / 3)
    checked = 0
    for (i, j, k) in [(3, 4, 5), (1, 1, 1), (7, 2, 9), (4, 8, 3)]:
        p = np.array([i, j, k])
        fv = lambda d: f3[tuple((p + d) % nn.astype(int))]
        e = np.eye(3, dtype=int)
        g = np.array([0.5 * (fv(e[c]) - fv(-e[c])) for c in range(3)]) * nn
        H = np.zeros((3, 3))
        for a in range(3):
            H[a, a] = (2 * fv(-e[a]) - 5 * fv(0 * e[a]) + 4 * fv(e[a]) - fv(2 * e[a])) * nn[a] ** 2
            for b in range(a + 1, 3):
                # mixed: product of the two 1-D first-derivative stencils at 0
                v = 0.25 * (fv(e[a] + e[b]) - fv(-e[a] + e[b]) - fv(e[a] - e[b]) + fv(-e[a] - e[b]))
                H[a, b] = H[b, a] = v * nn[a] * nn[b]
        gc = c2x.T @ g
        hc = c2x.T @ H @ c2x
        l2 = np.linalg.eigvalsh(hc)[1]
        o = k + n[2] * (j + n[1] * i)
        assert abs(s[o] - np.linalg.norm(gc) / (cst * fv(0 * e[0]) ** (4 / 3))) < 1e-10 * s[o] + 1e-14
        assert abs(abs(crho[o]) - 100 * fv(0 * e[0])) < 1e-12
        assert np.sign(crho[o]) == np.sign(l2)
        checked += 1
    assert checked == 4


def test_bader_attractor_rules_hand_worked():
    """bader@proc.f90:159-199 on two maxima, one on an atom (A at x = 1/4) and
    one off every atom (B at x = 5/8, 4.5 bohr from A): the nearest-atom rule
    sends both to A; RATOM 1 keeps A and appends B as a non-nuclear maximum at
    its grid position, RATOM 5 reaches A from B; NOATOMS numbers the maxima in
    discovery order (the sweep starts at the origin, 3 bohr from A and 4.5
    from B, so A first) and merges B into A when ratom exceeds 4.5 bohr."""
    x2c = np.diag([12.0, 12.0, 12.0])
    n = (24, 24, 24)
    a, b = (0.25, 0.5, 0.5), (0.625, 0.5, 0.5)
    f = _gauss_cell(n, [a, b], [2.0, 2.0], x2c)
    ia = 6 + 24 * (12 + 24 * 12)
    ib = 15 + 24 * (12 + 24 * 12)
    lab, info, xattr, _ = oracle.bader(f, n, x2c, [a], mode=0)
    assert info["nattr"] == 1 and np.all(lab == 1)
    lab, info, xattr, _ = oracle.bader(f, n, x2c, [a], mode=1, ratom=1.0)
    assert info["nattr"] == 2 and lab[ia] == 1 and lab[ib] == 2
    np.testing.assert_array_equal(xattr, [a, b])
    lab, info, xattr, _ = oracle.bader(f, n, x2c, [a], mode=1, ratom=5.0)
    assert info["nattr"] == 1 and np.all(lab == 1)
    lab, info, xattr, _ = oracle.bader(f, n, x2c, [a], mode=2, ratom=4.0)
    assert info["nattr"] == 2 and lab[ia] == 1 and lab[ib] == 2
    np.testing.assert_array_equal(xattr, [a, b])
    lab, info, xattr, _ = oracle.bader(f, n, x2c, [a], mode=2, ratom=5.0)
    assert info["nattr"] == 1 and np.all(lab == 1)
    np.testing.assert_array_equal(xattr, [a])
