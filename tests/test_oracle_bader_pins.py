"""Hand-traced pins of the oracle's Bader restatement (oracle/bader.c).

Each grid is two-dimensional (n2 = 1, so the j neighbours of a point are the
point itself) with a unit grid spacing (x2c = diag(n1, 1, n3) bohr, hence
lat2car = car2lat = identity and rho_grad_dir returns the central
differences themselves).  Rows of the literals below are k = 0, 1, ...;
columns are i = 0, 1, ...; points are written (i,k).  Attractors are atoms
placed on the grid maxima, so a label names the maximum a point belongs to.

The decisive steps of each case are traced by hand from critic2's
src/bader@proc.f90 in the comments: step_neargrid :455-494 (g scaled by
1/max|g|; pm = p + nint(g); dr = dr + g - nint(g); pm = pm + nint(dr);
dr = dr - nint(dr); dr reset only at a path's first step; an on-grid step
when pm is already on the path), rho_grad_dir :532-567 (an axis is zeroed
when both its neighbours are strictly lower), step_ongrid :500-527 (first
strict maximum of rho_c + (rho_n - rho_c)/|d| in d1, d2, d3 order),
max_neargrid :427-450 (stop after moving onto a known == 2 point),
the sweep :151-215 (a committed path relabels every point on it) and
refine_edge :300-422.

Every mutation switch of the oracle (a plausible mistake in one rule) fails
at least one of these pins (test_each_mutation_fails_a_pin).
"""
import numpy as np
import pytest

from oracle import oracle


def _grid(rows):
    F = np.asarray(rows, dtype=np.float64)
    n3, n1 = F.shape
    f = np.zeros((n1, 1, n3))
    f[:, 0, :] = F.T
    return f.reshape(-1, order="F"), (n1, 1, n3)


def run(rows, maxima, sweep_only=False):
    """Oracle labels as L[(i, k)], with one atom on each listed maximum."""
    f, n = _grid(rows)
    x2c = np.diag([float(n[0]), 1.0, float(n[2])])
    xat = np.array([[i / n[0], 0.0, k / n[2]] for (i, k) in maxima])
    lab, info, _, _ = oracle.bader(f, n, x2c, xat, sweep_only=sweep_only)
    L = lab.reshape((n[0], n[2]), order="F")
    return {(i, k): int(L[i, k]) for i in range(n[0]) for k in range(n[2])}, info


# ---------------------------------------------------------------- dr carry, refine pass 1
# A slope rising along k with a small +i tilt feeds two maxima, A = (0,6)
# (label 1) and B = (3,6) (label 2).
RAMP = [
    [10, 11, 12, 13, 11, 9],
    [13, 14, 15, 16, 14, 12],
    [16, 17, 18, 19, 17, 15],
    [19, 20, 21, 22, 20, 18],
    [22, 23, 24, 25, 23, 21],
    [30, 28, 29, 31, 28, 27],
    [40, 30, 32, 40, 30, 29],
    [7, 8, 9, 10, 8, 6],
]
RAMP_MAX = [(0, 6), (3, 6)]


def test_ramp_first_path_carries_dr():
    """The first start, (0,0), meets no known == 2 point, so its label is its
    free trajectory's end.  With gx = (rho(i+1) - rho(i-1))/2 and gz likewise:
      (0,0): gx = (11-9)/2 = 1, gz = (13-7)/2 = 3 -> g = (1/3, 1):
             pm = (0,1), dr_i = 1/3
      (0,1): gx = (14-12)/2 = 1, gz = (16-10)/2 = 3 -> dr_i = 2/3, nint = 1:
             pm = (1,2), dr_i = -1/3
      (1,2): g = (1/3, 1) ((18-16)/2, (20-14)/2) -> dr_i ~ 0: pm = (1,3)
      (1,3): g = (1/3, 1) -> dr_i = 1/3: pm = (1,4)
      (1,4): gx = (24-22)/2 = 1, gz = (28-20)/2 = 4 -> dr_i = 1/3 + 1/4:
             nint = 1: pm = (2,5), dr_i = -5/12
      (2,5): gx = (31-28)/2, gz = (32-24)/2 -> g = (3/8, 1), dr_i = -1/24:
             pm = (2,6)
      (2,6): gx = (40-30)/2 = 5; rho(2,7) = 9 and rho(2,5) = 29 are both
             below 32, so gz = 0 -> g = (1, 0): pm = (3,6)
      (3,6): both neighbours lower on every axis, g = 0 and is_max: B.
    With dr reset at every step the path climbs column 0 to A instead.
    Refinement keeps (0,0) in B: its re-ascent is the same trajectory."""
    sweep, _ = run(RAMP, RAMP_MAX, sweep_only=True)
    final, info = run(RAMP, RAMP_MAX)
    assert sweep[(0, 0)] == 2 and final[(0, 0)] == 2
    assert info["nattr"] == 2


def test_ramp_refine_pass1_reassigns():
    """(0,1) lies on the first path and takes B in the sweep (no later path
    passes it: it is never a start).  Refine pass 1 marks it (its neighbour
    (0,2) belongs to A, and it is no maximum) and re-ascends it with a fresh
    dr:
      (0,1): g = (1/3, 1) -> pm = (0,2), dr_i = 1/3
      (0,2): gx = (17-15)/2 = 1, gz = (19-13)/2 = 3 -> dr_i = 2/3:
             pm = (1,3), dr_i = -1/3
      (1,3): g = (1/3, 1) -> dr_i ~ 0: pm = (1,4)
      (1,4): g = (1/4, 1) -> dr_i = 1/4: pm = (1,5)
      (1,5): gx = (29-30)/2 = -1/2, gz = (30-23)/2 -> g = (-1/7, 1),
             dr_i ~ 0.107: pm = (1,6)
      (1,6): gx = (32-40)/2 = -4; rho(1,7) = 8 and rho(1,5) = 28 are below
             30, gz = 0 -> pm = (0,6) = A.
    Every point of this path is a marked edge point or next to one, so the
    marking cleared its known == 2 flag and the ascent runs to A: (0,1) is
    reassigned from 2 to 1."""
    sweep, _ = run(RAMP, RAMP_MAX, sweep_only=True)
    final, _ = run(RAMP, RAMP_MAX)
    assert sweep[(0, 1)] == 2
    assert final[(0, 1)] == 1


# ---------------------------------------------------------------- nint at +-0.5
TIE = [
    [10, 11, 12, 11, 10, 9],
    [14, 15, 16, 15, 14, 13],
    [18, 17, 22, 21, 16, 16],
    [30, 12, 32, 25, 18, 19],
    [9, 10, 11, 10, 9, 8],
    [6, 7, 8, 7, 6, 5],
]
TIE_MAX = [(0, 3), (2, 3)]  # A = 1, B = 2


def test_nint_rounds_half_away_from_zero():
    """  (0,0): gx = (11-9)/2 = 1, gz = (14-6)/2 = 4 -> g = (1/4, 1):
             pm = (0,1), dr_i = 1/4
      (0,1): gx = (15-13)/2 = 1, gz = (18-10)/2 = 4 -> dr_i = 1/2 exactly;
             Fortran nint(0.5) = 1: pm = (1,2), dr_i = -1/2
      (1,2): gx = (22-18)/2 = 2; rho(1,3) = 12 and rho(1,1) = 15 are below
             17, gz = 0 -> g = (1, 0): pm_i = 1 + 1 + nint(-1/2) = 1 + 1 - 1,
             back to (1,2), which is on the path: on-grid step.  Candidates
             17 + (rho - 17)/|d|: (0,3) -> 17 + 13/sqrt2 = 26.19,
             (2,2) -> 22, (2,3) -> 17 + 15/sqrt2 = 27.61 (largest): (2,3) = B.
    Rounding half to even instead leaves (0,1) at pm = (0,2) with dr_i = 1/2,
    where rho(1,2) = 17 and rho(5,2) = 16 are below 18 (g = (0, 1)), and the
    path ends at (0,3) = A."""
    sweep, _ = run(TIE, TIE_MAX, sweep_only=True)
    final, _ = run(TIE, TIE_MAX)
    assert sweep[(0, 0)] == 2 and final[(0, 0)] == 2
    # (0,1) is B from the first path, A from its own re-ascent:
    # (0,1) -> (0,2) (g = (1/4, 1)) -> (0,3) (g = (0, 1), dr_i = 1/4)
    assert sweep[(0, 1)] == 2 and final[(0, 1)] == 1


# ---------------------------------------------------------------- known == 2 stop, pass >= 2 rule
STOP = [
    [3, 0, 0, 3],
    [4, 2, 4, 3],
    [0, 0, 5, 1],
]
STOP_MAX = [(0, 1), (2, 2)]  # 1, 2


def test_path_stops_at_known2():
    """Sweep (starts in i-outer, k-inner order; n = (4, 1, 3)):
      start (0,0): g = (-3/4, 1) -> (3,1) (dr_i = 1/4); (3,1): gx = (4-4)/2
        = 0, gz = (1-3)/2 = -1 -> (3,0); (3,0): g = (1, 2/3) -> (0,1)
        (dr_k = -1/3); (0,1) is a maximum: label 1 on (0,0),(3,1),(3,0),(0,1).
      start (0,2): g = (-1, -1) -> (3,1) -> (3,0) -> (0,1): label 1.
      start (1,0): g = (-1, 2/3) -> (0,1): label 1.  Committing (0,1) checks
        its face neighbour (0,0): it and (0,1),(0,2),(1,0),(3,0) all carry 1,
        so known(0,0) = 2.
      start (1,1): g = 0 (both neighbours lower on k, equal 4s on i), not a
        maximum: on-grid step; (2,2) gives 2 + 3/sqrt2 = 4.12 > 4: new
        maximum (2,2), label 2.  start (1,2): g = (1, -0.4) -> (2,2): 2.
      start (2,0): gx = (3-0)/2, gz = (4-5)/2 -> g = (1, -1/3): (3,0),
        dr_k = -1/3; (3,0): g = (1, 2/3): pm_k = 0 + 1 + nint(-2/3) = 0,
        pm = (0,0), which is known == 2: the path stops with label 1.
    Without the stop it continues (0,0) -> (3,1), whose next point (3,0) is
    on the path, and the on-grid step from (3,1) reaches (2,2): label 2."""
    sweep, _ = run(STOP, STOP_MAX, sweep_only=True)
    assert sweep[(2, 0)] == 1 and sweep[(3, 0)] == 1 and sweep[(0, 0)] == 1
    # refine pass 1 re-ascends (2,0) with the flag cleared: (2,0) ends in 2
    final, info = run(STOP, STOP_MAX)
    assert final[(2, 0)] == 2
    assert info["refine_iters"] == 2


def test_refine_pass2_reads_known_at_p_plus_111():
    """Pass 1 changes only (2,0) (to -2, see above).  Pass 2's only source is
    (2,0): its 3x3x3 loop (with n2 = 1 every (d1, d3) repeats for d2 = -1, 0,
    1) negates every positive non-maximum around it, the seven points of
    columns i = 1, 2, 3 other than (2,0) and the maximum (2,2), with
    known = -1, and sets known(2,0) = -2 (a negative point with known 0).
    After the loop pt holds p + (1,1,1), i.e. (3,0,1), whose known is -1:
    the source is restored to +2.  Pass 2 therefore re-ascends 7 points;
    reading known(p) = -2 instead would keep (2,0) negative and re-ascend 8."""
    run(STOP, STOP_MAX)
    counts = oracle.refine_counts()
    assert len(counts) == 2 and counts[1] == 7


# ---------------------------------------------------------------- step_ongrid ties
ONGRID = [
    [1, 4, 1, 0],
    [4, 5, 4, 0],
    [9, 4, 9, 0],
    [2, 2, 2, 0],
]
ONGRID_MAX = [(0, 2), (2, 2)]


def test_ongrid_first_maximum_wins():
    """(1,1) = 5 has both neighbours lower on i (4, 4) and on k (4, 4): g = 0
    and it is no maximum, so every path through it takes an on-grid step.
    (0,2) and (2,2) (both 9) tie at 5 + 4/sqrt2; the strict > keeps the first
    in d1, d2, d3 order, d1 = -1: (0,2).  (1,1) is a marked edge point in
    refine pass 1 and its re-ascent is the same on-grid step."""
    final, _ = run(ONGRID, ONGRID_MAX)
    assert final[(1, 1)] == 1


# ---------------------------------------------------------------- a committed path relabels
RELABEL = [
    [2, 17, 16],
    [18, 9, 18],
    [16, 0, 9],
]
RELABEL_MAX = [(0, 1), (2, 1)]


def test_later_path_relabels_a_maximum():
    """n1 = 3, so i = 0 and i = 2 are neighbours: (0,1) and (2,1) (both 18)
    form a plateau.
      start (0,0): gx = (17-16)/2 = 1/2, gz = (18-16)/2 = 1 -> g = (1/2, 1):
        pm_i = 0 + 1 + nint(-1/2) = 0, dr_i = 1/2; pm = (0,1).
        (0,1): rho(1,1) = 9 < 18 but rho(2,1) = 18 is not, so gx = -9/2;
        gz = 0 (16 and 2 below 18): g = (-1, 0): pm_i = -1 + nint(1/2) = 0,
        back to (0,1), on the path: on-grid step, where no candidate exceeds
        18: no move.  New maximum (0,1): label 1.
      start (0,2): gz = (2-18)/2 < 0, gx = 0 (0 and 9 below 16): -> (0,1);
        (0,1) with dr = 0: g = (-1, 0) -> (2,1); (2,1): gx = (18-9)/2 > 0 ->
        (0,1), on the path: on-grid step, no move.  New maximum (2,1): label
        2 on (0,2), (0,1), (2,1): the maximum (0,1) now carries 2.
    Refinement does not mark (0,1) (it is a maximum), so it stays 2."""
    sweep, _ = run(RELABEL, RELABEL_MAX, sweep_only=True)
    final, _ = run(RELABEL, RELABEL_MAX)
    assert sweep[(0, 1)] == 2 and final[(0, 1)] == 2
    # (0,0) is relabelled by start (1,2)'s path, which passes (0,0) on its way
    assert final[(0, 0)] == 2


# ---------------------------------------------------------------- every mutation is caught
PINS = [
    test_ramp_first_path_carries_dr,
    test_ramp_refine_pass1_reassigns,
    test_nint_rounds_half_away_from_zero,
    test_path_stops_at_known2,
    test_refine_pass2_reads_known_at_p_plus_111,
    test_ongrid_first_maximum_wins,
    test_later_path_relabels_a_maximum,
]


@pytest.mark.parametrize("name", sorted(oracle.MUTATIONS))
def test_each_mutation_fails_a_pin(name):
    """Each mutation switch replaces one rule of bader@proc.f90 by a plausible
    mistake; at least one pin above must fail with it on."""
    failed = []
    with oracle.mutation(name):
        for pin in PINS:
            try:
                pin()
            except AssertionError:
                failed.append(pin.__name__)
    assert failed, f"mutation {name} passes every pin"


def test_pins_pass_without_mutation():
    for pin in PINS:
        pin()


# ---------------------------------------------------------------- DISCARD (bader@proc.f90:185-191)

def _run_noatoms(rows, discard):
    f, n = _grid(rows)
    x2c = np.diag([float(n[0]), 1.0, float(n[2])])
    lab, info, xattr, _ = oracle.bader(f, n, x2c, np.zeros((0, 3)), mode=2, ratom=1e-3, discard=discard)
    L = lab.reshape((n[0], n[2]), order="F")
    return {(i, k): int(L[i, k]) for i in range(n[0]) for k in range(n[2])}, info, xattr


def test_discard_keeps_rejected_basin_unlabelled():
    """NOATOMS: the maxima A = (0,6) and B = (3,6) of the RAMP grid are new
    attractors in discovery order.  The first start's path ends at B (see
    test_ramp_first_path_carries_dr), so B is attractor 1.  Discarding A
    (the expression is evaluated at x2c*dv = (0, 0, 6) bohr) leaves every
    path that ends there with label 0: one attractor, A's maximum 0, (0,0)
    in B."""
    calls = []

    def at_a(x):
        calls.append(tuple(x))
        return abs(x[0]) < 1e-9 and abs(x[2] - 6.0) < 1e-9

    L, info, xattr = _run_noatoms(RAMP, at_a)
    assert info["nattr"] == 1
    assert np.allclose(xattr[0], [3 / 6, 0.0, 6 / 8])
    assert L[(3, 6)] == 1 and L[(0, 6)] == 0 and L[(0, 0)] == 1
    assert (0.0, 0.0, 6.0) in calls and (3.0, 0.0, 6.0) in calls
    # without the expression both maxima are attractors, B first
    L2, info2, xattr2 = _run_noatoms(RAMP, None)
    assert info2["nattr"] == 2 and L2[(3, 6)] == 1 and L2[(0, 6)] == 2
    # a point left at 0 belonged to A's basin without the expression
    for p, v in L.items():
        if v == 0:
            assert L2[p] == 2, p


def test_discard_first_maximum_renumbers_the_next():
    """Discarding B, the first path's maximum, leaves that path at label 0
    and makes A (found later) attractor 1."""
    L, info, xattr = _run_noatoms(RAMP, lambda x: abs(x[0] - 3.0) < 1e-9 and abs(x[2] - 6.0) < 1e-9)
    assert info["nattr"] == 1
    assert np.allclose(xattr[0], [0.0, 0.0, 6 / 8])
    assert L[(3, 6)] == 0 and L[(0, 6)] == 1
