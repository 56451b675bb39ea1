"""The endgame estimates of the oracle (otr_endgame: ratio method on the
last Taylor coefficients of x(t + s), the condition estimate, |x|) against
closed forms, and the classification built on them
(phcpack_amd.tracker.classify_endpoints):

  x^2 = 1 - t        x = (1-t)^(1/2): a branch point at t = 1, winding 2
  x^3 = 1 - t        winding 3
  (1 - t) x = 1      x = 1/(1-t): a pole at t = 1 (the path diverges)
  x = 1 + t          regular: every coefficient beyond order 1 is zero
  x^2 = 1 - t, y = 2 + t, decoupled: the dominant component decides

The ratio method is exact for c (s* - s)^alpha, so alpha and the distance
rho = s* = 1 - t must come out to rounding; the Ada endgames
(Deformations/End_Games, *_vlprs_algorithm.adb) cannot be run here (no Ada
compiler), so their extrapolated winding numbers stay parity unpinned and
these closed forms pin the oracle instead."""
import numpy as np
import pytest

from oraclelib import oracle_endgame, oracle_track
from phcpack_amd.homotopy import (Homotopy, System, compile_homotopy, dense_system,
                                  total_degree_homotopy)
from phcpack_amd.tracker import (END_INFINITY, END_REGULAR, END_SINGULAR, TrackParams,
                                 classify_endpoints, points_to_planes)


def custom_homotopy(comps):
    hom = Homotopy(comps[0][1], comps[-1][1], 1.0)
    hom.components = [(np.asarray(w, np.complex128), s) for w, s in comps]
    return hom


def power_homotopy(c):
    """x^c - 1 + t: x(t) = (1 - t)^(1/c) from x = 1"""
    return custom_homotopy([([1, 0, 0], System(1, [[(1.0 + 0j, (c,)), (-1.0 + 0j, (0,))]])),
                            ([0, 1, 0], System(1, [[(1.0 + 0j, (0,))]]))])


def pole_homotopy():
    """(1 - t) x - 1: x(t) = 1/(1 - t)"""
    return custom_homotopy([([1, 0, 0], System(1, [[(1.0 + 0j, (1,)), (-1.0 + 0j, (0,))]])),
                            ([0, 1, 0], System(1, [[(-1.0 + 0j, (1,))]]))])


def at(hom, x, t, K=1, deg=8):
    prog = compile_homotopy(hom, deg)
    return oracle_endgame(prog, points_to_planes(np.array([x], np.complex128), K), K,
                          np.array([t]))


@pytest.mark.parametrize("K", [1, 2])
@pytest.mark.parametrize("c", [2, 3])
@pytest.mark.parametrize("t", [0.9, 0.999, 0.999999])
def test_branch_point_winding_number(K, c, t):
    e = at(power_homotopy(c), [(1 - t) ** (1.0 / c)], t, K)
    assert abs(e["alpha"][0] - 1.0 / c) < 1e-9
    assert abs(e["rho"][0] - (1 - t)) < 1e-9 * (1 - t)
    assert abs(e["rco"][0] - 1.0) < 1e-12
    cls, w = classify_endpoints([2], [t], e)
    assert cls[0] == END_SINGULAR and w[0] == c


@pytest.mark.parametrize("K", [1, 2])
@pytest.mark.parametrize("t", [0.9, 0.9999])
def test_pole_diverges(K, t):
    e = at(pole_homotopy(), [1.0 / (1 - t)], t, K)
    assert abs(e["alpha"][0] + 1.0) < 1e-9
    assert abs(e["rho"][0] - (1 - t)) < 1e-9 * (1 - t)
    assert abs(e["xmax"][0] - 1.0 / (1 - t)) < 1e-9 / (1 - t)
    cls, w = classify_endpoints([2], [t], e)
    assert cls[0] == END_INFINITY and w[0] == 0


def test_regular_series_terminates():
    hom = custom_homotopy([([1, 0, 0], System(1, [[(1.0 + 0j, (1,)), (-1.0 + 0j, (0,))]])),
                           ([0, 1, 0], System(1, [[(-1.0 + 0j, (0,))]]))])
    e = at(hom, [1.5], 0.5)
    assert e["alpha"][0] == np.inf and e["rho"][0] == np.inf
    cls, w = classify_endpoints([1], [1.0], e)
    assert cls[0] == END_REGULAR and w[0] == 1


def test_decoupled_dominant_component():
    """x^2 - 1 + t, y - 2 - t: y's series stops at order 1 (zero
    coefficients are passed over), x gives alpha = 1/2; J = diag(2x, 1)"""
    t = 0.99
    sq = System(2, [[(1.0 + 0j, (2, 0)), (-1.0 + 0j, (0, 0))], [(1.0 + 0j, (0, 1)), (-2.0 + 0j, (0, 0))]])
    lin = System(2, [[(1.0 + 0j, (0, 0))], [(-1.0 + 0j, (0, 0))]])
    hom = custom_homotopy([([1, 0, 0], sq), ([0, 1, 0], lin)])
    x = np.sqrt(1 - t)
    e = at(hom, [x, 2 + t], t)
    assert abs(e["alpha"][0] - 0.5) < 1e-9
    assert abs(e["rco"][0] - 2 * x) < 1e-3 * 2 * x
    assert abs(e["xmax"][0] - (2 + t)) < 1e-12


def test_classification_rules():
    est = dict(alpha=np.array([0.5, -1.0, 0.5, np.inf, 0.34, 0.5]),
               rho=np.array([0.01, 0.01, 0.3, np.inf, 0.011, 0.01]),
               rco=np.array([1e-3, 1e-3, 1e-3, 0.5, 1e-12, 1e-3]),
               xmax=np.array([1.0, 50.0, 1.0, 1.0, 1.0, 1e9]))
    status = np.array([2, 2, 2, 1, 1, 2])
    tend = np.array([0.99, 0.99, 0.99, 1.0, 0.99, 0.99])
    cls, w = classify_endpoints(status, tend, est)
    assert cls.tolist() == [END_SINGULAR, END_INFINITY, 0, END_REGULAR, END_SINGULAR, END_INFINITY]
    assert w.tolist() == [2, 0, 0, 1, 3, 0]


def test_rank_deficient_family_stalled_paths_diverge():
    """the rank-deficient family (top-degree forms of one equation a
    combination of two others) in 4 variables: the oracle's FP64 tracker
    finishes the regular paths, and the endgame estimates at the other
    paths' last points say a pole at t = 1 (diverging), as the solutions at
    infinity of the rank-deficient leading forms require"""
    rng = np.random.default_rng(3)
    hom, sols = total_degree_homotopy(dense_system(4, 2, rng, rank_deficient=1), rng)
    prog = compile_homotopy(hom, 8)
    o = oracle_track(prog, points_to_planes(sols), 1, TrackParams(tolend=1e-13))
    e = oracle_endgame(prog, o["endp"], 1, o["tend"])
    cls, w = classify_endpoints(o["status"], o["tend"], e)
    ok = o["status"] == 1
    assert ok.sum() >= 1 and (~ok).sum() >= 1
    assert np.all(cls[ok] == END_REGULAR)
    assert np.all(cls[~ok] == END_INFINITY), (o["status"][~ok], o["tend"][~ok], e["alpha"][~ok],
                                              e["rho"][~ok], e["xmax"][~ok])


def test_stopped_path_ends_at_its_last_accepted_point():
    """x^2 = 1 - t from x = 1: the path runs into the branch point at t = 1,
    the tracker stops it before (step below minstep or a singular
    Jacobian), and the end it reports is a point of the path at tend
    (x^2 = 1 - tend to the corrector's tolerance), not a rejected trial
    point; the endgame estimate there says winding number 2"""
    hom = power_homotopy(2)
    prog = compile_homotopy(hom, 8)
    o = oracle_track(prog, points_to_planes(np.array([[1.0 + 0j]])), 1, TrackParams(tolend=1e-13))
    assert o["status"][0] in (2, 4)
    x, t = o["endp"][0, 0, 0] + 1j * o["endp"][1, 0, 0], o["tend"][0]
    assert 0.99 < t < 1.0
    assert abs(x * x - (1.0 - t)) <= 1e-8 * (1.0 - t)
    e = oracle_endgame(prog, o["endp"], 1, o["tend"])
    cls, w = classify_endpoints(o["status"], o["tend"], e)
    assert cls[0] == END_SINGULAR and w[0] == 2


def test_tracking_continues_from_a_start_parameter():
    """a run stopped after 6 steps (maxsteps) and continued from its last
    point and tend reaches the endpoints of the uninterrupted run (to the
    end Newton's tolerance); tstart = 0 gives the uninterrupted run bit for
    bit"""
    rng = np.random.default_rng(11)
    hom, sols = total_degree_homotopy(dense_system(3, 2, rng), rng)
    prog = compile_homotopy(hom, 8)
    planes = points_to_planes(sols)
    full = oracle_track(prog, planes, 1, TrackParams(tolend=1e-13))
    assert np.all(full["status"] == 1)
    zero = oracle_track(prog, planes, 1, TrackParams(tolend=1e-13), tstart=np.zeros(len(sols)))
    for k in ("endp", "tend", "status", "nsteps", "work"):
        assert np.array_equal(zero[k], full[k]), k
    part = oracle_track(prog, planes, 1, TrackParams(tolend=1e-13, maxsteps=6))
    assert np.all(part["status"] == 3) and np.all(part["tend"] > 0) and np.all(part["tend"] < 1)
    rest = oracle_track(prog, part["endp"], 1, TrackParams(tolend=1e-13), tstart=part["tend"])
    assert np.all(rest["status"] == 1)
    za = full["endp"][0] + 1j * full["endp"][1]
    zb = rest["endp"][0] + 1j * rest["endp"][1]
    assert np.abs(za - zb).max() < 1e-10
