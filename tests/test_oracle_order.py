"""The predictor-order output of the controller in the oracle: a step of
series degree d (the controller's choice after deg0 coefficients) must be
the reference's step for a series of degree d.  Pinned by equalities with
the plain rule, which the oracle's own tests pin to the reference's step
rule:
  * one candidate (multiplier 1, degree d), deg0 = d: the same path, bit for
    bit, as the plain rule on the program compiled for degree d with the
    Pade numerator lowered by the same amount ([5/1] -> [3/1] for 8 -> 6);
    a wrong Fabry index, Pade degree or curvature root fails it;
  * one candidate of the full degree after deg0 < deg: the series resumed
    after the choice equals the series computed at once (the plain rule);
  * mixed candidates: every step's degree is its candidate's."""
import numpy as np
import pytest

from oraclelib import OracleController, oracle_track, oracle_track_trace
from phcpack_amd.controller import StepController, order_candidates
from phcpack_amd.homotopy import compile_homotopy, dense_system, total_degree_homotopy
from phcpack_amd.tracker import TrackParams, points_to_planes

FIELDS = ("endp", "tend", "status", "nsteps", "nrej", "dx", "work")


def _system(n, degree, seed):
    rng = np.random.default_rng(seed)
    hom, sols = total_degree_homotopy(dense_system(n, degree, rng), rng)
    return hom, sols


def _explore(ctl, npaths):
    return OracleController(2, npaths, ctl=ctl, seed=5, cap=512)


@pytest.mark.parametrize("pade,low", [((5, 1), (3, 1)), ((0, 0), (0, 0)), ((4, 2), (2, 2))])
@pytest.mark.parametrize("K", [1, 2])
def test_low_degree_candidate_is_the_rule_at_that_degree(K, pade, low):
    hom, sols = _system(3, 2, 21)
    tol = 1e-13 if K == 1 else 1e-28
    prog8, prog6 = compile_homotopy(hom, 8), compile_homotopy(hom, 6)
    planes = points_to_planes(sols)
    ctl = StepController(mults=(1.0,), degrees=(6,), deg0=6, pole_cap=0.0)
    got = oracle_track(prog8, planes, K, TrackParams(tolend=tol, pade_numdeg=pade[0],
                                                     pade_dendeg=pade[1]),
                       controller=_explore(ctl, len(sols)))
    ref = oracle_track(prog6, planes, K, TrackParams(tolend=tol, pade_numdeg=low[0],
                                                     pade_dendeg=low[1]))
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), f
    assert np.all(ref["status"] == 1)


@pytest.mark.parametrize("K", [1, 2])
def test_resumed_series_is_the_series_at_once(K):
    hom, sols = _system(4, 2, 22)
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28)
    prog = compile_homotopy(hom, 8)
    planes = points_to_planes(sols)
    for d0 in (3, 5, 7):
        ctl = StepController(mults=(1.0,), degrees=(0,), deg0=d0, pole_cap=0.0)
        got = oracle_track(prog, planes, K, prm, controller=_explore(ctl, len(sols)))
        ref = oracle_track(prog, planes, K, prm)
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (d0, f)


def test_every_step_takes_its_candidates_degree():
    hom, sols = _system(4, 2, 23)
    prog = compile_homotopy(hom, 8)
    mults, degrees = order_candidates(low=5)
    ctl = StepController(mults=mults, degrees=degrees, deg0=5)
    seen = set()
    for path in range(4):
        _, steps = oracle_track_trace(prog, points_to_planes(sols), 1, TrackParams(tolend=1e-13),
                                      path=path, controller=_explore(ctl, len(sols)))
        for s in steps:
            c = int(s["cand"])
            want = degrees[c] if degrees[c] else 8
            assert int(s["degree"]) == want
            seen.add(want)
    assert seen == {5, 8}
