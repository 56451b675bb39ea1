"""The learned controller's fallback in the oracle: a path the controller
does not finish is tracked again with the plain rule only when it stopped
at a t below retry_below (0: 0.99), so that paths that stop near t = 1 (a
singular or diverging endpoint, which the plain rule does not reach either)
are not tracked twice."""
import numpy as np

from oraclelib import OracleController, oracle_track
from phcpack_amd.controller import StepController, nparams
from phcpack_amd.homotopy import compile_homotopy, dense_system, total_degree_homotopy
from phcpack_amd.tracker import TrackParams, points_to_planes


def _runs(retry_below):
    rng = np.random.default_rng(10)
    hom, sols = total_degree_homotopy(dense_system(3, 2, rng), rng)
    prog = compile_homotopy(hom, 8)
    planes = points_to_planes(sols)
    plain = oracle_track(prog, planes, 1, TrackParams(tolend=1e-13))
    prm = TrackParams(tolend=1e-13, maxsteps=int(1.5 * plain["nsteps"].max()) + 1)
    mults = (1.0, 0.25)
    w = np.zeros(nparams(C=len(mults)), np.float32)
    w[-len(mults):] = 10.0  # every logit 10: always the quartered step
    ctl = StepController(weights=w, threshold=0.5, mults=mults, pole_cap=0.0)
    oc = OracleController(1, len(sols), ctl=ctl)
    oc.struct.retry_below = retry_below
    return plain, prm, oracle_track(prog, planes, 1, prm, controller=oc)


def test_early_stops_are_tracked_again():
    plain, prm, got = _runs(0.0)  # the default 0.99
    assert np.all(got["status"] == 1)
    assert np.array_equal(got["nsteps"], plain["nsteps"])
    assert np.array_equal(got["work"][0], plain["work"][0] + prm.maxsteps)


def test_late_stops_are_not_tracked_again():
    plain, prm, got = _runs(0.05)
    assert np.all(got["status"] == 3)  # maxsteps, the first attempt's
    assert np.all((got["tend"] >= 0.05) & (got["tend"] < 1.0))
    assert np.all(got["work"][0] == prm.maxsteps)
