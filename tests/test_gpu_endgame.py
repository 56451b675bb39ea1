"""The endgame kernel (phc_endgame) against the oracle's otr_endgame, bit for
bit: the closed forms of tests/test_oracle_endgame.py (branch points of
winding number 2 and 3, a pole, a regular path) and the last points of the
stalled paths of a rank-deficient system, in FP64 and double-double; then
the classification of a whole rank-deficient run on the GPU."""
import numpy as np
import pytest

from oraclelib import oracle_endgame
from phcpack_amd.tracker import (END_INFINITY, END_REGULAR, Tracker, TrackParams,
                                 classify_endpoints, points_to_planes)
from test_oracle_endgame import pole_homotopy, power_homotopy

pytestmark = pytest.mark.gpu


def same(g, o):
    for k in ("alpha", "rho", "rco", "xmax"):
        assert np.array_equal(g[k], o[k]), (k, g[k], o[k])


@pytest.mark.parametrize("K", [1, 2])
def test_closed_forms_match_oracle(K):
    t = np.array([0.9, 0.99, 0.999999])
    for hom, x in ((power_homotopy(2), np.sqrt(1 - t)), (power_homotopy(3), np.cbrt(1 - t)),
                   (pole_homotopy(), 1 / (1 - t))):
        tr = Tracker(hom, deg=8)
        planes = points_to_planes(x[:, None].astype(np.complex128), K)
        print(f"endgame K={K} closed form", flush=True)
        g = tr.endgame(planes, t, K=K)
        same(g, oracle_endgame(tr.prog, planes, K, t))


@pytest.mark.parametrize("K", [1, 2])
def test_rank_deficient_stalled_points_match_oracle(K):
    from phcpack_amd.homotopy import dense_system, total_degree_homotopy
    rng = np.random.default_rng(4)
    hom, sols = total_degree_homotopy(dense_system(6, 2, rng, rank_deficient=2), rng)
    tr = Tracker(hom, deg=8)
    print(f"endgame K={K} rank-deficient n=6", flush=True)
    o = tr.track(sols, K=1, params=TrackParams(tolend=1e-13))
    g = tr.endgame(o["endp"], o["tend"], K=K)
    same(g, oracle_endgame(tr.prog, o["endp"], K, o["tend"]))
    cls, _ = classify_endpoints(o["status"], o["tend"], g)
    ok = o["status"] == 1
    assert np.all(cls[ok] == END_REGULAR) and np.all(cls[~ok] == END_INFINITY)
    assert (~ok).sum() > 0
