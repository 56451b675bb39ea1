"""Learned step-size controller: the shared log2 approximation, the numpy
reference of the MLP gradient (pinned by central differences), and the
oracle's explore-mode logging."""
import math

import numpy as np
import pytest

from oraclelib import OracleController, _track_sigs, oracle_track
from phcpack_amd.controller import MULTS, NFEAT, StepController, forward_np, grad_np, init_weights
from phcpack_amd.homotopy import compile_homotopy, dense_system, total_degree_homotopy
from phcpack_amd.tracker import TrackParams, points_to_planes


def test_lg2_accuracy_and_special_values():
    o = _track_sigs()
    for x in np.concatenate([10.0 ** np.linspace(-300, 300, 601), [1.0, 2.0, 0.75, 3.0]]):
        assert abs(o.otr_lg2(float(x)) - math.log2(x)) < 2e-4
    assert o.otr_lg2(0.0) == -1100.0
    assert o.otr_lg2(-1.0) == -1100.0
    assert o.otr_lg2(float("inf")) == 1100.0


def test_numpy_gradient_by_central_differences():
    rng = np.random.default_rng(0)
    w = init_weights(rng).astype(np.float64)
    X = rng.standard_normal((7, NFEAT))
    cand = rng.integers(0, len(MULTS), 7)
    y = rng.integers(0, 2, 7).astype(np.float64)
    g, _ = grad_np(w, X, cand, y)
    idx = rng.choice(len(w), 40, replace=False)
    for i in idx:
        e = np.zeros_like(w)
        e[i] = 1e-6
        fp = grad_np(w + e, X, cand, y)[1]
        fm = grad_np(w - e, X, cand, y)[1]
        assert abs((fp - fm) / 2e-6 - g[i]) < 1e-6 + 1e-4 * abs(g[i])


def test_oracle_explore_logs_every_step():
    rng = np.random.default_rng(3)
    p = dense_system(3, 2, rng)
    hom, sols = total_degree_homotopy(p, rng)
    prog = compile_homotopy(hom, 8)
    oc = OracleController(2, len(sols), seed=5)
    out = oracle_track(prog, points_to_planes(sols), 1, TrackParams(tolend=1e-13), controller=oc)
    assert np.all(out["status"] == 1)
    # one record per predictor series (every step attempt)
    assert np.array_equal(oc.n, out["work"][0])
    X, c, y = oc.logs()
    assert set(np.unique(c)) <= set(range(len(MULTS)))
    # 0 = first attempt rejected, else the corrector's Newton iterations
    assert set(np.unique(y)) <= set(range(0, TrackParams().corsteps + 1))
    assert np.any(y >= 1)
    assert np.all(np.isfinite(X))


def test_oracle_learned_mode_runs_and_never_loses_paths():
    """With any weights the corrector and feedback judge every step: paths
    still converge (here the network is untrained and random)."""
    rng = np.random.default_rng(4)
    p = dense_system(3, 2, rng)
    hom, sols = total_degree_homotopy(p, rng)
    prog = compile_homotopy(hom, 8)
    ctl = StepController(seed=1, threshold=0.5)
    oc = OracleController(1, len(sols), ctl=ctl)
    out = oracle_track(prog, points_to_planes(sols), 2, TrackParams(), controller=oc)
    assert np.all(out["status"] == 1)
    ref = oracle_track(prog, points_to_planes(sols), 2, TrackParams())
    # same endpoints up to dd rounding (the path, not its steps, fixes them)
    diff = np.abs(out["endp"][:2].sum(0) - ref["endp"][:2].sum(0)).max()
    assert diff < 1e-20


def test_forward_np_matches_split_layout():
    rng = np.random.default_rng(2)
    w = init_weights(rng)
    X = rng.standard_normal((3, NFEAT))
    O, _ = forward_np(w, X)
    assert O.shape == (3, len(MULTS))
