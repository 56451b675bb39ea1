"""GPU parity of the controller: MLP forward/backward kernels against the
float64 numpy reference, training, and tracking in explore and learned
mode bit-identical to the oracle."""
import numpy as np
import pytest

from oraclelib import OracleController, oracle_track

pytestmark = pytest.mark.gpu


def test_mlp_forward_and_grad_match_numpy():
    from phcpack_amd.controller import MULTS, NFEAT, forward_np, grad_np, init_weights, mlp_forward, mlp_grad
    rng = np.random.default_rng(0)
    w = init_weights(rng)
    B = 301
    X = rng.standard_normal((B, NFEAT)).astype(np.float32)
    cand = rng.integers(0, len(MULTS), B).astype(np.int32)
    y = rng.integers(0, 2, B).astype(np.float32)
    O = mlp_forward(w, X)
    O_ref, _ = forward_np(w, X.astype(np.float64))
    assert np.allclose(O, O_ref, rtol=1e-4, atol=1e-4)
    g, loss = mlp_grad(w, X, cand, y)
    g_ref, loss_ref = grad_np(w, X.astype(np.float64), cand, y.astype(np.float64))
    assert abs(loss - loss_ref) < 1e-4 * max(1.0, abs(loss_ref))
    assert np.allclose(g, g_ref, rtol=1e-3, atol=1e-5 * np.abs(g_ref).max())


def test_mlp_training_reduces_loss():
    from phcpack_amd.controller import MULTS, NFEAT, StepController
    rng = np.random.default_rng(1)
    N = 4096
    X = rng.standard_normal((N, NFEAT)).astype(np.float32)
    cand = rng.integers(0, len(MULTS), N)
    # accepted when the candidate index is below a feature-dependent level
    y = (cand < 3 + 2 * np.tanh(X[:, 0])).astype(np.float32)
    ctl = StepController(seed=0)
    losses = ctl.fit(X, cand, y, epochs=20, batch=512, lr=3e-3)
    assert losses[-1] < 0.6 * losses[0]


def _system(n, degree, seed):
    from phcpack_amd.homotopy import dense_system, total_degree_homotopy
    rng = np.random.default_rng(seed)
    p = dense_system(n, degree, rng)
    hom, sols = total_degree_homotopy(p, rng)
    return hom, sols


PREDICTORS = [(0, 0), (5, 1), (4, 2)]  # (pade_numdeg, pade_dendeg): Taylor, [5/1], [4/2] Pade


@pytest.mark.parametrize("pade", PREDICTORS)
@pytest.mark.parametrize("K", [1, 2])
def test_explore_mode_matches_oracle(K, pade):
    from phcpack_amd.controller import StepController, StepLogs
    from phcpack_amd.tracker import Tracker, TrackParams, TrackBuffers, points_to_planes
    from phcpack_amd._lib import DeviceArray, check, load
    hom, sols = _system(3, 2, 9)
    tr = Tracker(hom, deg=8)
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28, pade_numdeg=pade[0], pade_dendeg=pade[1])
    ctl = StepController()
    logs = StepLogs(len(sols), 128)
    bufs = TrackBuffers(len(sols), 3, K)
    d = DeviceArray.from_host(points_to_planes(sols))
    tr.track_device(d, 1, len(sols), K, bufs, prm, controller=ctl.struct(2, logs, seed=11))
    check(load().phc_synchronize())
    g = bufs.to_host()
    oc = OracleController(2, len(sols), ctl=ctl, seed=11, cap=128)
    o = oracle_track(tr.prog, points_to_planes(sols), K, prm, controller=oc)
    for k in ("endp", "status", "nsteps", "nrej", "work"):
        assert np.array_equal(g[k], o[k]), k
    gx, gc, gy = logs.to_host()
    ox, oc_, oy = oc.logs()
    assert np.array_equal(gx, ox) and np.array_equal(gc, oc_) and np.array_equal(gy, oy)


@pytest.mark.parametrize("pade", PREDICTORS)
@pytest.mark.parametrize("K", [1, 2])
def test_learned_mode_matches_oracle(K, pade):
    from phcpack_amd.controller import StepController
    from phcpack_amd.tracker import Tracker, TrackParams, points_to_planes
    hom, sols = _system(4, 2, 10)
    tr = Tracker(hom, deg=8)
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28, pade_numdeg=pade[0], pade_dendeg=pade[1])
    ctl = StepController(seed=3, threshold=0.5)
    g = tr.track(sols, K=K, params=prm, controller=ctl.struct(1))
    oc = OracleController(1, len(sols), ctl=ctl)
    o = oracle_track(tr.prog, points_to_planes(sols), K, prm, controller=oc)
    for k in ("endp", "status", "nsteps", "nrej", "work"):
        assert np.array_equal(g[k], o[k]), k
    assert np.all(g["status"] == 1)


@pytest.mark.parametrize("K", [1, 2])
def test_unfinished_learned_paths_fall_back_to_the_plain_rule(K):
    """A controller that always picks a quarter of the pole step, with a
    step budget the plain rule meets and the quartered steps do not: every
    path fails its first attempt (maxsteps), is tracked again with the plain
    rule in the same launch and finishes; the work counters count both
    attempts; GPU and oracle agree bit for bit."""
    from phcpack_amd.controller import NFEAT, HIDDEN, StepController, nparams
    from phcpack_amd.tracker import Tracker, TrackParams, points_to_planes
    hom, sols = _system(4, 2, 10)
    tr = Tracker(hom, deg=8)
    base = TrackParams(tolend=1e-13 if K == 1 else 1e-28)
    plain = tr.track(sols, K=K, params=base)
    assert np.all(plain["status"] == 1)
    prm = TrackParams(tolend=base.tolend, maxsteps=int(1.5 * plain["nsteps"].max()) + 1)
    mults = (1.0, 0.25)
    w = np.zeros(nparams(C=len(mults)), np.float32)
    w[-len(mults):] = 10.0                      # every logit 10: the last candidate
    ctl = StepController(weights=w, threshold=0.5, mults=mults, pole_cap=0.0)
    g = tr.track(sols, K=K, params=prm, controller=ctl.struct(1))
    oc = OracleController(1, len(sols), ctl=ctl)
    o = oracle_track(tr.prog, points_to_planes(sols), K, prm, controller=oc)
    for k in ("endp", "status", "nsteps", "nrej", "work", "tend"):
        assert np.array_equal(g[k], o[k]), k
    assert np.all(g["status"] == 1)
    # the second attempt is the plain rule: same steps as plain tracking,
    # and the predictor work of the failed first attempt on top
    assert np.array_equal(g["nsteps"], plain["nsteps"])
    assert np.array_equal(g["work"][0], plain["work"][0] + prm.maxsteps)


@pytest.mark.parametrize("K", [1, 2])
def test_paths_stopped_late_are_not_tracked_again(K):
    """The same quartered-step controller with retry_below = 0.05: the paths
    stop (maxsteps) beyond t = 0.05, so none is tracked again: every status
    is maxsteps, the predictor work is the first attempt's alone; GPU and
    oracle agree bit for bit."""
    from phcpack_amd.controller import StepController, nparams
    from phcpack_amd.tracker import Tracker, TrackParams, points_to_planes
    hom, sols = _system(4, 2, 10)
    tr = Tracker(hom, deg=8)
    base = TrackParams(tolend=1e-13 if K == 1 else 1e-28)
    plain = tr.track(sols, K=K, params=base)
    prm = TrackParams(tolend=base.tolend, maxsteps=int(1.5 * plain["nsteps"].max()) + 1)
    mults = (1.0, 0.25)
    w = np.zeros(nparams(C=len(mults)), np.float32)
    w[-len(mults):] = 10.0
    ctl = StepController(weights=w, threshold=0.5, mults=mults, pole_cap=0.0)
    cs = ctl.struct(1)
    cs.retry_below = 0.05
    g = tr.track(sols, K=K, params=prm, controller=cs)
    oc = OracleController(1, len(sols), ctl=ctl)
    oc.struct.retry_below = 0.05
    o = oracle_track(tr.prog, points_to_planes(sols), K, prm, controller=oc)
    for k in ("endp", "status", "nsteps", "nrej", "work", "tend"):
        assert np.array_equal(g[k], o[k]), k
    assert np.all(g["status"] != 1) and np.all(g["tend"] >= 0.05) and np.all(g["tend"] < 1.0)
    assert np.all(g["work"][0] == prm.maxsteps)


def test_trained_controller_gains_over_plain_rule():
    """A/B on held-out systems: the controller trained (on the GPU, by the
    library's own kernels) on explore logs of other systems of the family
    needs fewer predictor steps than the plain pole-step rule and loses no
    path that the plain rule tracks."""
    from phcpack_amd.controller import train_controller
    from phcpack_amd.tracker import Tracker, TrackParams
    print("controller A/B", flush=True)
    prm = TrackParams(tolend=1e-13)
    trs, sts = [], []
    for seed in range(31, 37):
        hom, sols = _system(6, 2, seed)
        trs.append(Tracker(hom, deg=8))
        sts.append(sols)
    ctl, info = train_controller(trs, sts, prm, K=1, seed=5, epochs=30, batch=512,
                                 threshold=0.9)
    assert info["loss_last"] < info["loss_first"]
    steps_plain, steps_learned = 0, 0
    for seed in range(90, 94):            # held out
        hom, sols = _system(6, 2, seed)
        tr = Tracker(hom, deg=8)
        a = tr.track(sols, K=1, params=prm)
        b = tr.track(sols, K=1, params=prm, controller=ctl.struct(1))
        ok_a, ok_b = a["status"] == 1, b["status"] == 1
        assert ok_b.sum() >= ok_a.sum(), seed
        steps_plain += int(a["work"][0].sum())
        steps_learned += int(b["work"][0].sum())
    print(f"predictor steps plain {steps_plain} learned {steps_learned}", flush=True)
    assert steps_learned < 0.85 * steps_plain


@pytest.mark.parametrize("K", [1, 2])
def test_order_candidates_match_oracle(K):
    """joint (multiplier, series degree) candidates, explore and learned
    mode: the kernel's paths and step logs equal the oracle's bit for bit"""
    from phcpack_amd.controller import StepController, StepLogs, order_candidates
    from phcpack_amd.tracker import Tracker, TrackParams, TrackBuffers, points_to_planes
    from phcpack_amd._lib import DeviceArray, check, load
    hom, sols = _system(4, 2, 12)
    tr = Tracker(hom, deg=8)
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28)
    mults, degrees = order_candidates(low=5)
    ctl = StepController(seed=4, threshold=0.5, mults=mults, degrees=degrees, deg0=5)
    logs = StepLogs(len(sols), 256)
    bufs = TrackBuffers(len(sols), 4, K)
    d = DeviceArray.from_host(points_to_planes(sols))
    tr.track_device(d, 1, len(sols), K, bufs, prm, controller=ctl.struct(2, logs, seed=13))
    check(load().phc_synchronize())
    g = bufs.to_host()
    oc = OracleController(2, len(sols), ctl=ctl, seed=13, cap=256)
    o = oracle_track(tr.prog, points_to_planes(sols), K, prm, controller=oc)
    for k in ("endp", "status", "nsteps", "nrej", "work", "tend"):
        assert np.array_equal(g[k], o[k]), k
    gx, gc, gy = logs.to_host()
    ox, oc_, oy = oc.logs()
    assert np.array_equal(gx, ox) and np.array_equal(gc, oc_) and np.array_equal(gy, oy)
    assert len(set(gc.tolist())) > 6
    g = tr.track(sols, K=K, params=prm, controller=ctl.struct(1))
    o = oracle_track(tr.prog, points_to_planes(sols), K, prm,
                     controller=OracleController(1, len(sols), ctl=ctl))
    for k in ("endp", "status", "nsteps", "nrej", "work", "tend"):
        assert np.array_equal(g[k], o[k]), k


def test_low_degree_candidate_is_the_rule_at_that_degree():
    """one candidate (1, degree 6) after deg0 = 6 on a degree-8 program
    tracks every path as the plain rule on the degree-6 program with the
    [3/1] predictor, bit for bit, on the GPU"""
    from phcpack_amd.controller import StepController, StepLogs
    from phcpack_amd.tracker import Tracker, TrackParams, TrackBuffers, points_to_planes
    from phcpack_amd._lib import DeviceArray, check, load
    hom, sols = _system(4, 2, 14)
    tr8, tr6 = Tracker(hom, deg=8), Tracker(hom, deg=6)
    ctl = StepController(mults=(1.0,), degrees=(6,), deg0=6, pole_cap=0.0)
    logs = StepLogs(len(sols), 8)
    bufs = TrackBuffers(len(sols), 4, 2)
    d = DeviceArray.from_host(points_to_planes(sols))
    tr8.track_device(d, 1, len(sols), 2, bufs, TrackParams(), controller=ctl.struct(2, logs, seed=1))
    check(load().phc_synchronize())
    g = bufs.to_host()
    ref = tr6.track(sols, K=2, params=TrackParams(pade_numdeg=3, pade_dendeg=1))
    for k in ("endp", "status", "nsteps", "nrej", "work", "tend"):
        assert np.array_equal(g[k], ref[k]), k
