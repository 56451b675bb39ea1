"""The benchmark configurations themselves on the GPU, against the oracle:

C3  the headline system (12 dense quadratics, 4096 paths, seed 20261202):
    a random sample of 64 paths tracked bit-identically to the oracle in
    FP64 and in double-double, and all 4096 endpoints of the mixed method
    (FP64 tracking, dd/qd endpoint Newton) checked by their properties:
    residual <= 1e-24 in quad-double, 4096 distinct solutions.
C5  learned step control on dense cubics in 8 variables (6561 paths): a
    sample of paths in learned mode and in explore mode (with its step
    logs) bit-identical to the oracle.
"""
import numpy as np
import pytest

from oraclelib import OracleController, oracle_track

pytestmark = pytest.mark.gpu

SEED = 20261202


def headline_system(n=12, seed=SEED):
    from phcpack_amd.homotopy import dense_system, total_degree_homotopy
    rng = np.random.default_rng(seed)
    p = dense_system(n, 2, rng)
    hom, sols = total_degree_homotopy(p, rng)
    return p, hom, sols


@pytest.mark.parametrize("K", [1, 2])
def test_c3_sample_matches_oracle(K):
    from phcpack_amd.tracker import Tracker, TrackParams, points_to_planes
    p, hom, sols = headline_system()
    assert len(sols) == 4096
    tr = Tracker(hom, deg=8)
    rng = np.random.default_rng(64 + K)
    sample = np.sort(rng.choice(len(sols), 64, replace=False))
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28)
    print(f"C3 sample K={K} paths={len(sample)}", flush=True)
    g = tr.track(sols[sample], K=K, params=prm)
    o = oracle_track(tr.prog, points_to_planes(sols[sample]), K, prm, nthreads=8)
    for key in ("status", "nsteps", "nrej", "tend", "endp", "work", "dx"):
        assert np.array_equal(g[key], o[key]), key
    assert np.all(g["status"] == 1)


def test_c3_all_endpoints_by_properties():
    """the mixed method on all 4096 paths: FP64 tracking ([5/1] Pade
    predictor), double-double tracking from the start of the paths that FP64
    could not finish inside the interval, dd/qd endpoint Newton; every
    endpoint's residual max_i |p_i(x)| <= 1e-24 evaluated in quad-double at
    the endpoint, and the 4096 endpoints are distinct (the Bezout number of
    a generic system)"""
    from phcpack_amd.tracker import (Tracker, TrackParams, escalate_to_dd, planes_to_points,
                                     points_to_planes)
    p, hom, sols = headline_system()
    tr = Tracker(hom, deg=8)
    print("C3 all 4096 paths, mixed precision", flush=True)
    prm = TrackParams(tolend=1e-13)
    g = tr.track(sols, K=1, params=prm)
    starts = points_to_planes(sols)
    idx, esc = escalate_to_dd(tr, starts, g["status"], g["tend"], prm, endp=g["endp"])
    print(f"  FP64 statuses {np.bincount(g['status'], minlength=6).tolist()}, "
          f"escalated to dd tracking: {idx.tolist()}", flush=True)
    ok = g["status"] == 1
    assert np.all(ok | np.isin(np.arange(len(sols)), idx)), "a failure close to t = 1"
    r2 = tr.refine(g["endp"], 2, iters=3, tol=1e-28)
    x4 = np.zeros((8,) + g["endp"].shape[1:])
    x4[:2], x4[4:6] = r2["x"][:2], r2["x"][2:]
    if esc is not None:
        assert np.all(esc["status"] == 1)
        x4[:2, idx], x4[4:6, idx] = esc["endp"][:2], esc["endp"][2:]
    bad = np.nonzero((r2["status"] != 1) & ok)[0]
    if len(bad):
        r4 = tr.refine(x4[:, bad], 4, iters=4, tol=1e-56)
        assert np.all(r4["status"] == 1)
        x4[:, bad] = r4["x"]
    res = tr.refine(x4, 4, iters=0, tol=0.0)["res"]
    print(f"  max residual {res.max():.2e}, escalated to qd: {len(bad)}", flush=True)
    assert res.max() <= 1e-24
    z = planes_to_points(x4, 4)
    keys = {tuple(np.round(zi, 8)) for zi in z}
    assert len(keys) == 4096


def cubic_system(seed, n=8):
    from phcpack_amd.homotopy import dense_system, total_degree_homotopy
    rng = np.random.default_rng(seed)
    p = dense_system(n, 3, rng)
    return total_degree_homotopy(p, rng)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_c5_cubics_sample_matches_oracle(mode):
    """a held-out system of the learned-control family (seed + 110), 32
    sampled paths: plain rule (mode 0), learned (1) and explore (2, with its
    step logs), FP64 tracking as the benchmark runs it"""
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.controller import StepController, StepLogs
    from phcpack_amd.tracker import TrackBuffers, Tracker, TrackParams, points_to_planes
    hom, sols = cubic_system(SEED + 110)
    assert len(sols) == 6561
    tr = Tracker(hom, deg=8)
    rng = np.random.default_rng(5)
    starts = sols[np.sort(rng.choice(len(sols), 32, replace=False))]
    prm = TrackParams(tolend=1e-13)
    ctl = StepController(seed=3, threshold=0.5)
    print(f"C5 cubics n=8 mode={mode} paths={len(starts)}", flush=True)
    d = DeviceArray.from_host(points_to_planes(starts))
    bufs = TrackBuffers(len(starts), 8, 1)
    logs = StepLogs(len(starts), 256) if mode == 2 else None
    cs = None if mode == 0 else (ctl.struct(1) if mode == 1 else ctl.struct(2, logs, seed=17))
    tr.track_device(d, 1, len(starts), 1, bufs, prm, controller=cs)
    check(load().phc_synchronize())
    g = bufs.to_host()
    oc = None if mode == 0 else OracleController(mode, len(starts), ctl=ctl, seed=17, cap=256)
    o = oracle_track(tr.prog, points_to_planes(starts), 1, prm, controller=oc, nthreads=8 if mode == 0 else 1)
    for k in ("endp", "status", "nsteps", "nrej", "work", "tend"):
        assert np.array_equal(g[k], o[k]), k
    if mode == 2:
        gx, gc, gy = logs.to_host()
        ox, oc_, oy = oc.logs()
        assert np.array_equal(gx, ox) and np.array_equal(gc, oc_) and np.array_equal(gy, oy)
    # the Hessians of cubics are evaluated every step: some steps are set
    # by the curvature step
    print("  statuses", np.bincount(g["status"], minlength=6).tolist(),
          "pole/curv/max", g["work"][4:].sum(axis=1).tolist(), flush=True)
