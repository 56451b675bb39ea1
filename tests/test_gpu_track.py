"""GPU tracker and endpoint refinement against the C oracle.  Kernel and
oracle evaluate the same compiled program in the same operation order, so
endpoints, step counts and statuses must agree exactly."""
import numpy as np
import pytest

from oraclelib import oracle_refine, oracle_track

pytestmark = pytest.mark.gpu


def make(n, degree, seed, rank_deficient=0, k=1):
    from phcpack_amd.homotopy import dense_system, total_degree_homotopy
    rng = np.random.default_rng(seed)
    p = dense_system(n, degree, rng, rank_deficient)
    hom, sols = total_degree_homotopy(p, rng, k=k)
    return p, hom, sols


@pytest.mark.parametrize("K", [1, 2])
@pytest.mark.parametrize("n,degree,seed", [(3, 2, 1), (4, 2, 2), (3, 3, 3)])
def test_track_matches_oracle(K, n, degree, seed):
    from phcpack_amd.tracker import Tracker, TrackParams, points_to_planes
    p, hom, sols = make(n, degree, seed)
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28)
    tr = Tracker(hom, deg=8)
    print(f"track K={K} n={n} degree={degree} seed={seed} paths={len(sols)}", flush=True)
    g = tr.track(sols, K=K, params=prm)
    o = oracle_track(tr.prog, points_to_planes(sols), K, prm)
    assert np.array_equal(g["status"], o["status"])
    assert np.array_equal(g["nsteps"], o["nsteps"])
    assert np.array_equal(g["nrej"], o["nrej"])
    assert np.array_equal(g["tend"], o["tend"])
    assert np.array_equal(g["endp"], o["endp"])
    assert np.array_equal(g["work"], o["work"])
    assert np.all(g["status"] == 1)


@pytest.mark.parametrize("K", [1, 2])
@pytest.mark.parametrize("n,degree,seed,npaths", [(13, 2, 4, 12), (16, 1, 5, 1), (17, 1, 6, 1)])
def test_track_matches_oracle_wider(K, n, degree, seed, npaths):
    """Dimensions around the 16-column limit of the one-phase-per-pivot
    Gauss-Jordan inverse (rows split unevenly over the lane groups at 13,
    four rows per group at 16, the phased inverse at 17)."""
    from phcpack_amd.tracker import Tracker, TrackParams, points_to_planes
    p, hom, sols = make(n, degree, seed)
    sols = sols[:npaths]
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28)
    tr = Tracker(hom, deg=8)
    print(f"track K={K} n={n} degree={degree} seed={seed} paths={len(sols)}", flush=True)
    g = tr.track(sols, K=K, params=prm)
    o = oracle_track(tr.prog, points_to_planes(sols), K, prm)
    for key in ("status", "nsteps", "nrej", "tend", "endp", "work"):
        assert np.array_equal(g[key], o[key]), key
    assert np.all(g["status"] == 1)


@pytest.mark.parametrize("K,waves", [(1, 12), (1, 8), (2, 8)])
@pytest.mark.parametrize("n,degree,seed,num,den", [(3, 2, 1, 5, 1), (4, 2, 2, 4, 1),
                                                   (3, 3, 3, 2, 1), (4, 2, 2, 4, 0),
                                                   (3, 2, 1, 6, 0), (4, 2, 2, 4, 2),
                                                   (3, 3, 3, 1, 2), (3, 3, 3, 3, 2)])
def test_track_pade_matches_oracle(K, waves, n, degree, seed, num, den):
    """The [num/den] Pade predictor (den 1, 2) and the Taylor polynomial of
    lower degree (den 0): kernel and oracle agree bit for bit, in the 12-
    and 8-wavefront variants of the FP64 kernel and in double-double."""
    from phcpack_amd.tracker import Tracker, TrackParams, points_to_planes
    p, hom, sols = make(n, degree, seed)
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28, pade_numdeg=num, pade_dendeg=den)
    tr = Tracker(hom, deg=8)
    print(f"track pade [{num}/{den}] K={K} waves={waves} n={n} degree={degree} seed={seed} "
          f"paths={len(sols)}", flush=True)
    g = tr.track(sols, K=K, params=prm, waves=waves)
    o = oracle_track(tr.prog, points_to_planes(sols), K, prm)
    for key in ("status", "nsteps", "nrej", "tend", "endp", "work"):
        assert np.array_equal(g[key], o[key]), key
    assert np.all(g["status"] == 1), np.bincount(g["status"], minlength=6)


@pytest.mark.parametrize("K", [1, 2])
def test_track_pade_rational_path(K):
    """x(t) = (a + b t)/(r - t): the [5/1] predictor is exact, so every
    corrector converges at its first Newton iteration (plus the extra one
    unless the update vanishes); kernel and oracle agree."""
    from phcpack_amd.homotopy import Homotopy, System
    from phcpack_amd.tracker import Tracker, TrackParams, planes_to_points, points_to_planes
    a, b, r = 0.6 + 0.2j, -0.3 + 0.9j, 1.25
    hom = Homotopy(System(1, [[(1.0 + 0j, (1,))]]), System(1, [[(-1.0 + 0j, (0,))]]), 1.0)
    hom.components = [(np.array([r, -1.0, 0.0], np.complex128), hom.target),
                      (np.array([a, b, 0.0], np.complex128), hom.start)]
    tr = Tracker(hom, deg=8)
    prm = TrackParams(pade_numdeg=5, pade_dendeg=1, endsteps=0)
    sols = np.array([[a / r]])
    print(f"track pade rational path K={K}", flush=True)
    g = tr.track(sols, K=K, params=prm)
    o = oracle_track(tr.prog, points_to_planes(sols), K, prm)
    for key in ("status", "nsteps", "nrej", "tend", "endp", "work"):
        assert np.array_equal(g[key], o[key]), key
    assert g["status"][0] == 1 and g["work"][3, 0] <= g["work"][1, 0] <= 2 * g["work"][3, 0]
    want = (a + b) / (r - 1)
    assert abs(planes_to_points(g["endp"], K)[0, 0] - want) < 1e-13 * abs(want)


def test_track_pade_degrees_refused():
    """denominator degrees other than 0, 1 and 2, and degrees that leave
    the error estimate's coefficient L+M+2 beyond the series, are refused"""
    from phcpack_amd._lib import PhcError
    from phcpack_amd.tracker import Tracker, TrackParams
    p, hom, sols = make(3, 2, 7)
    tr = Tracker(hom, deg=8)
    for num, den in ((5, 3), (7, 2), (8, 1), (0, 1), (0, 2), (9, 0), (6, 1), (5, 2), (7, 0)):
        with pytest.raises(PhcError, match="pade"):
            tr.track(sols, K=1, params=TrackParams(pade_numdeg=num, pade_dendeg=den))


@pytest.mark.parametrize("waves", [1, 3, 8])
def test_track_wave_counts_agree(waves):
    """Wavefronts of one workgroup track their paths independently: the
    result does not depend on how many share a workgroup."""
    from phcpack_amd.tracker import Tracker, TrackParams
    p, hom, sols = make(4, 2, 5)
    tr = Tracker(hom, deg=8)
    prm = TrackParams()
    a = tr.track(sols, K=2, params=prm, waves=8)
    b = tr.track(sols, K=2, params=prm, waves=waves)
    assert np.array_equal(a["endp"], b["endp"])


@pytest.mark.parametrize("K", [2, 4, pytest.param(8, marks=pytest.mark.skip(
    reason="octo-double device refinement is switched off: its kernel (non-inlined helpers "
           "reaching LDS through FLAT instructions) stalled a device run; cause not established"))])
def test_refine_matches_oracle(K):
    from phcpack_amd.tracker import Tracker, TrackParams
    p, hom, sols = make(4, 2, 7)
    tr = Tracker(hom, deg=8)
    g = tr.track(sols, K=1, params=TrackParams(tolend=1e-13))
    x = g["endp"]
    r = tr.refine(x, K, iters=3, tol=1e-60)
    o = oracle_refine(tr.prog0, x, K, iters=3, tol=1e-60)
    for key in ("x", "dx", "res", "status"):
        assert np.array_equal(r[key], o[key]), key
    assert r["res"].max() < {2: 1e-28, 4: 1e-55, 8: 1e-100}[K]


def test_refine_octo_double_refused():
    """K = 8 on the device is refused with an error, never silently degraded."""
    from phcpack_amd._lib import PhcError
    from phcpack_amd.tracker import Tracker, TrackParams
    p, hom, sols = make(3, 2, 7)
    tr = Tracker(hom, deg=8)
    g = tr.track(sols, K=1, params=TrackParams(tolend=1e-13))
    with pytest.raises(PhcError, match="octo-double"):
        tr.refine(g["endp"], 8, iters=1, tol=1e-60)


def test_decoupled_closed_form():
    """x_i^2 = a_i: every endpoint is a sign pattern of sqrt(a)."""
    import itertools
    from phcpack_amd.homotopy import System, total_degree_homotopy
    from phcpack_amd.tracker import Tracker, TrackParams, planes_to_points
    rng = np.random.default_rng(11)
    n = 4
    a = np.exp(2j * np.pi * rng.uniform(size=n)) * rng.uniform(0.5, 2.0, n)
    eqs = []
    for i in range(n):
        e = [0] * n
        e[i] = 2
        eqs.append([(1.0 + 0j, tuple(e)), (complex(-a[i]), (0,) * n)])
    hom, sols = total_degree_homotopy(System(n, eqs), rng)
    g = Tracker(hom, deg=8).track(sols, K=2, params=TrackParams())
    x = planes_to_points(g["endp"], 2)
    want = np.array([np.array(s) * np.sqrt(a) for s in itertools.product([1, -1], repeat=n)])
    for xi in x:
        assert np.min(np.abs(want - xi).max(axis=1)) < 1e-14
    # all sign patterns reached
    got = {tuple(np.sign(np.round((xi / np.sqrt(a)).real)).astype(int)) for xi in x}
    assert len(got) == 2 ** n


@pytest.mark.parametrize("K", [1, 2])
def test_rank_deficient_n14_matches_oracle(K):
    """The dimension-14 family whose last 2 top-degree forms are combinations
    of two others (rank-deficient leading forms): a sample of 96 of its 16384
    paths, the ones that diverge or stall near t = 1 included, tracked
    bit-identically to the oracle; the program at n = 14 still leaves LDS for
    8 FP64 or 4 dd wavefronts per workgroup."""
    from phcpack_amd.tracker import Tracker, TrackParams, points_to_planes
    p, hom, sols = make(14, 2, 20261202, rank_deficient=2)
    tr = Tracker(hom, deg=8)
    prog_b, path_b = tr.dev.lds_bytes(K)
    assert prog_b + (8 if K == 1 else 4) * path_b <= 160 * 1024
    rng = np.random.default_rng(4)
    sample = np.sort(rng.choice(len(sols), 96, replace=False))
    starts = sols[sample]
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28)
    print(f"rank-deficient n=14 K={K} sample={len(sample)}", flush=True)
    g = tr.track(starts, K=K, params=prm)
    o = oracle_track(tr.prog, points_to_planes(starts), K, prm, nthreads=8)
    for k in ("status", "nsteps", "nrej", "tend", "endp", "work"):
        assert np.array_equal(g[k], o[k]), k
    print("  statuses", np.bincount(g["status"], minlength=6).tolist(), flush=True)
    # about a quarter of the paths reach regular solutions, the rest do not finish
    assert 0 < np.sum(g["status"] == 1) < len(sample)


@pytest.mark.parametrize("K", [2, 4])
def test_refine_only_finished_paths(K):
    """refinement with the tracker's statuses as mask: finished paths get
    the same result as without the mask, the others are copied through with
    their tracker status and infinite residual"""
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.tracker import TrackBuffers, Tracker, TrackParams, points_to_planes
    p, hom, sols = make(6, 2, 3, rank_deficient=1)
    tr = Tracker(hom, deg=8)
    npaths, n = len(sols), hom.n
    d_start = DeviceArray.from_host(points_to_planes(sols))
    b = TrackBuffers(npaths, n, 1)
    tr.track_device(d_start, 1, npaths, 1, b, TrackParams(tolend=1e-13))
    check(load().phc_synchronize())
    tst = b.status.to_host()
    assert 0 < np.sum(tst == 1) < npaths

    def run(only):
        out = DeviceArray((2 * K, npaths, n))
        dx, res = DeviceArray((npaths,)), DeviceArray((npaths,))
        st = DeviceArray((npaths,), np.int32)
        tr.refine_device(b.endp, 1, npaths, K, out, dx, res, st, iters=3, tol=1e-50, only=only)
        check(load().phc_synchronize())
        return out.to_host(), dx.to_host(), res.to_host(), st.to_host()

    x0, dx0, r0, s0 = run(None)
    x1, dx1, r1, s1 = run(b.status)
    fin = tst == 1
    assert np.array_equal(x1[:, fin], x0[:, fin]) and np.array_equal(r1[fin], r0[fin])
    assert np.array_equal(s1[fin], s0[fin]) and np.array_equal(dx1[fin], dx0[fin])
    e = b.endp.to_host()
    assert np.array_equal(x1[0, ~fin], e[0, ~fin]) and np.array_equal(x1[K, ~fin], e[1, ~fin])
    assert np.all(x1[1:K, ~fin] == 0) and np.all(np.isinf(r1[~fin]))
    assert np.array_equal(s1[~fin], tst[~fin])


def test_autotune_picks_a_valid_launch_shape():
    """the launch-shape tuner times every candidate and keeps the fastest;
    results are the same whatever it picks"""
    from phcpack_amd._lib import DeviceArray
    from phcpack_amd.autotune import tune_waves
    from phcpack_amd.tracker import TrackBuffers, Tracker, TrackParams, points_to_planes
    p, hom, sols = make(5, 2, 11)
    tr = Tracker(hom, deg=8)
    prm = TrackParams(tolend=1e-13)
    d = DeviceArray.from_host(points_to_planes(sols))
    bufs = TrackBuffers(len(sols), 5, 1)
    best, table = tune_waves(tr, d, 1, len(sols), 1, bufs, prm, candidates=(4, 8, 12))
    assert best in (4, 8, 12) and set(table) == {4, 8, 12} and tr.waves[1] == best
    a = tr.track(sols, K=1, params=prm)
    b = tr.track(sols, K=1, params=prm, waves=4)
    assert np.array_equal(a["endp"], b["endp"])


def test_mixed_endpoints_checked_by_octo_double_oracle():
    """FP64 tracking, dd endpoint Newton and the quad-double residual on the
    GPU, then the octo-double residual of the same endpoints on the CPU
    (oracle): every finished endpoint has residual <= 1e-24 in octo-double,
    and the GPU's quad-double residual agrees with it far below that"""
    from phcpack_amd.tracker import Tracker, TrackParams
    p, hom, sols = make(5, 2, 13)
    tr = Tracker(hom, deg=8)
    g = tr.track(sols, K=1, params=TrackParams(tolend=1e-13))
    r2 = tr.refine(g["endp"], 2, iters=3, tol=1e-28)
    ok = (g["status"] == 1) & (r2["status"] == 1)
    assert ok.sum() == len(sols)
    x4 = np.zeros((8,) + r2["x"].shape[1:])
    x4[:2], x4[4:6] = r2["x"][:2], r2["x"][2:]
    res_qd = tr.refine(x4, 4, iters=0, tol=0.0)["res"]
    res_od = oracle_refine(tr.prog0, r2["x"], 8, iters=0, tol=0.0)["res"]
    print(f"  max od residual {res_od.max():.3e}, max |qd - od| {np.abs(res_qd - res_od).max():.3e}",
          flush=True)
    assert res_od.max() <= 1e-24
    assert np.abs(res_qd - res_od).max() <= 1e-40


@pytest.mark.parametrize("K", [1, 2])
def test_track_from_start_parameters_matches_oracle(K):
    """phc_track_from: paths continued from the last points of a run stopped
    after 6 steps, each at its own tend, bit-identical to the oracle"""
    from phcpack_amd.tracker import Tracker, TrackParams
    p, hom, sols = make(3, 2, 11)
    tr = Tracker(hom, deg=8)
    print(f"track from tstart K={K}", flush=True)
    part = tr.track(sols, K=1, params=TrackParams(tolend=1e-13, maxsteps=6))
    assert np.all(part["status"] == 3)
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28)
    g = tr.track(part["endp"], K=K, params=prm, tstart=part["tend"])
    o = oracle_track(tr.prog, part["endp"], K, prm, tstart=part["tend"])
    for key in ("status", "nsteps", "nrej", "tend", "endp", "work", "dx"):
        assert np.array_equal(g[key], o[key]), key
    assert np.all(g["status"] == 1)
