#!/usr/bin/env python3
"""Throughput of the batched path tracker on one MI355X.

Metric: paths/s whose endpoint residual max_i |p_i(x)| is <= 1e-24, for the
learned + mixed-precision tracker (FP64 tracking with the learned step
controller, double-double Newton at the endpoints, quad-double Newton for
endpoints that do not converge in double-double), beside the plain
double-double tracker and the plain-rule mixed tracker on the same system.
The controller is trained before the timed region on explore-mode step logs
of two other systems of the same family (seeds 20261203, 20261204) and its
acceptance threshold is tuned on a third (20261205); the evaluation system
is never seen in training.  The residual of every endpoint is evaluated at
the exact endpoint in quad-double arithmetic on the GPU (rounding ~1e-60,
far below the 1e-24 threshold).

Workload: total-degree homotopy gamma*(1-t)*q + t*p with p = 12 dense
quadratics in 12 variables (all 91 monomials of degree <= 2, complex
coefficients on the unit circle), q_i = x_i^2 - c_i, 4096 paths, power
series of degree 8 at every step, seed 20261202.  Every method (and the CPU
baseline) follows the reference's step rule (pole step, Hesse-Pade
curvature step, mixed-residual feedback and corrector) and predicts the
next point with the [5/1] Pade approximant of the series per component,
the reference's default for a degree-8 series; --predictor pade4/2 or
taylor6 choose another.  One step = tracking all 4096 paths from t = 0 to
t = 1 plus the endpoint refinement, inputs resident on the device.  The
line also carries the rank-deficient family (14 quadratics, 16384 paths)
with the same controller, and the CPU baseline on one core and on all
host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-learned]
                      [--no-rankdef] [--predictor pade5/1|pade4/2|taylor6|...]
                      [--workload ...]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "paths/sec at ≤1e-24 endpoint residual, learned+mixed vs plain dd on 1 GPU"
SEED = 20261202
RES_TOL = 1e-24
FP64_PEAK_TFLOPS = 78.6
TRAIN_SEEDS = (SEED + 1, SEED + 2)   # controller training systems (same family)
TUNE_SEED = SEED + 3                 # threshold tuning system


class Events:
    """HIP events on the stream the kernels run on (the default stream)."""

    def __init__(self, lib):
        self.lib = lib
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        lib.phc_event_create(ctypes.byref(self.a))
        lib.phc_event_create(ctypes.byref(self.b))

    def start(self):
        self.lib.phc_event_record(self.a, None)

    def stop_ms(self):
        self.lib.phc_event_record(self.b, None)
        ms = ctypes.c_float()
        self.lib.phc_event_elapsed_ms(self.a, self.b, ctypes.byref(ms))
        return float(ms.value)


def build_workload(n=12, seed=SEED):
    from phcpack_amd.homotopy import dense_system, total_degree_homotopy
    rng = np.random.default_rng(seed)
    p = dense_system(n, 2, rng)
    hom, sols = total_degree_homotopy(p, rng)
    return p, hom, sols


def run_plain(tr, d_start, npaths, bufs, ev, prm):
    ev.start()
    tr.track_device(d_start, 1, npaths, 2, bufs, prm)
    return ev.stop_ms()


def run_mixed(tr, d_start, starts, npaths, bufs, ref, ev, prm, lib, controller=None):
    """FP64 tracking (plain pole-step rule, or the learned controller when
    given: a path it does not finish is tracked again with the plain rule
    inside the same kernel), then dd Newton at the endpoints of the paths the
    tracker finished (the others are passed over on the device).  Paths that
    FP64 tracking could not finish inside the interval (a step below minstep
    at t < 0.99, say) are tracked again in double-double from their start
    (precision escalation).  Finished paths whose dd Newton does not converge
    are refined in quad-double; with the learned controller, paths that
    still fail are tracked again with the plain rule (a path the controller
    lost is recovered, not dropped).
    Returns (tracking ms, endpoint ms, qd-escalated, re-tracked, dd-escalated)."""
    from phcpack_amd._lib import check
    from phcpack_amd.tracker import escalate_to_dd
    ev.start()
    tr.track_device(d_start, 1, npaths, 1, bufs, prm, controller=controller)
    t_track = ev.stop_ms()
    ev.start()
    tr.refine_device(bufs.endp, 1, npaths, 2, ref["x2"], ref["dx"], ref["res"], ref["st"],
                     iters=3, tol=1e-28, only=bufs.status)
    t_ref = ev.stop_ms()
    t0 = time.perf_counter()
    st = ref["st"].to_host()
    status1 = bufs.status.to_host()
    tracked = status1 == 1
    bad = np.nonzero((st != 1) & tracked)[0]
    fix = {}
    # precision escalation: interior failures of FP64 tracking, tracked in dd
    esc_idx, esc = escalate_to_dd(tr, starts, status1, bufs.tend.to_host(), prm, endp=bufs.endp)
    if esc is not None:
        xe = to_qd(esc["endp"], 2)
        for j, p in enumerate(esc_idx):
            fix[int(p)] = (xe[:, j], bool(esc["status"][j] == 1))
    nres = 0
    lost = np.zeros(0, np.int64)
    if len(bad):
        # escalate the unconverged endpoints to quad-double
        r4 = tr.refine(ref["x2"].to_host()[:, bad], 4, iters=4, tol=1e-56)
        ok4 = r4["status"] == 1
        for j, p in enumerate(bad):
            fix[int(p)] = (r4["x"][:, j], bool(ok4[j]))
        lost = bad[~ok4]
    if controller is not None:
        if len(lost):
            o = tr.track(starts[:, lost], K=1, params=prm)
            r2 = tr.refine(o["endp"], 2, iters=3, tol=1e-28)
            x4 = to_qd(r2["x"], 2)
            ok = (o["status"] == 1) & (r2["status"] == 1)
            esc = (o["status"] == 1) & ~ok
            if esc.any():
                r4b = tr.refine(x4[:, esc], 4, iters=4, tol=1e-56)
                x4[:, esc] = r4b["x"]
                ok[esc] = r4b["status"] == 1
            for j, p in enumerate(lost):
                fix[int(p)] = (x4[:, j], bool(ok[j]))
            nres = len(lost)
    check(lib.phc_synchronize())
    t_ref += (time.perf_counter() - t0) * 1e3
    ref["fix"] = fix
    return t_track, t_ref, len(bad), nres, len(esc_idx)


def qd_residuals(tr, x_planes):
    """max_i |p_i(x)| in quad-double at t = 1 (verification, not timed)."""
    return tr.refine(x_planes, 4, iters=0, tol=0.0)["res"]


def to_qd(x_planes, K):
    out = np.zeros((8,) + x_planes.shape[1:])
    out[:K] = x_planes[:K]
    out[4:4 + K] = x_planes[K:]
    return out


def count_lost(plain_x, plain_ok, mixed_x, mixed_ok, tol=1e-8):
    """solutions found by the plain tracker that the mixed one misses"""
    a = plain_x[plain_ok]
    b = mixed_x[mixed_ok]
    if len(a) == 0:
        return 0
    if len(b) == 0:
        return len(a)
    lost = 0
    for i in range(0, len(a), 256):
        d = np.abs(a[i:i + 256, None, :] - b[None, :, :]).max(axis=2).min(axis=1)
        lost += int((d > tol).sum())
    return lost


def distinct(x, ok, tol=1e-8):
    pts = x[ok]
    keys = set()
    for q in np.round(pts / tol).astype(np.complex128):
        keys.add(tuple(np.round(q.real).astype(np.int64)) + tuple(np.round(q.imag).astype(np.int64)))
    return len(keys)


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE passes, copy-kernel calibrated), or
    None when no summary exists for it."""
    f = os.path.join(ROOT, "profiles", "tracker_c3_hbm_traffic.json")
    try:
        with open(f) as fh:
            return float(json.load(fh)["bytes_per_launch"][kernel])
    except (OSError, KeyError, ValueError):
        return None


def host_cpu():
    """the host's CPU model and the cores this process may use"""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return dict(model=model, nproc=avail, os_cpu_count=os.cpu_count())


def cpu_baseline(hom, sols, seconds=15.0, pade=(5, 1)):
    """The oracle's plain dd tracker (the reference's step rule restated)
    on a bounded random sample of the paths: first on one core, as the
    reference's host path runs, then on all usable host cores (capped at
    16) over independent paths as the Ada Tasking trackers do."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oraclelib import oracle_track
    from phcpack_amd.homotopy import compile_homotopy
    from phcpack_amd.tracker import TrackParams, points_to_planes
    prog = compile_homotopy(hom, 8)
    cpu = host_cpu()
    threads = max(1, min(16, cpu["nproc"]))
    prm = TrackParams(tolend=1e-28, pade_numdeg=pade[0], pade_dendeg=pade[1])
    order = np.random.default_rng(SEED + 7).permutation(len(sols))

    def timed(m, nth):
        pick = np.sort(order[:m])
        t0 = time.perf_counter()
        out = oracle_track(prog, points_to_planes(sols[pick]), 2, prm, nthreads=nth)
        return time.perf_counter() - t0, int((out["status"] == 1).sum())

    # one core: calibrate on 2 paths, then about a third of the time budget
    dt, _ = timed(2, 1)
    m1 = int(min(len(sols), max(2, 2 / max(dt, 1e-9) * seconds / 3)))
    dt1, ok1 = timed(m1, 1)
    # all cores: the rest of the budget
    mt = int(min(len(sols), max(threads, m1 / dt1 * threads * seconds * 2 / 3)))
    mt = max(threads, (mt // threads) * threads)
    dtt, okt = timed(mt, threads)
    return dict(value=round(mt / dtt, 3), unit="paths/s", cores=threads, kind="port",
                sample=f"random {mt} of {len(sols)} paths (seed {SEED + 7}), plain dd tracker "
                       f"(oracle/track.c, the reference's step rule), {threads} threads, "
                       f"{dtt:.1f} s, {okt} succeeded",
                one_core={"value": round(m1 / dt1, 3), "unit": "paths/s", "cores": 1,
                          "sample": f"random {m1} paths, {dt1:.1f} s, {ok1} succeeded"},
                cpu_model=cpu["model"], nproc=cpu["nproc"])


def parse_predictor(name, deg=8):
    """'taylorL' -> (L, 0), 'padeL/M' -> (L, M), M = 1, 2: TrackParams'
    pade_numdeg and pade_dendeg; L + M + 2 <= deg (the error estimate of
    the curvature step reads coefficient L + M + 2)"""
    import re
    m = re.fullmatch(r"taylor(\d+)|pade(\d+)/([12])", name)
    if not m:
        raise SystemExit(f"--predictor {name!r}: expected taylorL, padeL/1 or padeL/2")
    if m.group(1) is not None:
        L = int(m.group(1))
        if not 1 <= L <= deg - 2:
            raise SystemExit(f"--predictor {name!r}: Taylor degree must be in 1..{deg - 2}")
        return (L, 0)
    L, M = int(m.group(2)), int(m.group(3))
    if not 1 <= L <= deg - 2 - M:
        raise SystemExit(f"--predictor {name!r}: Pade numerator degree must be in "
                         f"1..{deg - 2 - M}")
    return (L, M)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("tracker", "qr", "newton", "rankdef", "learned"),
                    default="tracker",
                    help="tracker: the headline line; qr: complex quad-double blocked QR "
                         "least-squares solve of 1024 x 1024; newton: batched dd row Newton; "
                         "rankdef: 14 quadratics with rank-deficient leading forms, 16384 "
                         "paths; learned: step controller trained on 10 dense cubic systems "
                         "(n = 8, 65610 paths), evaluated on 5 held-out systems")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip controller training; the headline is then the plain-rule mixed tracker")
    ap.add_argument("--predictor", default="pade5/1",
                    help="tracker predictor of every method (and of the CPU baseline): "
                         "padeL/M (the [L/M] Pade approximant, M = 1, 2, L <= 6 - M; the "
                         "reference's default [5/1]) or taylorL (L <= 6) of the degree-8 series")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-rankdef", action="store_true",
                    help="leave the rank-deficient n = 14 block out of the headline line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also launch one 256 MiB device copy kernel (HBM counter scale)")
    ap.add_argument("--save-controller", metavar="NPZ",
                    help="write the trained step controller to this file")
    ap.add_argument("--load-controller", metavar="NPZ",
                    help="use a saved step controller instead of training one (profiled runs)")
    args = ap.parse_args()
    args.pade = parse_predictor(args.predictor)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the paths are "
              "independent and would be split evenly across GPUs, but this build measures "
              "one GPU.", file=sys.stderr)
        return 2
    if args.workload == "qr":
        return run_qr(args)
    if args.workload == "newton":
        return run_newton(args)
    if args.workload == "rankdef":
        return run_rankdef(args)
    if args.workload == "learned":
        return run_learned(args)
    return run(args)


def candidate_sets(orders=True):
    """the candidate sets the controller is trained with: step multipliers
    only (series degree 8), and, with orders, the joint step-size and
    predictor-order candidates: every multiplier with degree 8 or 6, the
    choice made after the series to degree 6"""
    from phcpack_amd.controller import MULTS, order_candidates
    sets = [("multiplier", MULTS, None, 0)]
    if orders:
        m, d = order_candidates(low=6)
        sets.append(("multiplier+order", m, d, 6))
    return sets


def train_and_tune(trackers, starts, tune_tr, tune_sols, prm, ev, epochs=25, orders=True):
    """For every candidate set: explore-mode logs of the training systems,
    then two networks, one per label (accepted at the first attempt;
    accepted within 2 Newton iterations), each with its threshold and pole
    cap tuned by time on the tuning system; the fastest network is kept.
    Returns (controller or None, summary)."""
    from phcpack_amd.controller import (StepController, collect_training_logs, train_controller,
                                        tune_threshold)

    def timer(fn):
        ev.start()
        fn()
        return ev.stop_ms()

    best_ctl, best_ms, runs = None, None, []
    for name, mults, degrees, deg0 in candidate_sets(orders):
        explore = StepController(mults=mults, degrees=degrees, deg0=deg0)
        logs = collect_training_logs(trackers, starts, prm, K=1, seed=7, controller=explore)
        for mn in (None, 2):
            ctl, info = train_controller(trackers, starts, prm, K=1, seed=7, epochs=epochs,
                                         max_newton=mn, logs=logs, mults=mults, degrees=degrees,
                                         deg0=deg0)
            best, table = tune_threshold(ctl, tune_tr, tune_sols, prm, K=1, timer=timer)
            ms = table[f"{best[0]}/{best[1]}"]["ms"] if best is not None else None
            info.update(candidates=name, threshold=best[0] if best else None,
                        pole_cap=best[1] if best else None, tuned_ms=ms,
                        tune_table={str(k): v for k, v in table.items()})
            runs.append(info)
            if ms is not None and (best_ms is None or ms < best_ms):
                best_ctl, best_ms = ctl, ms
    chosen = min((r for r in runs if r["tuned_ms"] is not None), key=lambda r: r["tuned_ms"],
                 default=None)
    summary = dict(chosen or runs[0])
    summary["networks"] = [{k: r[k] for k in ("candidates", "max_newton", "accept_rate",
                                              "loss_last", "threshold", "pole_cap", "tuned_ms")}
                           for r in runs]
    return best_ctl, summary


def train_learned(n, prm, ev):
    """Train the step controller on systems of the same family with other
    seeds (explore-mode logs of FP64 tracking), then tune its threshold on
    a further system.  Not part of the timed region."""
    from phcpack_amd.tracker import Tracker
    t0 = time.perf_counter()
    trs, sts = [], []
    for s in TRAIN_SEEDS:
        _, hom, sols = build_workload(n, s)
        trs.append(Tracker(hom, deg=8))
        sts.append(sols)
    _, hom, sols = build_workload(n, TUNE_SEED)
    ctl, info = train_and_tune(trs, sts, Tracker(hom, deg=8), sols, prm, ev)
    info.update(train_seeds=list(TRAIN_SEEDS), tune_seed=TUNE_SEED,
                seconds=round(time.perf_counter() - t0, 2))
    return ctl, info


def verify(tr, out, ref, K_end):
    """endpoints (qd planes), qd residuals and success mask of one method"""
    ok_track = out["status"] == 1
    if ref is None:
        x = to_qd(out["endp"], K_end)
    else:
        x = to_qd(ref["x2"].to_host(), 2)
        for p, (xq, okp) in ref["fix"].items():
            x[:, p] = xq
            ok_track[p] = okp
    res = qd_residuals(tr, x)
    ok = ok_track & (res <= RES_TOL)
    return x, res, ok


def run(args):
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.flops import tracker_flops
    from phcpack_amd.tracker import TrackBuffers, Tracker, TrackParams, planes_to_points, points_to_planes
    lib = load()
    p, hom, sols = build_workload()
    npaths, n = len(sols), hom.n
    tr = Tracker(hom, deg=8)
    h_start = points_to_planes(sols)
    d_start = DeviceArray.from_host(h_start)
    ev = Events(lib)
    pnum, pden = args.pade
    prm_plain = TrackParams(tolend=1e-28, pade_numdeg=pnum, pade_dendeg=pden)
    prm_mixed = TrackParams(tolend=1e-13, pade_numdeg=pnum, pade_dendeg=pden)
    b_plain = TrackBuffers(npaths, n, 2)
    b_mixed = TrackBuffers(npaths, n, 1)
    b_learn = TrackBuffers(npaths, n, 1)

    def refbufs():
        return dict(x2=DeviceArray((4, npaths, n)), dx=DeviceArray((npaths,)),
                    res=DeviceArray((npaths,)), st=DeviceArray((npaths,), np.int32))

    ref_m, ref_l = refbufs(), refbufs()
    if args.calibrate_copy:
        big = DeviceArray((32 * 1024 * 1024,))
        dst = DeviceArray((32 * 1024 * 1024,))
        check(lib.phc_device_copy(dst.ptr, big.ptr, big.nbytes, None))
        check(lib.phc_synchronize())

    if args.no_learned:
        ctl, train_info = None, None
    elif args.load_controller:
        from phcpack_amd.controller import StepController
        ctl = StepController.load(args.load_controller)
        train_info = {"loaded_from": args.load_controller}
    else:
        ctl, train_info = train_learned(n, prm_mixed, ev)
    if ctl is not None and args.save_controller:
        ctl.save(args.save_controller)
    cstruct = ctl.struct(1) if ctl is not None else None
    # launch shape per precision, tuned on this system before the timed region
    from phcpack_amd.autotune import tune_waves
    tuned = {}
    for K_, b_, p_, c_ in ((2, b_plain, prm_plain, None), (1, b_mixed, prm_mixed, None)):
        tuned[K_] = tune_waves(tr, d_start, 1, npaths, K_, b_, p_, controller=c_)

    for _ in range(args.warmup):
        run_plain(tr, d_start, npaths, b_plain, ev, prm_plain)
        run_mixed(tr, d_start, h_start, npaths, b_mixed, ref_m, ev, prm_mixed, lib)
        if cstruct is not None:
            run_mixed(tr, d_start, h_start, npaths, b_learn, ref_l, ev, prm_mixed, lib, cstruct)
    check(lib.phc_synchronize())

    def timed(fn):
        rows = []
        check(lib.phc_synchronize())
        w0 = time.perf_counter()
        for _ in range(args.steps):
            rows.append(fn())
        check(lib.phc_synchronize())
        return rows, (time.perf_counter() - w0) / args.steps * 1e3

    rows_l, wall_l = (timed(lambda: run_mixed(tr, d_start, h_start, npaths, b_learn, ref_l, ev,
                                              prm_mixed, lib, cstruct))
                      if cstruct is not None else (None, None))
    rows_m, wall_m = timed(lambda: run_mixed(tr, d_start, h_start, npaths, b_mixed, ref_m, ev,
                                             prm_mixed, lib))
    rows_p, wall_p = timed(lambda: run_plain(tr, d_start, npaths, b_plain, ev, prm_plain))

    # ---- verification (not timed): quad-double residuals at t = 1 ----
    op, om = b_plain.to_host(), b_mixed.to_host()
    xp, res_p, ok_p = verify(tr, op, None, 2)
    xm, res_m, ok_m = verify(tr, om, ref_m, 2)
    cp, cm = planes_to_points(xp, 4), planes_to_points(xm, 4)
    t_plain = float(np.mean(rows_p))
    v_plain = ok_p.sum() / (t_plain / 1e3)

    def block(out, rows, wall, x, res, ok, c):
        t_track = float(np.mean([r[0] for r in rows]))
        t_ref = float(np.mean([r[1] for r in rows]))
        t = t_track + t_ref
        v = ok.sum() / (t / 1e3)
        return v, t, t_track, {
            "paths_per_s": round(float(v), 1), "ms_per_step": round(t, 3),
            "good_paths": int(ok.sum()), "max_residual": float(res.max()),
            "track_ms": round(t_track, 3), "refine_ms": round(t_ref, 3),
            "escalated_paths": int(rows[-1][2]), "retracked_paths": int(rows[-1][3]),
            "dd_escalated_paths": int(rows[-1][4]),
            "steps_per_path": float(out["nsteps"].mean()),
            "rejections_per_path": float(out["nrej"].mean()),
            "newton_steps_per_s": round(int(out["work"][1].sum()) / (t / 1e3), 1),
            "distinct_solutions": distinct(c, ok),
            "lost_paths_vs_plain": int(count_lost(cp, ok_p, c, ok)),
            "wall_ms_per_step": round(wall, 3)}

    v_m, t_m, tt_m, blk_m = block(om, rows_m, wall_m, xm, res_m, ok_m, cm)
    head = dict(v=v_m, t=t_m, tt=tt_m, out=om, lost=blk_m["lost_paths_vs_plain"], name="mixed")
    blk_l = None
    if rows_l is not None:
        ol = b_learn.to_host()
        xl, res_l, ok_l = verify(tr, ol, ref_l, 2)
        cl = planes_to_points(xl, 4)
        v_l, t_l, tt_l, blk_l = block(ol, rows_l, wall_l, xl, res_l, ok_l, cl)
        blk_l["threshold"] = ctl.threshold
        blk_l["pole_cap"] = ctl.pole_cap
        blk_l["training"] = train_info
        head = dict(v=v_l, t=t_l, tt=tt_l, out=ol, lost=blk_l["lost_paths_vs_plain"],
                    name="learned+mixed")

    fl_track = tracker_flops(tr.prog, 1, head["out"]["work"])
    fl_plain = tracker_flops(tr.prog, 2, op["work"])
    achieved = fl_track / (head["tt"] / 1e3) / 1e12
    line = {
        "metric": METRIC,
        "value": round(float(head["v"]), 1),
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["t"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "total-degree homotopy, 12 dense complex quadratics in 12 variables, "
                        f"4096 paths, degree-8 series, {args.predictor} predictor, seed 20261202",
            "method": head["name"] + ": FP64 tracking + dd endpoint Newton (qd escalation)",
            "paths": npaths, "n": n, "series_degree": 8, "predictor": args.predictor,
            "seed": SEED,
            "residual_tol": RES_TOL,
            "residual_check": "max_i |p_i(x)| at the exact endpoint, quad-double on the GPU",
        },
        "plain_dd": {"paths_per_s": round(float(v_plain), 1), "ms_per_step": round(t_plain, 3),
                     "good_paths": int(ok_p.sum()), "max_residual": float(res_p.max()),
                     "achieved_tflops": round(fl_plain / (t_plain / 1e3) / 1e12, 3),
                     "steps_per_path": float(op["nsteps"].mean()),
                     "wall_ms_per_step": round(wall_p, 3)},
        "mixed": blk_m,
        "learned_mixed": blk_l,
        "speedup_vs_plain": round(float(head["v"] / v_plain), 3) if v_plain > 0 else None,
        "speedup_mixed_vs_plain": round(float(v_m / v_plain), 3) if v_plain > 0 else None,
        "lost_paths_vs_plain": int(head["lost"]),
        "waves_per_workgroup": {"plain_dd": tuned[2][0], "fp64": tuned[1][0],
                                "tuning_ms": {str(k): {str(w): round(ms, 2) for w, ms in v[1].items()}
                                              for k, v in tuned.items()}},
        "roofline": {
            "bound": "fp64-valu",
            "achieved": round(achieved, 4),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
            "traffic": measured_traffic("track_kernel<1>"),
            "traffic_source": "profiles/tracker_c3_hbm_traffic.md",
            "kernel": "track_kernel<1> (FP64 tracking, " + head["name"] + ")",
            "note": "compute bound on the FP64 vector ALU (no MFMA: the multiple-double "
                    "arithmetic is scalar); MI355X FP64 peak 78.6 TFLOP/s for both VALU and "
                    "MFMA; flops counted from the program and per-path work counters "
                    "(phcpack_amd/flops.py, FMA = 2)",
            "flops_per_launch": float(fl_track),
            "kernel_ms": round(head["tt"], 3),
        },
    }
    # the rank-deficient family (n = 14, 16384 paths) with the same controller
    if not args.no_rankdef:
        line["rankdef_n14"] = rankdef_block(args, lib, ev, cstruct)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(hom, sols, args.cpu_seconds, args.pade)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


def md_lead_diff(lib, K, da, db, nelem):
    """max over the elements of |leading limb of a - b|, the difference taken
    in K-limb arithmetic on the GPU, for complex planar [2K][nelem] arrays"""
    from phcpack_amd._lib import DeviceArray, check
    dc = DeviceArray((2 * K, nelem))
    for h in (0, 1):
        off = h * K * nelem * 8
        check(lib.phc_md_batch(1, K, nelem, ctypes.c_void_p(da.ptr.value + off),
                               ctypes.c_void_p(db.ptr.value + off),
                               ctypes.c_void_p(dc.ptr.value + off), None))
    c = dc.to_host()
    return float(max(np.abs(c[0]).max(), np.abs(c[K]).max()))


def pad_limbs(d, K, K2):
    """complex planar [2K][...] device array -> [2K2][...] with zero limbs"""
    from phcpack_amd._lib import DeviceArray
    h = d.to_host()
    out = np.zeros((2 * K2,) + h.shape[1:])
    out[:K], out[K2:K2 + K] = h[:K], h[K:]
    return DeviceArray.from_host(out)


def octo_double_checks(lib, K, n, dA, dx, db, dx0):
    """max |A x - b| and max |x - x0| (leading limbs) evaluated in
    octo-double on the GPU for the K-limb solve A x = b"""
    from phcpack_amd._lib import DeviceArray, check
    from phcpack_amd.baqr import gemm_device
    A8, x8, b8, x08 = (pad_limbs(d, K, 8) for d in (dA, dx, db, dx0))
    r8 = DeviceArray.zeros((16, n))
    gemm_device(8, False, False, n, 1, n, (A8, 0, n, 1), (x8, 0, 1, 0), (r8, 0, 1, 0))
    check(lib.phc_synchronize())
    return {"max_abs_Ax_minus_b": md_lead_diff(lib, 8, r8, b8, n),
            "max_abs_x_minus_x0": md_lead_diff(lib, 8, x8, x08, n),
            "note": "A, x, b, x0 padded to eight limbs; r = A x by the octo-double complex GEMM"}


def run_qr(args, n=1024, szt=128, K=4, seed=SEED):
    """Blocked Householder QR least-squares solve in complex quad-double:
    A (n x n, re and im uniform in [-1,1]) = QR, y = Q^H b, R x = y by the
    tiled back substitution, b = A x0 with x0 random; inputs resident on the
    device.  Checks (outside the timed region, on the GPU): max |Q^H Q - I|,
    max |QR - A|, max |x - x0| in quad-double, and the residual max |A x - b|
    and max |x - x0| of the quad-double solution in octo-double."""
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.baqr import BlockedQR, gemm_device, upper_tiled_solver_device
    from phcpack_amd.flops import BUILD, baqr_canonical, baqr_work, complex_mac
    lib = load()
    rng = np.random.default_rng(seed)
    P = 2 * K
    A = np.zeros((P, n, n))
    A[0], A[K] = rng.uniform(-1, 1, (n, n)), rng.uniform(-1, 1, (n, n))
    x0 = np.zeros((P, n))
    x0[0], x0[K] = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    dA, dx0 = DeviceArray.from_host(A), DeviceArray.from_host(x0)
    db = DeviceArray.zeros((P, n))
    gemm_device(K, False, False, n, 1, n, (dA, 0, n, 1), (dx0, 0, 1, 0), (db, 0, 1, 0))
    # the GEMM register block and k split of the plan, tuned on this matrix
    # (the tile size is the workload's: 8 tiles of 128)
    from phcpack_amd.autotune import QR_BLOCKS, tune_qr
    best, table = tune_qr(K, n, n, dA, szts=[szt],
                          log=lambda s: print(s, file=sys.stderr, flush=True))
    qr = BlockedQR(K, n, n, szt)
    qr.set_tuning(best[1], best[2])
    dQ, dR = DeviceArray((P, n, n)), DeviceArray((P, n, n))
    dy, dx, dbw = DeviceArray((P, n)), DeviceArray((P, n)), DeviceArray((P, n))
    ev = Events(lib)

    def step():
        qr.factor_device(dA, dQ, dR)
        qr.qhb_device(dQ, db, dy)
        upper_tiled_solver_device(K, n, szt, n // szt, dR, dy, dx)

    for _ in range(args.warmup):
        step()
    check(lib.phc_synchronize())
    t_qr, t_all = [], []
    for _ in range(args.steps):
        ev.start()
        qr.factor_device(dA, dQ, dR)
        t_qr.append(ev.stop_ms())
        ev.start()
        qr.qhb_device(dQ, db, dy)
        upper_tiled_solver_device(K, n, szt, n // szt, dR, dy, dx)
        t_all.append(t_qr[-1] + ev.stop_ms())
    check(lib.phc_synchronize())
    # checks in quad-double on the GPU (R's diagonal tiles now hold inverses: refactor)
    qr.factor_device(dA, dQ, dR)
    dI = DeviceArray.zeros((P, n, n))
    gemm_device(K, True, False, n, n, n, (dQ, 0, 1, n), (dQ, 0, n, 1), (dI, 0, n, 1))
    eye = np.zeros((P, n, n))
    eye[0] = np.eye(n)
    qq = md_lead_diff(lib, K, dI, DeviceArray.from_host(eye), n * n)
    dQR = DeviceArray.zeros((P, n, n))
    gemm_device(K, False, False, n, n, n, (dQ, 0, n, 1), (dR, 0, n, 1), (dQR, 0, n, 1))
    qra = md_lead_diff(lib, K, dQR, dA, n * n)
    xerr = md_lead_diff(lib, K, dx, dx0, n)
    # the error of the quad-double solution in octo-double on the GPU: A, x
    # and b padded with zero limbs, r = A x - b by the octo-double complex
    # GEMM, and x - x0 (exact in octo-double)
    od = octo_double_checks(lib, K, n, dA, dx, db, dx0)
    # the real QR of the real part of A with the same plan (reported beside
    # the metric, not part of it)
    dAr = DeviceArray.from_host(np.ascontiguousarray(A[:K]))
    dQr, dRr = DeviceArray((K, n, n)), DeviceArray((K, n, n))
    qr.factor_real_device(dAr, dQr, dRr)
    check(lib.phc_synchronize())
    t_real = []
    for _ in range(max(args.steps, 1)):
        ev.start()
        qr.factor_real_device(dAr, dQr, dRr)
        t_real.append(ev.stop_ms())
    ms_qr, ms = float(np.mean(t_qr)), float(np.mean(t_all))
    executed = baqr_work(n, n, szt) * complex_mac(K)
    cm, ca = baqr_canonical(n)
    canonical = cm * BUILD[K]["mul"] + ca * BUILD[K]["add"]
    line = {
        "metric": "complex quad-double blocked QR least-squares solves/sec (1024x1024)",
        "value": round(1e3 / ms, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (quad-double)", "data": "synthetic",
        "config": {"workload": f"complex quad-double {n}x{n}, {n // szt} tiles of {szt} columns, "
                               "entries uniform in [-1,1], b = A x0", "seed": seed},
        "qr_ms": round(ms_qr, 3),
        "real_qr_ms": round(float(np.mean(t_real)), 3),
        "tuning": {"gemm_block": QR_BLOCKS[best[1]], "split_target": best[2],
                   "qr_ms_by_block_and_split": {f"{QR_BLOCKS[b]}/{sp}": round(v, 2)
                                                for (_, b, sp), v in sorted(table.items())}},
        "checks": {"max_abs_QHQ_minus_I": qq, "max_abs_QR_minus_A": qra, "max_abs_x_minus_x0": xerr,
                   "octo_double": od},
        "roofline": {"bound": "fp64-valu", "achieved": round(executed / (ms_qr / 1e3) / 1e12, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(executed / (ms_qr / 1e3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                     "traffic": None, "kernel": "blocked QR (all kernels)",
                     "flops_executed": executed, "flops_canonical": canonical,
                     "canonical_tflops": round(canonical / (ms_qr / 1e3) / 1e12, 3)},
    }
    print(json.dumps(line), flush=True)
    return 0


def rankdef_block(args, lib, ev, cstruct=None, n=14, rank_deficient=2, seed=SEED):
    """Mixed-precision tracking of a rank-deficient family: 14 dense
    quadratics (16384 total-degree paths) whose last 2 top-degree forms are
    combinations of the top-degree forms of two others, so the leading forms
    are rank-deficient and part of the paths diverge or end at singular
    points.  FP64 tracking + dd endpoint Newton of the finished paths (qd for
    those dd does not settle) against the plain dd tracker, and with a
    learned controller (cstruct) when one is given; paths/s count endpoints
    with max_i |p_i(x)| <= 1e-24 in quad-double."""
    from phcpack_amd._lib import DeviceArray, check
    from phcpack_amd.autotune import tune_waves
    from phcpack_amd.flops import tracker_flops
    from phcpack_amd.homotopy import dense_system, total_degree_homotopy
    from phcpack_amd.tracker import TrackBuffers, Tracker, TrackParams, planes_to_points, points_to_planes
    rng = np.random.default_rng(seed)
    p = dense_system(n, 2, rng, rank_deficient=rank_deficient)
    hom, sols = total_degree_homotopy(p, rng)
    npaths = len(sols)
    tr = Tracker(hom, deg=8)
    h_start = points_to_planes(sols)
    d_start = DeviceArray.from_host(h_start)
    bm, bp, bl = TrackBuffers(npaths, n, 1), TrackBuffers(npaths, n, 2), TrackBuffers(npaths, n, 1)

    def refbufs():
        return dict(x2=DeviceArray((4, npaths, n)), dx=DeviceArray((npaths,)),
                    res=DeviceArray((npaths,)), st=DeviceArray((npaths,), np.int32))

    ref, ref_l = refbufs(), refbufs()
    pnum, pden = args.pade
    prm_m = TrackParams(tolend=1e-13, pade_numdeg=pnum, pade_dendeg=pden)
    prm_p = TrackParams(tolend=1e-28, pade_numdeg=pnum, pade_dendeg=pden)
    tuned = {K_: tune_waves(tr, d_start, 1, npaths, K_, b_, p_)
             for K_, b_, p_ in ((2, bp, prm_p), (1, bm, prm_m))}
    for _ in range(args.warmup):
        run_mixed(tr, d_start, h_start, npaths, bm, ref, ev, prm_m, lib)
        run_plain(tr, d_start, npaths, bp, ev, prm_p)
        if cstruct is not None:
            run_mixed(tr, d_start, h_start, npaths, bl, ref_l, ev, prm_m, lib, cstruct)
    check(lib.phc_synchronize())
    rows = [run_mixed(tr, d_start, h_start, npaths, bm, ref, ev, prm_m, lib) for _ in range(args.steps)]
    rows_l = ([run_mixed(tr, d_start, h_start, npaths, bl, ref_l, ev, prm_m, lib, cstruct)
               for _ in range(args.steps)] if cstruct is not None else None)
    t_plain = float(np.mean([run_plain(tr, d_start, npaths, bp, ev, prm_p) for _ in range(args.steps)]))
    check(lib.phc_synchronize())
    om, op = bm.to_host(), bp.to_host()
    xm, res_m, ok_m = verify(tr, om, ref, 2)
    xp, res_p, ok_p = verify(tr, op, None, 2)
    cm, cp = planes_to_points(xm, 4), planes_to_points(xp, 4)
    v_p = ok_p.sum() / (t_plain / 1e3)
    names = ["running", "success", "minstep", "maxsteps", "singular", "diverged"]
    cls = lambda st: {names[i]: int(c) for i, c in enumerate(np.bincount(st, minlength=6)) if c}  # noqa: E731

    def end_classes(o, K):
        """endgame estimates at every path's last point (phc_endgame, timed)
        and the classes of classify_endpoints"""
        from phcpack_amd.tracker import END_NAMES, classify_endpoints
        planes = np.ascontiguousarray(o["endp"])
        d_x = DeviceArray.from_host(planes)
        d_t = DeviceArray.from_host(np.ascontiguousarray(o["tend"]))
        est = {k: DeviceArray((npaths,)) for k in ("alpha", "rho", "rco", "xmax")}
        ev.start()
        tr.endgame_device(d_x, K, npaths, d_t, K, est["alpha"], est["rho"], est["rco"],
                          est["xmax"])
        ms = ev.stop_ms()
        e = {k: v.to_host() for k, v in est.items()}
        cls, wind = classify_endpoints(o["status"], o["tend"], e)
        out = {END_NAMES[i]: int(c) for i, c in enumerate(np.bincount(cls, minlength=4)) if c}
        ws = wind[cls == 2]
        return {"classes": out, "winding_numbers": {str(int(w)): int(c) for w, c in
                                                    zip(*np.unique(ws, return_counts=True))},
                "endgame_ms": round(ms, 3)}

    def method(o, rows_, x, res, ok, c):
        t_track = float(np.mean([r[0] for r in rows_]))
        t = float(np.mean([r[0] + r[1] for r in rows_]))
        fl = tracker_flops(tr.prog, 1, o["work"])
        return ok.sum() / (t / 1e3), {
            "ms": round(t, 3), "track_ms": round(t_track, 3),
            "paths_per_s": round(float(ok.sum() / (t / 1e3)), 1),
            "good_paths": int(ok.sum()), "tracker_statuses": cls(o["status"]),
            "escalated_paths": int(rows_[-1][2]), "retracked_paths": int(rows_[-1][3]),
            "dd_escalated_paths": int(rows_[-1][4]),
            "max_good_residual": float(res[ok].max()) if ok.any() else None,
            "distinct_solutions": distinct(c, ok),
            "steps_per_path": float(o["nsteps"].mean()),
            "lost_paths_vs_plain": int(count_lost(cp, ok_p, c, ok)),
            "fp64_tflops": round(fl / (t_track / 1e3) / 1e12, 3),
            "endgame": end_classes(o, 1)}

    v_m, blk_m = method(om, rows, xm, res_m, ok_m, cm)
    blk = {
        "workload": "total-degree homotopy, 14 dense complex quadratics, top-degree forms of "
                    f"{rank_deficient} replaced by random combinations of two others, "
                    f"{npaths} paths, degree-8 series, {args.predictor} predictor, seed {seed}",
        "mixed": blk_m,
        "plain_dd": {"ms": round(t_plain, 3), "paths_per_s": round(float(v_p), 1),
                     "good_paths": int(ok_p.sum()), "tracker_statuses": cls(op["status"]),
                     "max_good_residual": float(res_p[ok_p].max()) if ok_p.any() else None,
                     "distinct_solutions": distinct(cp, ok_p),
                     "steps_per_path": float(op["nsteps"].mean()),
                     "endgame": end_classes(op, 2)},
        "speedup_mixed_vs_plain": round(float(v_m / v_p), 3) if v_p > 0 else None,
        "waves_per_workgroup": {"plain_dd": tuned[2][0], "fp64": tuned[1][0]},
    }
    v_head, lost = v_m, blk_m["lost_paths_vs_plain"]
    if rows_l is not None:
        ol = bl.to_host()
        xl, res_l, ok_l = verify(tr, ol, ref_l, 2)
        v_l, blk_l = method(ol, rows_l, xl, res_l, ok_l, planes_to_points(xl, 4))
        blk["learned_mixed"] = blk_l
        blk["speedup_learned_vs_plain"] = round(float(v_l / v_p), 3) if v_p > 0 else None
        v_head, lost = v_l, blk_l["lost_paths_vs_plain"]
    blk["value"] = round(float(v_head), 1)
    blk["lost_paths_vs_plain"] = int(lost)
    return blk


def run_rankdef(args):
    """the rank-deficient family on its own line (plain dd vs mixed)"""
    from phcpack_amd._lib import load
    lib = load()
    blk = rankdef_block(args, lib, Events(lib))
    line = {
        "metric": "paths/sec at <=1e-24 endpoint residual, 14 dense quadratics with "
                  "rank-deficient leading forms (16384 paths), mixed vs plain dd",
        "value": blk["value"], "unit": "paths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": blk["mixed"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic", "config": {"workload": blk["workload"], "residual_tol": RES_TOL},
        "rankdef": blk, "roofline": None,
    }
    head = blk.get("learned_mixed") or blk.get("mixed") or {}
    if head.get("fp64_tflops") is not None:
        # the FP64 tracking launch of the headline method, flops counted as for
        # the 12-quadratics line (phcpack_amd/flops.py, FMA = 2)
        line["roofline"] = {"bound": "fp64-valu", "achieved": head["fp64_tflops"], "peak": 78.6,
                            "unit": "TFLOP/s", "frac": round(head["fp64_tflops"] / 78.6, 4),
                            "traffic": None,
                            "kernel": "track_kernel<1> (FP64 tracking, 16384 paths)"}
    print(json.dumps(line), flush=True)
    return 0


def run_learned(args, n=8, degree=3, ntrain=10, neval=5, seed=SEED):
    """Learned step control on dense cubics in 8 variables (all 165
    monomials of degree <= 3, 3^8 = 6561 total-degree paths per system).
    The controller is trained on explore-mode step logs of ntrain systems
    (seeds seed+100..; 65610 paths), its threshold and pole cap tuned on the
    last training system, then evaluated on neval held-out systems
    (seeds seed+110..): FP64 tracking with the learned controller vs the
    plain pole-step rule, both followed by the same dd endpoint Newton (qd
    escalation).  value = paths/s at <= 1e-24 of the learned method over all
    held-out systems; a path the learned controller loses is tracked again
    with the plain rule (its time counts)."""
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.homotopy import dense_system, total_degree_homotopy
    from phcpack_amd.tracker import TrackBuffers, Tracker, TrackParams, planes_to_points, points_to_planes
    lib = load()
    ev = Events(lib)
    prm = TrackParams(tolend=1e-13, pade_numdeg=args.pade[0], pade_dendeg=args.pade[1])

    def system(s):
        rng = np.random.default_rng(s)
        p = dense_system(n, degree, rng)
        hom, sols = total_degree_homotopy(p, rng)
        return Tracker(hom, deg=8), sols

    t0 = time.perf_counter()
    train = [system(seed + 100 + i) for i in range(ntrain)]
    ctl, info = train_and_tune([t for t, _ in train], [s for _, s in train], train[-1][0],
                               train[-1][1], prm, ev)
    info.update(train_paths=int(sum(len(s) for _, s in train)),
                seconds=round(time.perf_counter() - t0, 2))
    if ctl is None:
        print(json.dumps({"metric": "learned step control", "value": None,
                          "error": "no threshold keeps every path", "training": info}))
        return 1
    cstruct = ctl.struct(1)
    rows = []
    tot = dict(t_plain=0.0, t_learn=0.0, good_plain=0, good_learn=0, lost=0, paths=0)
    for i in range(neval):
        s = seed + 110 + i
        tr, sols = system(s)
        npaths = len(sols)
        h_start = points_to_planes(sols)
        d_start = DeviceArray.from_host(h_start)

        def bufs():
            return TrackBuffers(npaths, n, 1), dict(
                x2=DeviceArray((4, npaths, n)), dx=DeviceArray((npaths,)),
                res=DeviceArray((npaths,)), st=DeviceArray((npaths,), np.int32))

        (bp, rp), (bl, rl) = bufs(), bufs()
        for _ in range(args.warmup):
            run_mixed(tr, d_start, h_start, npaths, bp, rp, ev, prm, lib)
            run_mixed(tr, d_start, h_start, npaths, bl, rl, ev, prm, lib, cstruct)
        check(lib.phc_synchronize())
        tp = [run_mixed(tr, d_start, h_start, npaths, bp, rp, ev, prm, lib) for _ in range(args.steps)]
        tl = [run_mixed(tr, d_start, h_start, npaths, bl, rl, ev, prm, lib, cstruct)
              for _ in range(args.steps)]
        op, ol = bp.to_host(), bl.to_host()
        xp, res_p, ok_p = verify(tr, op, rp, 2)
        xl, res_l, ok_l = verify(tr, ol, rl, 2)
        cp, cl = planes_to_points(xp, 4), planes_to_points(xl, 4)
        ms_p = float(np.mean([r[0] + r[1] for r in tp]))
        ms_l = float(np.mean([r[0] + r[1] for r in tl]))
        lost = int(count_lost(cp, ok_p, cl, ok_l))
        rows.append({"seed": s, "paths": npaths,
                     "plain": {"ms": round(ms_p, 3), "good_paths": int(ok_p.sum()),
                               "steps_per_path": float(op["nsteps"].mean()),
                               "rejections_per_path": float(op["nrej"].mean())},
                     "learned": {"ms": round(ms_l, 3), "good_paths": int(ok_l.sum()),
                                 "steps_per_path": float(ol["nsteps"].mean()),
                                 "rejections_per_path": float(ol["nrej"].mean()),
                                 "retracked_paths": int(tl[-1][3]),
                                 "dd_escalated_paths": int(tl[-1][4])},
                     "lost_paths_vs_plain": lost})
        tot["t_plain"] += ms_p
        tot["t_learn"] += ms_l
        # FP64 flops of the learned tracking launches (phcpack_amd/flops.py)
        from phcpack_amd.flops import tracker_flops
        tot["fl_learn"] = tot.get("fl_learn", 0.0) + tracker_flops(tr.prog, 1, ol["work"])
        tot["tt_learn"] = tot.get("tt_learn", 0.0) + float(np.mean([r[0] for r in tl]))
        tot["good_plain"] += int(ok_p.sum())
        tot["good_learn"] += int(ok_l.sum())
        tot["lost"] += lost
        tot["paths"] += npaths
    v_l = tot["good_learn"] / (tot["t_learn"] / 1e3)
    v_p = tot["good_plain"] / (tot["t_plain"] / 1e3)
    line = {
        "metric": "paths/sec at <=1e-24 endpoint residual, learned vs plain step control "
                  "(dense cubics n=8, held-out systems)",
        "value": round(v_l, 1), "unit": "paths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(tot["t_learn"] / neval, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"total-degree homotopies of {n} dense complex cubics (6561 paths "
                               f"each); controller trained on {ntrain} systems, evaluated on "
                               f"{neval} held-out systems; FP64 tracking ({args.predictor} "
                               "predictor) + dd endpoint Newton",
                   "train_seeds": [seed + 100 + i for i in range(ntrain)],
                   "eval_seeds": [seed + 110 + i for i in range(neval)]},
        "plain_rule": {"paths_per_s": round(v_p, 1), "good_paths": tot["good_plain"],
                       "ms_total": round(tot["t_plain"], 3)},
        "learned": {"paths_per_s": round(v_l, 1), "good_paths": tot["good_learn"],
                    "ms_total": round(tot["t_learn"], 3)},
        "speedup_learned_vs_plain_rule": round(v_l / v_p, 3),
        "lost_paths_vs_plain": tot["lost"],
        "systems": rows,
        "training": info,
        "roofline": None,
    }
    if tot.get("tt_learn", 0.0) > 0.0:
        ach = tot["fl_learn"] / (tot["tt_learn"] / 1e3) / 1e12
        line["roofline"] = {"bound": "fp64-valu", "achieved": round(ach, 3), "peak": 78.6,
                            "unit": "TFLOP/s", "frac": round(ach / 78.6, 4), "traffic": None,
                            "kernel": "track_kernel<1> (FP64 tracking, learned controller)"}
    print(json.dumps(line), flush=True)
    return 0


def run_newton(args, nsys=8192, seed=SEED):
    """Batched staggered row Newton in double-double: nsys systems of 8
    polynomials in 8 variables, 8 monomials of 3 distinct variables each
    with exponents uniform in {1, 2} (the common-factor trick), coefficient
    series of degree 15 uniform in [-1,1], 8 Newton steps from a random
    constant start series, one wavefront per system.  Outside the timed
    region: 16 systems checked bit for bit against the oracle, and the
    oracle's row Newton timed on one host core (the reference's host path
    is single-threaded)."""
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.newton import RowSystem
    lib = load()
    rng = np.random.default_rng(seed)
    dim, nper, nvar, deg, K = 8, 8, 3, 15, 2
    D = deg + 1
    nmono = dim * nper
    monoptr = np.arange(0, nmono + 1, nper, dtype=np.int32)
    nvr = np.full(nmono, nvar, np.int32)
    idx = np.concatenate([np.sort(rng.choice(dim, nvar, replace=False)) for _ in range(nmono)]).astype(np.int32)
    idxptr = np.concatenate([[0], np.cumsum(nvr)]).astype(np.int32)
    exps = rng.choice(np.array([1, 2], np.int32), size=len(idx)).astype(np.int32)
    sysd = RowSystem(dim, monoptr, nvr, idxptr, idx, exps)
    cst = np.zeros((K, nsys, dim, D))
    cff = np.zeros((K, nsys, nmono, D))
    cst[0] = rng.uniform(-1, 1, (nsys, dim, D))
    cff[0] = rng.uniform(-1, 1, (nsys, nmono, D))
    x0 = np.zeros((K, nsys, dim, D))
    x0[0, :, :, 0] = rng.uniform(-1, 1, (nsys, dim))
    dc, df = DeviceArray.from_host(cst), DeviceArray.from_host(cff)
    dx0 = DeviceArray.from_host(x0)
    dx = DeviceArray((K, nsys, dim, D))
    steps = DeviceArray((nsys,), np.int32)
    ev = Events(lib)
    for _ in range(args.warmup):
        check(lib.phc_device_copy(dx.ptr, dx0.ptr, dx.nbytes, None))
        sysd.row_newton_device(K, deg, 8, nsys, dc, False, df, False, dx, {"steps": steps})
    check(lib.phc_synchronize())
    ms = []
    for _ in range(args.steps):
        check(lib.phc_device_copy(dx.ptr, dx0.ptr, dx.nbytes, None))
        ev.start()
        sysd.row_newton_device(K, deg, 8, nsys, dc, False, df, False, dx, {"steps": steps})
        ms.append(ev.stop_ms())
    check(lib.phc_synchronize())
    st = steps.to_host()
    t = float(np.mean(ms))
    xg = dx.to_host()
    # parity and the CPU baseline (outside the timed region)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oraclelib import RowSystem as ORowSystem, oracle_row_newton
    nchk, bad = 16, 0
    t0 = time.perf_counter()
    for b in range(nchk):
        osys = ORowSystem(dim, deg, monoptr, nvr, idxptr, idx, np.ascontiguousarray(cst[:, b]),
                          np.ascontiguousarray(cff[:, b]), exps)
        o = oracle_row_newton(K, osys, np.ascontiguousarray(x0[:, b]), 8)
        bad += int(not np.array_equal(o["x"], xg[:, b]))
    cpu_s = (time.perf_counter() - t0) / nchk
    line = {
        "metric": "dd row Newton systems/sec (8 polynomials, 8 monomials of 3 variables with "
                  "exponents in {1,2}, degree 15, 8 steps)",
        "value": round(nsys / (t / 1e3), 1), "unit": "systems/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (double-double)", "data": "synthetic",
        "config": {"workload": f"{nsys} independent systems, one wavefront each", "seed": seed},
        "newton_steps_per_s": round(float(st.sum()) / (t / 1e3), 1),
        "parity": {"systems_checked": nchk, "bit_identical_to_oracle": nchk - bad},
        "roofline": None,
        "cpu_baseline": {"value": round(1.0 / cpu_s, 2), "unit": "systems/s", "cores": 1,
                         "kind": "port", "sample": f"{nchk} of the {nsys} systems, the oracle's "
                         "row Newton (bit-identical to the compiled reference)"},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
