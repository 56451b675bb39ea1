#!/usr/bin/env python3
"""Throughput of the batched path tracker on one MI355X.

Metric: paths/s whose endpoint residual max_i |p_i(x)| is <= 1e-24, for the
mixed-precision tracker (FP64 tracking, double-double Newton at the
endpoints, quad-double Newton for endpoints that do not converge in
double-double), beside the plain double-double tracker on the same system.
The residual of every endpoint is evaluated at the exact endpoint in
quad-double arithmetic on the GPU (rounding ~1e-60, far below the 1e-24
threshold); octo-double residuals of sampled endpoints are checked against
the CPU oracle by the tests.

Workload: total-degree homotopy gamma*(1-t)*q + t*p with p = 12 dense
quadratics in 12 variables (all 91 monomials of degree <= 2, complex
coefficients on the unit circle), q_i = x_i^2 - c_i, 4096 paths, predictor
series of degree 8, seed 20261202.  One step = tracking all 4096 paths from
t = 0 to t = 1 plus the endpoint refinement, inputs resident on the device.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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


def run_mixed(tr, d_start, npaths, bufs, ref, ev, prm, lib):
    """FP64 tracking, then dd Newton; paths whose dd Newton does not
    converge are refined in quad-double."""
    from phcpack_amd._lib import check
    ev.start()
    tr.track_device(d_start, 1, npaths, 1, bufs, prm)
    t_track = ev.stop_ms()
    ev.start()
    tr.refine_device(bufs.endp, 1, npaths, 2, ref["x2"], ref["dx"], ref["res"], ref["st"],
                     iters=3, tol=1e-28)
    t_ref = ev.stop_ms()
    st = ref["st"].to_host()
    bad = np.nonzero(st != 1)[0]
    escal = 0
    if len(bad):
        # escalate the unconverged paths to quad-double
        x2 = ref["x2"].to_host()[:, bad]
        t0 = time.perf_counter()
        r4 = tr.refine(x2, 4, iters=4, tol=1e-56)
        check(lib.phc_synchronize())
        escal = len(bad)
        t_ref += (time.perf_counter() - t0) * 1e3
        ref["escalated"] = (bad, r4["x"])
    else:
        ref["escalated"] = None
    return t_track, t_ref, escal


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


def cpu_baseline(hom, sols, seconds=15.0):
    """The oracle's plain dd tracker on a bounded sample of the paths, all
    host threads (capped at 16)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oraclelib import oracle_track
    from phcpack_amd.homotopy import compile_homotopy
    from phcpack_amd.tracker import TrackParams, points_to_planes
    prog = compile_homotopy(hom, 8)
    threads = max(1, min(16, os.cpu_count() or 1))
    prm = TrackParams(tolend=1e-28)
    # calibrate on one path per thread, then size the sample to ~seconds
    t0 = time.perf_counter()
    oracle_track(prog, points_to_planes(sols[:threads]), 2, prm, nthreads=threads)
    dt = time.perf_counter() - t0
    rate = threads / max(dt, 1e-9)
    m = int(min(len(sols), max(threads, rate * seconds)))
    m = max(threads, (m // threads) * threads)
    t0 = time.perf_counter()
    out = oracle_track(prog, points_to_planes(sols[:m]), 2, prm, nthreads=threads)
    dt = time.perf_counter() - t0
    ok = int((out["status"] == 1).sum())
    return dict(value=round(m / dt, 3), unit="paths/s", cores=threads, kind="port",
                sample=f"first {m} of {len(sols)} paths, plain dd tracker (oracle/track.c), "
                       f"{threads} threads, {dt:.1f} s, {ok} succeeded")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also launch one 256 MiB device copy kernel (HBM counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the paths are "
              "independent and would be split evenly across GPUs, but this build measures "
              "one GPU.", file=sys.stderr)
        return 2
    return run(args)


def run(args):
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.flops import tracker_flops, refine_flops
    from phcpack_amd.tracker import TrackBuffers, Tracker, TrackParams, planes_to_points, points_to_planes
    lib = load()
    p, hom, sols = build_workload()
    npaths, n = len(sols), hom.n
    tr = Tracker(hom, deg=8)
    d_start = DeviceArray.from_host(points_to_planes(sols))
    ev = Events(lib)
    prm_plain = TrackParams(tolend=1e-28)
    prm_mixed = TrackParams(tolend=1e-13)
    b_plain = TrackBuffers(npaths, n, 2)
    b_mixed = TrackBuffers(npaths, n, 1)
    ref = dict(x2=DeviceArray((4, npaths, n)), dx=DeviceArray((npaths,)),
               res=DeviceArray((npaths,)), st=DeviceArray((npaths,), np.int32))
    if args.calibrate_copy:
        big = DeviceArray((32 * 1024 * 1024,))
        dst = DeviceArray((32 * 1024 * 1024,))
        check(lib.phc_device_copy(dst.ptr, big.ptr, big.nbytes, None))
        check(lib.phc_synchronize())

    for _ in range(args.warmup):
        run_plain(tr, d_start, npaths, b_plain, ev, prm_plain)
        run_mixed(tr, d_start, npaths, b_mixed, ref, ev, prm_mixed, lib)
    check(lib.phc_synchronize())

    plain_ms, track_ms, refine_ms, mixed_ms = [], [], [], []
    escal = 0
    check(lib.phc_synchronize())
    wall0 = time.perf_counter()
    for _ in range(args.steps):
        a, b, e = run_mixed(tr, d_start, npaths, b_mixed, ref, ev, prm_mixed, lib)
        track_ms.append(a)
        refine_ms.append(b)
        mixed_ms.append(a + b)
        escal = e
    check(lib.phc_synchronize())
    wall_mixed = (time.perf_counter() - wall0) / args.steps * 1e3
    wall0 = time.perf_counter()
    for _ in range(args.steps):
        plain_ms.append(run_plain(tr, d_start, npaths, b_plain, ev, prm_plain))
    check(lib.phc_synchronize())
    wall_plain = (time.perf_counter() - wall0) / args.steps * 1e3

    # ---- verification (not timed): octo-double residuals at t = 1 ----
    op = b_plain.to_host()
    om = b_mixed.to_host()
    xp = op["endp"]
    res_p = qd_residuals(tr, to_qd(xp, 2))
    x2 = ref["x2"].to_host()
    xm = to_qd(x2, 2)
    if ref.get("escalated") is not None:
        bad, x4 = ref["escalated"]
        xm[:, bad] = x4
    res_m = qd_residuals(tr, xm)
    ok_p = (op["status"] == 1) & (res_p <= RES_TOL)
    ok_m = (om["status"] == 1) & (res_m <= RES_TOL)
    cp, cm = planes_to_points(xp, 2), planes_to_points(xm, 4)
    lost = count_lost(cp, ok_p, cm, ok_m)

    t_mixed = float(np.mean(mixed_ms))
    t_plain = float(np.mean(plain_ms))
    v_mixed = ok_m.sum() / (t_mixed / 1e3)
    v_plain = ok_p.sum() / (t_plain / 1e3)
    fl_track = tracker_flops(tr.prog, 1, om["work"])
    fl_plain = tracker_flops(tr.prog, 2, op["work"])
    t_track = float(np.mean(track_ms))
    achieved = fl_track / (t_track / 1e3) / 1e12
    newton_iters = int(om["work"][1].sum())
    line = {
        "metric": METRIC,
        "value": round(float(v_mixed), 1),
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "total-degree homotopy, 12 dense complex quadratics in 12 variables, "
                        "4096 paths, degree-8 series predictor, seed 20261202",
            "method": "mixed: FP64 tracking + dd endpoint Newton (qd/od escalation)",
            "paths": npaths, "n": n, "series_degree": 8, "seed": SEED,
            "residual_tol": RES_TOL,
            "residual_check": "max_i |p_i(x)| at the exact endpoint, quad-double on the GPU",
        },
        "plain_dd": {"paths_per_s": round(float(v_plain), 1), "ms_per_step": round(t_plain, 3),
                     "good_paths": int(ok_p.sum()), "max_residual": float(res_p.max()),
                     "achieved_tflops": round(fl_plain / (t_plain / 1e3) / 1e12, 3),
                     "steps_per_path": float(op["nsteps"].mean())},
        "mixed": {"good_paths": int(ok_m.sum()), "max_residual": float(res_m.max()),
                  "track_ms": round(t_track, 3), "refine_ms": round(float(np.mean(refine_ms)), 3),
                  "escalated_paths": int(escal), "steps_per_path": float(om["nsteps"].mean()),
                  "newton_steps_per_s": round(newton_iters / (t_mixed / 1e3), 1),
                  "distinct_solutions": distinct(cm, ok_m)},
        "speedup_mixed_vs_plain": round(float(v_mixed / v_plain), 3) if v_plain > 0 else None,
        "lost_paths_vs_plain": int(lost),
        "learned_controller": "not yet built",
        "wall_ms_per_step": {"mixed": round(wall_mixed, 3), "plain": round(wall_plain, 3)},
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 4),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
            "traffic": measured_traffic("track_kernel<1>"),
            "traffic_source": "profiles/tracker_c3_hbm_traffic.md",
            "kernel": "track_kernel<1> (FP64 tracking)",
            "note": "FP64 vector ALU; MI355X FP64 peak 78.6 TFLOP/s for both VALU and MFMA; "
                    "flops counted from the program and per-path work counters "
                    "(phcpack_amd/flops.py, FMA = 2)",
            "flops_per_launch": float(fl_track),
            "kernel_ms": round(t_track, 3),
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(hom, sols, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
