"""Launch-shape autotuner of the batched tracker.

The tracker's persistent kernel runs a workgroup of W wavefronts per compute
unit, W bounded by the LDS the compiled program and the per-path state take
and by the register file.  More wavefronts hide more latency but share a
workgroup's issue slots; the best W depends on the system and the precision.
The tuner times each candidate W on the caller's own start points (the
results do not depend on W: every wavefront tracks its paths independently)
and keeps the fastest, per tracker and precision.
"""
import ctypes

from ._lib import check, load


class _Timer:
    def __init__(self):
        lib = load()
        self.lib = lib
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib.phc_event_create(ctypes.byref(self.a)))
        check(lib.phc_event_create(ctypes.byref(self.b)))

    def __call__(self, fn):
        check(self.lib.phc_event_record(self.a, None))
        fn()
        check(self.lib.phc_event_record(self.b, None))
        ms = ctypes.c_float()
        check(self.lib.phc_event_elapsed_ms(self.a, self.b, ctypes.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            self.lib.phc_event_destroy(self.a)
            self.lib.phc_event_destroy(self.b)
        except Exception:
            pass


def tune_waves(tracker, d_start, KS, npaths, K, bufs, params, candidates=(6, 8, 10, 12),
               controller=None, reps=1):
    """Time phc_track for every candidate waves-per-workgroup on the given
    start points; returns (best, {waves: ms}) and records the choice in
    tracker.waves[K] (used by track_device when waves is not given)."""
    timer = _Timer()
    cap = 12 if K == 1 else 8
    table = {}
    for w in candidates:
        if w > cap:
            continue
        run = lambda: tracker.track_device(d_start, KS, npaths, K, bufs, params, waves=w,  # noqa: E731
                                           controller=controller)
        run()  # warm (program upload to LDS, code object load)
        ms = min(timer(run) for _ in range(reps))
        table[w] = ms
    best = min(table, key=table.get)
    tracker.waves[K] = best
    return best, table
