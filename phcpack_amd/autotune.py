"""Launch-shape autotuners: the batched tracker and the blocked QR.

The tracker's persistent kernel runs a workgroup of W wavefronts per compute
unit, W bounded by the LDS the compiled program and the per-path state take
and by the register file.  More wavefronts hide more latency but share a
workgroup's issue slots; the best W depends on the system and the precision.
The tuner times each candidate W on the caller's own start points (the
results do not depend on W: every wavefront tracks its paths independently)
and keeps the fastest, per tracker and precision.

The blocked QR's tile size szt (where the caller may choose it: ncols must
be a multiple), the register block of its matrix products and the split of
their long k ranges over workgroups trade per-column latency (the panel)
against the throughput of the trailing updates; the best choice depends on
the limb count, the shape and the chip.  tune_qr times the whole
factorization for every candidate on the caller's matrix and keeps the
fastest.
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


QR_BLOCKS = {1: "1x1", 2: "2x1", 3: "1x2", 4: "2x2"}


def tune_qr(K, nrows, ncols, A_dev, szts=None, blocks=(1, 2, 3, 4), splits=(256, 1024, 4096),
            reps=2, log=None):
    """Time the blocked QR of the device matrix A_dev ([2K][nrows][ncols])
    for every tile size in szts (default: the divisors of ncols among 32,
    64, 128), GEMM register block and split target; returns
    (best (szt, block, split), {(szt, block, split): ms}).  log: a callable
    that receives one line per candidate."""
    from ._lib import DeviceArray
    from .baqr import BlockedQR
    if szts is None:
        szts = [s for s in (32, 64, 128) if ncols % s == 0] or [ncols]
    timer = _Timer()
    dQ, dR = DeviceArray((2 * K, nrows, nrows)), DeviceArray((2 * K, nrows, ncols))
    table = {}
    for szt in szts:
        qr = BlockedQR(K, nrows, ncols, szt)
        for blk in blocks:
            for spl in splits:
                qr.set_tuning(blk, spl)
                run = lambda: qr.factor_device(A_dev, dQ, dR)  # noqa: E731
                run()  # capture of the launch sequence, code objects
                ms = min(timer(run) for _ in range(reps))
                table[(szt, blk, spl)] = ms
                if log:
                    log(f"qr tune K={K} {nrows}x{ncols} szt={szt} block={QR_BLOCKS[blk]} "
                        f"split={spl}: {ms:.2f} ms")
        del qr
    best = min(table, key=table.get)
    return best, table
