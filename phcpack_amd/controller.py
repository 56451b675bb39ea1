"""Learned step-size controller of the batched tracker.

The tracker forms 16 features per step (log2 Fabry radius, log2 previous
step, t, log2 of the largest series coefficients of orders 1..8, log2 of the
last four Newton corrections, rejections of the previous step) and an FP32
MLP 16 -> 64 -> 64 -> C gives one logit per candidate multiplier of the pole
step (C = 6 candidates 1, sqrt 2, 2, 2 sqrt 2, 4, 4 sqrt 2).  In learned mode
the tracker takes the largest candidate whose logit clears the threshold;
the pole-step rule of the reference
(Deformations/Curves/standard_predictor_convolutions.adb:1142-1185), the
predictor feedback and the Newton corrector stay on as judge and fallback.

Training data are the tracker's own step logs: in explore mode every step
tries a random candidate and records whether its first attempt was accepted
and how many Newton iterations the corrector then took; the label is
"accepted" or "accepted within max_newton iterations".
The network is trained on the GPU by the library's forward/backward kernels
(phc_mlp_train: tiled FP32 products, ReLU, binary cross-entropy, Adam).
"""
import ctypes

import numpy as np

from ._lib import DeviceArray, PhcController, check, load

NFEAT, HIDDEN = 16, 64
MULTS = tuple(float(2.0 ** (0.5 * i)) for i in range(6))


def order_candidates(mults=MULTS, low=6):
    """(multipliers, series degrees) of the joint step-size and
    predictor-order candidates: every multiplier with the program's series
    degree (0) and with degree `low`, the lower degree after the full one
    of the same multiplier, since learned mode takes the last candidate that
    clears the threshold (a step with fewer series orders costs less)."""
    m, d = [], []
    for x in mults:
        m += [x, x]
        d += [0, low]
    return tuple(m), tuple(d)


def nparams(F=NFEAT, H=HIDDEN, C=len(MULTS)):
    return F * H + H + H * H + H + H * C + C


def init_weights(rng, F=NFEAT, H=HIDDEN, C=len(MULTS)):
    """He-initialized weights in the phc_mlp layout (float32)."""
    w = []
    w.append(rng.standard_normal((F, H)) * np.sqrt(2.0 / F))
    w.append(np.zeros(H))
    w.append(rng.standard_normal((H, H)) * np.sqrt(2.0 / H))
    w.append(np.zeros(H))
    w.append(rng.standard_normal((H, C)) * np.sqrt(1.0 / H))
    w.append(np.zeros(C))
    return np.concatenate([a.reshape(-1) for a in w]).astype(np.float32)


def split(w, F=NFEAT, H=HIDDEN, C=len(MULTS)):
    o = 0
    out = []
    for shape in ((F, H), (H,), (H, H), (H,), (H, C), (C,)):
        n = int(np.prod(shape))
        out.append(w[o:o + n].reshape(shape))
        o += n
    return out


def forward_np(w, X, F=NFEAT, H=HIDDEN, C=len(MULTS)):
    """float64 reference forward pass; returns logits and activations"""
    W1, b1, W2, b2, W3, b3 = (a.astype(np.float64) for a in split(w, F, H, C))
    Z1 = X @ W1 + b1
    H1 = np.maximum(Z1, 0)
    Z2 = H1 @ W2 + b2
    H2 = np.maximum(Z2, 0)
    return H2 @ W3 + b3, (X, H1, H2)


def grad_np(w, X, cand, y, F=NFEAT, H=HIDDEN, C=len(MULTS)):
    """float64 gradient of the mean BCE of logit cand[i] against y[i]"""
    W1, b1, W2, b2, W3, b3 = (a.astype(np.float64) for a in split(w, F, H, C))
    O, (X, H1, H2) = forward_np(w, X, F, H, C)
    B = len(X)
    o = O[np.arange(B), cand]
    p = 1.0 / (1.0 + np.exp(-o))
    dO = np.zeros_like(O)
    dO[np.arange(B), cand] = (p - y) / B
    loss = np.mean(np.maximum(o, 0) - o * y + np.log1p(np.exp(-np.abs(o))))
    dW3 = H2.T @ dO
    db3 = dO.sum(0)
    dZ2 = (dO @ W3.T) * (H2 > 0)
    dW2 = H1.T @ dZ2
    db2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * (H1 > 0)
    dW1 = X.T @ dZ1
    db1 = dZ1.sum(0)
    g = np.concatenate([a.reshape(-1) for a in (dW1, db1, dW2, db2, dW3, db3)])
    return g, loss


def mlp_forward(w, X):
    lib = load()
    X = np.ascontiguousarray(X, np.float32)
    dw, dx = DeviceArray.from_host(np.ascontiguousarray(w, np.float32)), DeviceArray.from_host(X)
    C = len(MULTS)
    do = DeviceArray((len(X), C), np.float32)
    check(lib.phc_mlp_forward(NFEAT, HIDDEN, C, dw.ptr, len(X), dx.ptr, do.ptr, None))
    return do.to_host()


def mlp_grad(w, X, cand, y):
    lib = load()
    C = len(MULTS)
    dw = DeviceArray.from_host(np.ascontiguousarray(w, np.float32))
    dx = DeviceArray.from_host(np.ascontiguousarray(X, np.float32))
    dc = DeviceArray.from_host(np.ascontiguousarray(cand, np.int32))
    dy = DeviceArray.from_host(np.ascontiguousarray(y, np.float32))
    dg = DeviceArray((len(w),), np.float32)
    dl = DeviceArray((1,), np.float32)
    check(lib.phc_mlp_grad(NFEAT, HIDDEN, C, dw.ptr, len(X), dx.ptr, dc.ptr, dy.ptr, dg.ptr,
                           dl.ptr, None))
    return dg.to_host(), float(dl.to_host()[0])


class StepController:
    """Weights, input standardization and candidates of the controller."""

    def __init__(self, weights=None, mu=None, sd=None, threshold=0.8, mults=MULTS, seed=0,
                 pole_cap=1.0, degrees=None, deg0=0):
        self.mults = tuple(mults)
        self.C = len(self.mults)
        # series degree of every candidate's step (0: the program's) and the
        # degree computed before the choice (0: the program's)
        self.degrees = tuple(int(d) for d in (degrees or (0,) * self.C))
        if len(self.degrees) != self.C:
            raise ValueError("one series degree per candidate multiplier")
        self.deg0 = int(deg0)
        rng = np.random.default_rng(seed)
        self.w = init_weights(rng, C=self.C) if weights is None else np.asarray(weights, np.float32)
        self.mu = np.zeros(NFEAT, np.float32) if mu is None else np.asarray(mu, np.float32)
        self.sd = np.ones(NFEAT, np.float32) if sd is None else np.asarray(sd, np.float32)
        self.threshold = threshold
        self.pole_cap = pole_cap  # multiplied steps stay <= pole_cap * Fabry radius
        self._dev = None

    def save(self, path):
        """weights, standardization, candidates, threshold and pole cap as
        an .npz file (plain arrays, no pickled objects)"""
        np.savez(path, w=self.w, mu=self.mu, sd=self.sd, mults=np.asarray(self.mults),
                 threshold=np.float64(self.threshold), pole_cap=np.float64(self.pole_cap),
                 degrees=np.asarray(self.degrees, np.int32), deg0=np.int32(self.deg0))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            degrees = tuple(int(d) for d in z["degrees"]) if "degrees" in z else None
            deg0 = int(z["deg0"]) if "deg0" in z else 0
            return cls(weights=z["w"], mu=z["mu"], sd=z["sd"], threshold=float(z["threshold"]),
                       mults=tuple(float(m) for m in z["mults"]), pole_cap=float(z["pole_cap"]),
                       degrees=degrees, deg0=deg0)

    @property
    def logit_threshold(self):
        return float(np.log(self.threshold / (1.0 - self.threshold)))

    def device(self):
        if self._dev is None:
            self._dev = (DeviceArray.from_host(self.w), DeviceArray.from_host(self.mu),
                         DeviceArray.from_host(self.sd))
        return self._dev

    def struct(self, mode, logs=None, seed=0):
        c = PhcController()
        c.mode = mode
        c.ncand = self.C
        for i, m in enumerate(self.mults):
            c.mult[i] = m
        c.deg0 = self.deg0
        for i, d in enumerate(self.degrees):
            c.degree[i] = d
        if mode == 1:
            dw, dmu, dsd = self.device()
            c.weights, c.mu, c.sd = dw.ptr.value, dmu.ptr.value, dsd.ptr.value
        c.logit_threshold = self.logit_threshold
        c.seed = seed
        c.pole_cap = self.pole_cap
        if logs is not None:
            c.log_capacity = logs.cap
            c.log_x, c.log_c = logs.x.ptr.value, logs.c.ptr.value
            c.log_y, c.log_n = logs.y.ptr.value, logs.n.ptr.value
        return c

    def fit(self, X, cand, y, epochs=30, batch=1024, lr=3e-3, seed=0):
        """standardize the features and train on the GPU; returns epoch losses"""
        X = np.asarray(X, np.float32)
        self.mu = X.mean(0).astype(np.float32)
        self.sd = np.maximum(X.std(0), 1e-3).astype(np.float32)
        rng = np.random.default_rng(seed)
        perm = rng.permutation(len(X))
        Xn = ((X - self.mu) / self.sd)[perm]
        lib = load()
        dw = DeviceArray.from_host(self.w)
        dx = DeviceArray.from_host(np.ascontiguousarray(Xn, np.float32))
        dc = DeviceArray.from_host(np.ascontiguousarray(np.asarray(cand)[perm], np.int32))
        dy = DeviceArray.from_host(np.ascontiguousarray(np.asarray(y)[perm], np.float32))
        losses = (ctypes.c_float * epochs)()
        check(lib.phc_mlp_train(NFEAT, HIDDEN, self.C, dw.ptr, len(X), dx.ptr, dc.ptr, dy.ptr,
                                epochs, batch, lr, seed, losses, None))
        self.w = dw.to_host()
        self._dev = None
        return np.array(losses[:])


class StepLogs:
    """device buffers for explore-mode step logs"""

    def __init__(self, npaths, cap=256):
        self.cap = cap
        self.x = DeviceArray((npaths, cap, NFEAT), np.float32)
        self.c = DeviceArray((npaths, cap), np.int32)
        self.y = DeviceArray((npaths, cap), np.int32)
        self.n = DeviceArray.zeros((npaths,), np.int32)

    def to_host(self):
        n = self.n.to_host()
        x, c, y = self.x.to_host(), self.c.to_host(), self.y.to_host()
        keep = np.arange(self.cap)[None, :] < n[:, None]
        return x[keep], c[keep], y[keep]


def collect_logs(tracker, starts, params, seed=1, cap=256, K=1, controller=None):
    """track `starts` with random candidates and return (X, cand, y, outputs):
    y = the Newton iterations of a step accepted at its first attempt, 0 for
    a rejected first attempt"""
    from .tracker import TrackBuffers, points_to_planes
    ctl = controller or StepController()
    planes = points_to_planes(starts)
    npaths = planes.shape[1]
    logs = StepLogs(npaths, cap)
    d = DeviceArray.from_host(planes)
    bufs = TrackBuffers(npaths, tracker.n, K)
    tracker.track_device(d, 1, npaths, K, bufs, params, controller=ctl.struct(2, logs, seed))
    check(load().phc_synchronize())
    return logs.to_host() + (bufs.to_host(),)


def collect_training_logs(trackers, starts, params, K=1, seed=1, cap=256, controller=None):
    """explore-mode logs of every (tracker, starts) pair, concatenated:
    (features, candidates, Newton iterations or 0); controller: the
    candidates to explore (default: the multipliers MULTS)"""
    X, C, Y = [], [], []
    for i, (tr, st) in enumerate(zip(trackers, starts)):
        x, c, y, _ = collect_logs(tr, st, params, seed=seed + i, cap=cap, K=K,
                                  controller=controller)
        X.append(x)
        C.append(c)
        Y.append(y)
    return np.concatenate(X), np.concatenate(C), np.concatenate(Y)


def step_labels(Y, max_newton=None):
    """1 for a step accepted at its first attempt (within max_newton Newton
    iterations when given), else 0"""
    ok = Y >= 1
    if max_newton is not None:
        ok &= Y <= max_newton
    return ok.astype(np.float32)


def train_controller(trackers, starts, params, K=1, seed=1, cap=256, epochs=30, batch=1024,
                     lr=3e-3, threshold=0.8, max_newton=None, logs=None, mults=MULTS,
                     degrees=None, deg0=0):
    """Explore-mode logs from every (tracker, starts) pair (or the given
    logs), then GPU training on step_labels(., max_newton).  The candidates
    are the multipliers mults with the series degrees `degrees` (0: the
    program's; deg0 the degree computed before the choice).

    Returns the trained StepController and a summary dict (samples,
    positive rate, first and last epoch loss)."""
    explore = StepController(mults=mults, degrees=degrees, deg0=deg0)
    X, C, Y = logs if logs is not None else collect_training_logs(trackers, starts, params, K,
                                                                  seed, cap, explore)
    y = step_labels(Y, max_newton)
    ctl = StepController(threshold=threshold, seed=seed, mults=mults, degrees=degrees, deg0=deg0)
    losses = ctl.fit(X, C, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return ctl, dict(samples=int(len(X)), accept_rate=float(y.mean()), max_newton=max_newton,
                     loss_first=float(losses[0]), loss_last=float(losses[-1]))


def tune_threshold(ctl, tracker, starts, params, K=1,
                   grid=((0.5, 1.0), (0.7, 1.0), (0.8, 1.0), (0.9, 1.0), (0.7, 1.25),
                         (0.9, 0.75), (0.9, 1.5), (0.95, 0.75), (0.7, 1.5), (0.8, 1.5),
                         (0.7, 2.0), (0.9, 2.0)),
                   timer=None, repeats=2):
    """Pick (acceptance-probability threshold, pole cap) with the fewest
    predictor steps (or the shortest time when `timer(fn)` is given, the
    least of `repeats` runs) among those that lose no successful path
    relative to the plain pole-step rule on a tuning system.  Sets them on
    ctl; returns (choice or None, table)."""
    from .tracker import TrackBuffers, points_to_planes
    planes = points_to_planes(starts)
    npaths = planes.shape[1]
    d = DeviceArray.from_host(planes)
    bufs = TrackBuffers(npaths, tracker.n, K)

    def run(c):
        tracker.track_device(d, 1, npaths, K, bufs, params, controller=c)
        check(load().phc_synchronize())

    def measure(c):
        ms = min(timer(lambda: run(c)) for _ in range(repeats)) if timer else None
        if ms is None:
            run(c)
        out = bufs.to_host()
        return dict(ok=int((out["status"] == 1).sum()), pred=int(out["work"][0].sum()),
                    ms=ms)

    base = measure(None)
    table = {"plain": base}
    best, best_key = None, None
    for th, cap in grid:
        ctl.threshold, ctl.pole_cap = th, cap
        r = measure(ctl.struct(1))
        table[f"{th}/{cap}"] = r
        if r["ok"] < base["ok"]:
            continue
        key = r["ms"] if r["ms"] is not None else r["pred"]
        if best_key is None or key < best_key:
            best, best_key = (th, cap), key
    if best is not None:
        ctl.threshold, ctl.pole_cap = best
    return best, table
