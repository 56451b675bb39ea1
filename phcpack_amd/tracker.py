"""Host side of the batched path tracker: compiled homotopies on the device,
tracking, endpoint refinement and residuals through the C-ABI.

The interface mirrors the reference's batched tracker entry
gpumanypaths_dd(mode, verbose, regamma, imgamma)
(Ada/PHCtoC/gpumanypaths_dd.cpp:18-57) and its Path_Parameters
(Ada/PHCtoC/path_parameters.ads:14-60), reduced to what the series tracker
uses, with the step-control defaults of
Deformations/Curves/homotopy_continuation_parameters.adb:11-22.
"""
import ctypes
import dataclasses
from dataclasses import dataclass

import numpy as np

from ._lib import DeviceArray, P_double, P_int, PhcProgram, PhcTrackParams, check, load
from .homotopy import compile_homotopy

STATUS = {0: "running", 1: "success", 2: "step below minstep", 3: "maxsteps", 4: "singular",
          5: "diverged"}


@dataclass
class TrackParams:
    maxsteps: int = 1000
    corsteps: int = 4
    endsteps: int = 8
    fabry_offset: int = 2
    maxstep: float = 0.1
    minstep: float = 1e-6
    pbeta: float = 0.5
    alpha: float = 1e-3
    tolcor: float = 1e-8
    tolend: float = 1e-28
    divergence: float = 1e8
    # predictor: the [pade_numdeg/M] Pade approximant (pade_dendeg M = 1, 2;
    # the reference's default [5/1]) or the Taylor polynomial (pade_dendeg 0,
    # degree pade_numdeg, deg - 2 if 0); pade_numdeg + M + 2 <= deg
    pade_numdeg: int = 5
    pade_dendeg: int = 1
    cbeta: float = 0.005

    def to_c(self):
        return PhcTrackParams(self.maxsteps, self.corsteps, self.endsteps, self.fabry_offset,
                              self.maxstep, self.minstep, self.pbeta, self.alpha, self.tolcor,
                              self.tolend, self.divergence, self.pade_numdeg, self.pade_dendeg,
                              self.cbeta)


def program_struct(prog):
    """ctypes phc_program over the numpy arrays of a Program (the arrays are
    kept alive on the returned object)."""
    keep = {}

    def ia(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        if a.size == 0:
            a = np.zeros(1, np.int32)
        keep[id(a)] = a
        return a.ctypes.data_as(P_int)

    def da(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep[id(a)] = a
        return a.ctypes.data_as(P_double)

    s = PhcProgram(
        prog.n, prog.deg, prog.ncomp, prog.wdeg, prog.G,
        prog.nprod, prog.nlev, prog.nslots, prog.nfactor, ia(prog.nodes), ia(prog.lev_ptr),
        len(prog.pairs) // 4, ia(prog.pairs),
        len(prog.term_pair), ia(prog.seg_ptr), ia(prog.term_pair), da(prog.term_coef),
        len(prog.jpairs) // 2, ia(prog.jpairs),
        len(prog.jterms), ia(prog.jent_ptr), ia(prog.jterms),
        da(prog.weights), da(prog.hsig),
        len(prog.hterms) // 3, ia(prog.hent_ptr), ia(prog.hterms),
        len(prog.hpairs) // 2, ia(prog.hpairs))
    s._keep = keep
    return s


def points_to_planes(x, K=1):
    """complex128 [npaths][n] -> [2K][npaths][n] limb planes (exact)."""
    x = np.asarray(x, dtype=np.complex128)
    out = np.zeros((2 * K,) + x.shape)
    out[0] = x.real
    out[K] = x.imag
    return out


def planes_to_points(p, K):
    """[2K][npaths][n] -> complex128 (rounded to double)."""
    return p[:K].sum(axis=0) + 1j * p[K:].sum(axis=0)


class DeviceProgram:
    """A compiled homotopy uploaded to the device (phc_tracker handle)."""

    def __init__(self, prog):
        lib = load()
        self.prog = prog
        self.cstruct = program_struct(prog)
        h = ctypes.c_void_p()
        check(lib.phc_tracker_create(ctypes.byref(h), ctypes.byref(self.cstruct)))
        self.handle = h

    def lds_bytes(self, K):
        """(program bytes, bytes per path) of LDS at precision K"""
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        check(load().phc_tracker_lds_bytes(self.handle, K, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def __del__(self):
        try:
            if self.handle:
                load().phc_tracker_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class TrackBuffers:
    """Device outputs of one tracking run."""

    def __init__(self, npaths, n, K):
        self.endp = DeviceArray((2 * K, npaths, n))
        self.tend = DeviceArray((npaths,))
        self.status = DeviceArray((npaths,), np.int32)
        self.nsteps = DeviceArray((npaths,), np.int32)
        self.nrej = DeviceArray((npaths,), np.int32)
        self.dx = DeviceArray((npaths,))
        self.work = DeviceArray((8, npaths), np.int32)

    def to_host(self):
        return dict(endp=self.endp.to_host(), tend=self.tend.to_host(),
                    status=self.status.to_host(), nsteps=self.nsteps.to_host(),
                    nrej=self.nrej.to_host(), dx=self.dx.to_host(), work=self.work.to_host())


class Tracker:
    """Batched tracker for one homotopy (predictor series degree deg)."""

    def __init__(self, hom, deg=8):
        self.hom = hom
        self.n = hom.n
        self.deg = deg
        self.prog = compile_homotopy(hom, deg)
        self.prog0 = compile_homotopy(hom, 0)
        self.dev = DeviceProgram(self.prog)
        self.dev0 = DeviceProgram(self.prog0)
        self.waves = {}  # precision -> waves per workgroup chosen by phcpack_amd.autotune

    def track_device(self, d_start, KS, npaths, K, bufs, params=None, waves=None, stream=None,
                     controller=None, d_tstart=None):
        """controller: a PhcController (or None for the plain step rule);
        waves: wavefronts per workgroup (default: the tuned choice, else 12);
        d_tstart: device array of start parameters (None: every path at 0).
        Calls on one tracker are ordered by the library: a call on another
        stream waits on the device for the tracker's previous call."""
        if waves is None:
            waves = self.waves.get(K, 12)
        p = (params or TrackParams()).to_c()
        cptr = ctypes.byref(controller) if controller is not None else None
        check(load().phc_track_from(self.dev.handle, K, npaths, d_start.ptr, KS,
                                    d_tstart.ptr if d_tstart is not None else None,
                                    bufs.endp.ptr, bufs.tend.ptr, bufs.status.ptr,
                                    bufs.nsteps.ptr, bufs.nrej.ptr, bufs.dx.ptr, bufs.work.ptr,
                                    ctypes.byref(p), cptr, waves, stream))

    def track(self, starts, K=2, params=None, waves=None, controller=None, tstart=None):
        """starts: complex128 [npaths][n] (or [2KS][npaths][n] planes);
        tstart: start parameter per path (None: 0)."""
        planes = starts if starts.dtype == np.float64 else points_to_planes(starts)
        KS = planes.shape[0] // 2
        npaths = planes.shape[1]
        d_start = DeviceArray.from_host(np.ascontiguousarray(planes))
        d_t = None if tstart is None else DeviceArray.from_host(
            np.ascontiguousarray(tstart, np.float64))
        bufs = TrackBuffers(npaths, self.n, K)
        self.track_device(d_start, KS, npaths, K, bufs, params, waves, controller=controller,
                          d_tstart=d_t)
        check(load().phc_synchronize())
        return bufs.to_host()

    def refine_device(self, d_x, KS, npaths, K, d_out, d_dx, d_res, d_st, iters, tol, t=1.0,
                      waves=8, stream=None, only=None):
        """only: device int array of tracker statuses; paths not at status 1
        are copied through unrefined"""
        check(load().phc_refine(self.dev0.handle, K, npaths, t, d_x.ptr, KS, d_out.ptr, iters,
                                tol, d_dx.ptr, d_res.ptr, d_st.ptr,
                                only.ptr if only is not None else None, waves, stream))

    def endgame_device(self, d_x, KS, npaths, d_tend, K, d_alpha, d_rho, d_rco, d_xmax,
                       waves=8, stream=None):
        """the ratio-method estimates at the paths' last points (phc_endgame)"""
        check(load().phc_endgame(self.dev.handle, K, npaths, d_x.ptr, KS, d_tend.ptr,
                                 d_alpha.ptr, d_rho.ptr, d_rco.ptr, d_xmax.ptr, waves, stream))

    def endgame(self, planes, tend, K=1, waves=8):
        """dict of alpha, rho, rco, xmax [npaths] at the points
        [2KS][npaths][n] and parameters tend [npaths] (phc_endgame)"""
        KS, npaths = planes.shape[0] // 2, planes.shape[1]
        d_x = DeviceArray.from_host(np.ascontiguousarray(planes, np.float64))
        d_t = DeviceArray.from_host(np.ascontiguousarray(tend, np.float64))
        out = {k: DeviceArray((npaths,)) for k in ("alpha", "rho", "rco", "xmax")}
        self.endgame_device(d_x, KS, npaths, d_t, K, out["alpha"], out["rho"], out["rco"],
                            out["xmax"], waves)
        check(load().phc_synchronize())
        return {k: v.to_host() for k, v in out.items()}

    def refine(self, planes, K, iters, tol, t=1.0, waves=8):
        """Newton at t in precision K from [2KS][npaths][n] planes."""
        KS = planes.shape[0] // 2
        npaths = planes.shape[1]
        d_x = DeviceArray.from_host(planes)
        d_out = DeviceArray((2 * K, npaths, self.n))
        d_dx, d_res = DeviceArray((npaths,)), DeviceArray((npaths,))
        d_st = DeviceArray((npaths,), np.int32)
        self.refine_device(d_x, KS, npaths, K, d_out, d_dx, d_res, d_st, iters, tol, t, waves)
        check(load().phc_synchronize())
        return dict(x=d_out.to_host(), dx=d_dx.to_host(), res=d_res.to_host(),
                    status=d_st.to_host())


# statuses of paths that stopped inside [0, 1) for numerical reasons: step
# below minstep, maxsteps reached, singular Jacobian (not divergence)
INTERIOR_FAILURES = (2, 3, 4)


def interior_failures(status, tend, t_interior=0.99):
    """indices of the paths of a tracking run that failed at t < t_interior
    with a status of INTERIOR_FAILURES.  With the gamma trick the paths are
    regular for t < 1, so such a failure is one of precision or step
    control; failures close to t = 1 belong to singular or diverging
    endpoints, which more precision in the tracker does not repair."""
    status = np.asarray(status)
    return np.nonzero(np.isin(status, INTERIOR_FAILURES) & (np.asarray(tend) < t_interior))[0]


def escalate_to_dd(tr, start_planes, status, tend, params=None, t_interior=0.99, endp=None):
    """Precision escalation of the mixed method: the paths that FP64 tracking
    could not finish inside the interval (interior_failures) are tracked on
    in double-double with the plain step rule.  With endp (the FP64 run's
    last points, host planes or a DeviceArray: a stopped path ends at its
    last accepted point, at t = tend) the dd run continues each path from
    there, and the endgame estimates at those points first pass over the
    stops that a singularity of the end explains (a pole or a branch point
    at t = 1: classify_endpoints), which more precision would not finish
    either; without endp the paths are tracked again from start_planes
    (host [2KS][npaths][n]) at t = 0.  Returns (indices, tracker output of
    those paths in double-double, or None when there are none)."""
    status = np.asarray(status)
    tend = np.asarray(tend)
    idx = interior_failures(status, tend, t_interior)
    ep = None
    if len(idx) and endp is not None:
        ep = endp.to_host() if hasattr(endp, "to_host") else np.asarray(endp)
        est = tr.endgame(np.ascontiguousarray(ep[:, idx]), tend[idx], K=1)
        cls, _ = classify_endpoints(status[idx], tend[idx], est)
        idx = idx[cls == END_UNKNOWN]
    if len(idx) == 0:
        return idx, None
    prm = dataclasses.replace(params or TrackParams(), tolend=1e-28)
    if ep is None:
        return idx, tr.track(np.ascontiguousarray(start_planes[:, idx]), K=2, params=prm)
    return idx, tr.track(np.ascontiguousarray(ep[:, idx]), K=2, params=prm, tstart=tend[idx])


# classes of path ends
END_UNKNOWN, END_REGULAR, END_SINGULAR, END_INFINITY = 0, 1, 2, 3
END_NAMES = {0: "undetermined", 1: "regular", 2: "singular", 3: "at infinity"}


def classify_endpoints(status, tend, est, rco_tol=1e-8, big=1e8, near=0.25):
    """Where each path went as t -> 1, from the tracker's status and end
    parameter and the phc_endgame estimates est (alpha, rho, rco, xmax).
    Returns (class [npaths] in END_*, winding number [npaths], 0 when not
    estimated):
      regular      the tracker finished and rco > rco_tol;
      at infinity  the tracker reported divergence, or |x| > big, or the
                   nearest singularity is a pole (alpha < -1/4) at t = 1;
      singular     a branch point at t = 1 (0 < alpha < 0.9: winding number
                   round(1/alpha)), or a finished path with rco <= rco_tol;
    the singularity counts as at t = 1 when |rho - (1 - tend)| <= near
    (1 - tend).  Otherwise undetermined (a path stopped by step control away
    from any singularity of its end, which more precision may finish)."""
    status = np.asarray(status)
    tend = np.asarray(tend, np.float64)
    alpha, rho, rco, xmax = (np.asarray(est[k], np.float64) for k in ("alpha", "rho", "rco", "xmax"))
    n = len(status)
    cls = np.zeros(n, np.int32)
    wind = np.zeros(n, np.int32)
    dist = 1.0 - tend
    with np.errstate(invalid="ignore", divide="ignore"):
        at_end = np.isfinite(rho) & (dist > 0) & (np.abs(rho - dist) <= near * dist)
        branch = at_end & (alpha > 0.0) & (alpha < 0.9)
        pole = at_end & (alpha < -0.25)
        w = np.where(branch, np.rint(1.0 / np.where(branch, alpha, 1.0)), 0).astype(np.int32)
    ok = status == 1
    cls[ok & (rco > rco_tol)] = END_REGULAR
    inf = (status == 5) | (xmax > big) | (pole & ~ok)
    cls[inf & (cls == 0)] = END_INFINITY
    sing = (branch & ~ok) | (ok & ~(rco > rco_tol))
    cls[sing & (cls == 0)] = END_SINGULAR
    wind[cls == END_SINGULAR] = w[cls == END_SINGULAR]
    wind[cls == END_REGULAR] = 1
    return cls, wind
