"""ctypes access to the CPU oracle (oracle/liboracle.so) and, when it was
built, to the compiled reference host code (oracle/_ref/libphcref.so).
Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libphcref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

P_d = ctypes.POINTER(ctypes.c_double)
P_i = ctypes.POINTER(ctypes.c_int)

_o = None
_r = None


def oracle():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        _o = ctypes.CDLL(ORACLE)
    return _o


def ref():
    """The compiled reference, or None where it was not built."""
    global _r
    if _r is None and os.path.exists(REF):
        _r = ctypes.CDLL(REF)
    return _r


def pd(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(P_d)


def pi(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(P_i)


def rand_md(rng, K, shape, lo=-1.0, hi=1.0):
    """Random normalized K-limb numbers [K][*shape], built like the
    reference's random_quad_double: r0 + r1*eps + r2*eps^2 + ... with
    renormalizing additions (oracle omd_add_d)."""
    n = int(np.prod(shape))
    out = np.zeros((K, n))
    out[0] = rng.uniform(lo, hi, n)
    if K > 1:
        o = oracle()
        eps = 2.0 ** -52
        buf = np.zeros(K)
        for j in range(n):
            buf[:] = 0.0
            buf[0] = out[0, j]
            for i in range(1, K):
                o.omd_add_d(K, pd(buf), ctypes.c_double(rng.uniform(-1, 1) * eps ** i), pd(buf))
            out[:, j] = buf
    return out.reshape((K,) + tuple(shape))


def rand_cmd(rng, K, shape):
    """complex: [2K][*shape] (real limbs then imaginary limbs)"""
    return np.concatenate([rand_md(rng, K, shape), rand_md(rng, K, shape)])


def ulp_err(a, b, K):
    """max |a-b| relative to |a| using the leading limbs, in units of eps^K"""
    diff = np.abs(np.sum(a - b, axis=0))
    scale = np.maximum(np.abs(a[0]), 1e-300)
    return float(np.max(diff / scale)) / (2.0 ** -52) ** K


# ----------------------------------------------------------------- tracker --

def _track_sigs():
    import ctypes as C
    from phcpack_amd._lib import PhcController, PhcProgram, PhcTrackParams
    o = oracle()
    o.otr_track.argtypes = [C.POINTER(PhcProgram), C.c_int, C.c_int, C.c_int, C.c_int, P_d, C.c_int,
                            P_d, P_d, P_d, P_i, P_i, P_i, P_d, P_i, C.POINTER(PhcTrackParams),
                            C.POINTER(PhcController)]
    o.otr_track_mt.argtypes = [C.POINTER(PhcProgram), C.c_int, C.c_int, P_d, C.c_int, P_d, P_d,
                               P_d, P_i, P_i, P_i, P_d, P_i, C.POINTER(PhcTrackParams),
                               C.POINTER(PhcController), C.c_int]
    o.otr_lg2.argtypes = [C.c_double]
    o.otr_lg2.restype = C.c_double
    o.otr_refine.argtypes = [C.POINTER(PhcProgram), C.c_int, C.c_int, C.c_double, P_d, C.c_int,
                             P_d, C.c_int, C.c_double, P_d, P_d, P_i]
    o.otr_eval.argtypes = [C.POINTER(PhcProgram), C.c_int, C.c_double, P_d, P_d, P_d]
    o.otr_series.argtypes = [C.POINTER(PhcProgram), C.c_int, C.c_double, P_d, P_d]
    o.otr_series.restype = C.c_int
    o.otr_endgame.argtypes = [C.POINTER(PhcProgram), C.c_int, C.c_int, P_d, C.c_int, P_d, P_d,
                              P_d, P_d, P_d]
    o.otr_pade_coefficients.argtypes = [C.c_int, P_d, C.c_int, C.c_int, P_d]
    o.otr_pade_error.argtypes = [C.c_int, P_d, C.c_int, C.c_int, P_d, P_d]
    o.otr_root.argtypes = [C.c_double, C.c_int]
    o.otr_root.restype = C.c_double
    o.otr_sigma_max_sq.argtypes = [C.c_int, P_d, P_d, P_d]
    o.otr_sigma_max_sq.restype = C.c_double
    o.otr_mixed_residual.argtypes = [C.POINTER(PhcProgram), C.c_int, C.c_double, P_d]
    o.otr_mixed_residual.restype = C.c_double
    o.otr_hessian_sum.argtypes = [C.POINTER(PhcProgram), C.c_int, C.c_double, P_d]
    o.otr_hessian_sum.restype = C.c_double
    o.otr_track_trace.argtypes = [C.POINTER(PhcProgram), C.c_int, C.c_int, C.c_int, P_d, C.c_int,
                                  P_d, P_d, P_i, P_i, P_i, P_d, P_i, C.POINTER(PhcTrackParams),
                                  C.POINTER(PhcController), P_d, C.c_int]
    o.otr_track_trace.restype = C.c_int
    o.otr_pade_evaluate.argtypes = [C.c_int, P_d, C.c_int, C.c_int, P_d, C.c_double, P_d]
    return o


class OracleController:
    """host-memory controller for the oracle (same layout as phc_controller);
    mode 1 needs a StepController, mode 2 allocates the step logs"""

    def __init__(self, mode, npaths, ctl=None, seed=0, cap=256):
        import ctypes as C
        from phcpack_amd._lib import PhcController
        from phcpack_amd.controller import NFEAT, StepController
        ctl = ctl or StepController()
        self.keep = []
        c = PhcController()
        c.mode, c.ncand = mode, ctl.C
        for i, m in enumerate(ctl.mults):
            c.mult[i] = m
        c.deg0 = ctl.deg0
        for i, d in enumerate(ctl.degrees):
            c.degree[i] = d
        for name, arr in (("weights", ctl.w), ("mu", ctl.mu), ("sd", ctl.sd)):
            a = np.ascontiguousarray(arr, np.float32)
            self.keep.append(a)
            setattr(c, name, a.ctypes.data)
        c.logit_threshold = ctl.logit_threshold
        c.seed = seed
        c.pole_cap = ctl.pole_cap
        self.cap = cap
        self.x = np.zeros((npaths, cap, NFEAT), np.float32)
        self.c = np.zeros((npaths, cap), np.int32)
        self.y = np.zeros((npaths, cap), np.int32)
        self.n = np.zeros(npaths, np.int32)
        c.log_capacity = cap
        c.log_x, c.log_c, c.log_y, c.log_n = (a.ctypes.data for a in (self.x, self.c, self.y, self.n))
        self.struct = c

    def logs(self):
        keep = np.arange(self.cap)[None, :] < self.n[:, None]
        return self.x[keep], self.c[keep], self.y[keep]


def oracle_track(prog, starts_planes, K, params, nthreads=1, paths=None, controller=None,
                 tstart=None):
    """Track with the C oracle; starts_planes [2KS][npaths][n];
    controller: an OracleController or None; tstart [npaths] the start
    parameter of every path (None: 0)."""
    import ctypes as C
    from phcpack_amd.tracker import program_struct
    o = _track_sigs()
    ps = program_struct(prog)
    st = np.ascontiguousarray(starts_planes, dtype=np.float64)
    KS, npaths, n = st.shape[0] // 2, st.shape[1], st.shape[2]
    out = dict(endp=np.zeros((2 * K, npaths, n)), tend=np.zeros(npaths),
               status=np.zeros(npaths, np.int32), nsteps=np.zeros(npaths, np.int32),
               nrej=np.zeros(npaths, np.int32), dx=np.zeros(npaths),
               work=np.zeros((8, npaths), np.int32))
    p = params.to_c()
    cp = C.byref(controller.struct) if controller is not None else None
    ts = None if tstart is None else np.ascontiguousarray(tstart, dtype=np.float64)
    tp = pd(ts) if ts is not None else None
    if nthreads > 1:
        o.otr_track_mt(C.byref(ps), K, npaths, pd(st), KS, tp, pd(out["endp"]), pd(out["tend"]),
                       pi(out["status"]), pi(out["nsteps"]), pi(out["nrej"]), pd(out["dx"]),
                       pi(out["work"]), C.byref(p), cp, nthreads)
    else:
        p0, p1 = (0, npaths) if paths is None else paths
        o.otr_track(C.byref(ps), K, npaths, p0, p1, pd(st), KS, tp, pd(out["endp"]),
                    pd(out["tend"]), pi(out["status"]), pi(out["nsteps"]), pi(out["nrej"]),
                    pd(out["dx"]), pi(out["work"]), C.byref(p), cp)
    return out


TRACE_FIELDS = ("t", "r", "pole", "curv", "eta", "nrm", "which", "rule", "step1", "step",
                "nfb", "ncr", "mix", "nit", "accepted", "sigma1", "hs2", "t_after", "cand",
                "degree")


def oracle_track_trace(prog, starts_planes, K, params, path=0, controller=None, cap=2000):
    """One path with the oracle's step log: (the outputs as oracle_track's,
    filled for that path only; per-step dicts with the fields of
    TRACE_FIELDS)"""
    import ctypes as C
    from phcpack_amd.tracker import program_struct
    o = _track_sigs()
    ps = program_struct(prog)
    st = np.ascontiguousarray(starts_planes, dtype=np.float64)
    KS, npaths, n = st.shape[0] // 2, st.shape[1], st.shape[2]
    out = dict(endp=np.zeros((2 * K, npaths, n)), tend=np.zeros(npaths),
               status=np.zeros(npaths, np.int32), nsteps=np.zeros(npaths, np.int32),
               nrej=np.zeros(npaths, np.int32), dx=np.zeros(npaths),
               work=np.zeros((8, npaths), np.int32))
    trace = np.zeros((cap, 20))
    p = params.to_c()
    cp = C.byref(controller.struct) if controller is not None else None
    nt = o.otr_track_trace(C.byref(ps), K, npaths, path, pd(st), KS, pd(out["endp"]),
                           pd(out["tend"]), pi(out["status"]), pi(out["nsteps"]),
                           pi(out["nrej"]), pd(out["dx"]), pi(out["work"]), C.byref(p), cp,
                           pd(trace), cap)
    steps = [dict(zip(TRACE_FIELDS, row)) for row in trace[:nt]]
    return out, steps


def oracle_root(x, k):
    return _track_sigs().otr_root(x, k)


def oracle_sigma_max(M):
    """sqrt of the oracle's estimate of the largest eigenvalue of M^H M"""
    M = np.ascontiguousarray(np.asarray(M, np.complex128))
    n = M.shape[0]
    a = np.ascontiguousarray(M.view(np.float64).reshape(-1))
    P, Q = np.zeros(2 * n * n), np.zeros(2 * n * n)
    return float(np.sqrt(_track_sigs().otr_sigma_max_sq(n, pd(a), pd(P), pd(Q))))


def oracle_mixed_residual(prog, x_planes, K, t):
    import ctypes as C
    from phcpack_amd.tracker import program_struct
    o = _track_sigs()
    ps = program_struct(prog)
    x = np.ascontiguousarray(x_planes, dtype=np.float64)
    return o.otr_mixed_residual(C.byref(ps), K, t, pd(x))


def oracle_hessian_sum(prog, x_planes, K, t):
    import ctypes as C
    from phcpack_amd.tracker import program_struct
    o = _track_sigs()
    ps = program_struct(prog)
    x = np.ascontiguousarray(x_planes, dtype=np.float64)
    return o.otr_hessian_sum(C.byref(ps), K, t, pd(x))


def oracle_pade_error(K, c, L, M):
    """c float64 [m][2K] with c[0] the constant term: (coefficients after
    the [L/M] construction, x0, leading limbs of the error estimate)"""
    o = _track_sigs()
    c = np.array(c, dtype=np.float64, order="C")
    x0 = c[0].copy()
    err = np.zeros(2)
    o.otr_pade_error(K, pd(c), L, M, pd(x0), pd(err))
    return c, x0, complex(err[0], err[1])


def oracle_pade(K, c, L, pden, s):
    """The tracker's predictor on one component: c float64 [m][2K] (per
    coefficient its K real limbs, then its K imaginary limbs).  Returns the
    coefficients after the [L/pden] construction (unchanged for pden 0),
    the constant term x0 and the predicted value [2K] at step s."""
    o = _track_sigs()
    c = np.array(c, dtype=np.float64, order="C")
    x0 = c[0].copy()
    if pden >= 1:
        o.otr_pade_coefficients(K, pd(c), L, pden, pd(x0))
    out = np.zeros(2 * K)
    o.otr_pade_evaluate(K, pd(c), L, pden, pd(x0), s, pd(out))
    return c, x0, out


def oracle_refine(prog, planes, K, iters, tol, t=1.0):
    import ctypes as C
    from phcpack_amd.tracker import program_struct
    o = _track_sigs()
    ps = program_struct(prog)
    x = np.ascontiguousarray(planes, dtype=np.float64)
    KS, npaths, n = x.shape[0] // 2, x.shape[1], x.shape[2]
    out = dict(x=np.zeros((2 * K, npaths, n)), dx=np.zeros(npaths), res=np.zeros(npaths),
               status=np.zeros(npaths, np.int32))
    o.otr_refine(C.byref(ps), K, npaths, t, pd(x), KS, pd(out["x"]), iters, tol, pd(out["dx"]),
                 pd(out["res"]), pi(out["status"]))
    return out


def oracle_eval(prog, x_planes, K, t):
    """value [2K][n] and Jacobian [2K][n*n] of the program at one point [2K][n]"""
    import ctypes as C
    from phcpack_amd.tracker import program_struct
    o = _track_sigs()
    ps = program_struct(prog)
    n = prog.n
    x = np.ascontiguousarray(x_planes, dtype=np.float64)
    val, jac = np.zeros((2 * K, n)), np.zeros((2 * K, n * n))
    o.otr_eval(C.byref(ps), K, t, pd(x), pd(val), pd(jac))
    return val, jac


def oracle_endgame(prog, planes, K, tend):
    """alpha, rho, rco, xmax [npaths] of otr_endgame at the points [2KS][npaths][n]"""
    import ctypes as C
    from phcpack_amd.tracker import program_struct
    o = _track_sigs()
    ps = program_struct(prog)
    x = np.ascontiguousarray(planes, dtype=np.float64)
    KS, npaths = x.shape[0] // 2, x.shape[1]
    t = np.ascontiguousarray(tend, dtype=np.float64)
    out = {k: np.zeros(npaths) for k in ("alpha", "rho", "rco", "xmax")}
    o.otr_endgame(C.byref(ps), K, npaths, pd(x), KS, pd(t), pd(out["alpha"]), pd(out["rho"]),
                  pd(out["rco"]), pd(out["xmax"]))
    return out


def oracle_series(prog, x_planes, K, t):
    import ctypes as C
    from phcpack_amd.tracker import program_struct
    o = _track_sigs()
    ps = program_struct(prog)
    x = np.ascontiguousarray(x_planes, dtype=np.float64)
    s = np.zeros((2 * K, prog.n, prog.deg + 1))
    ok = o.otr_series(C.byref(ps), K, t, pd(x), pd(s))
    return s, ok


# ------------------------------------------------------------------ Newton --

class RowSystem:
    """A real polynomial system for the row Newton method: polynomial i has
    monomials monoptr[i]..monoptr[i+1]-1, monomial m the nvr[m] distinct
    variables idx[idxptr[m]:idxptr[m+1]]; cst [K][dim][D], cff [K][nmono][D]."""

    def __init__(self, dim, deg, monoptr, nvr, idxptr, idx, cst, cff, exp=None):
        self.dim, self.deg = dim, deg
        self.monoptr = np.ascontiguousarray(monoptr, np.int32)
        self.nvr = np.ascontiguousarray(nvr, np.int32)
        self.idxptr = np.ascontiguousarray(idxptr, np.int32)
        self.idx = np.ascontiguousarray(idx, np.int32)
        self.cst = np.ascontiguousarray(cst, np.float64)
        self.cff = np.ascontiguousarray(cff, np.float64)
        # exponents parallel to idx (None: all ones)
        self.exp = None if exp is None else np.ascontiguousarray(exp, np.int32)

    @property
    def nmono(self):
        return int(self.monoptr[-1])


def ref_cyclic_row_system(seed, dim, deg):
    """the reference's own dd row-Newton test system (needs oracle/_ref)"""
    r = ref()
    nidx = ctypes.c_int()
    nmono = r.ref_dbl2_cyclic_row_system(ctypes.c_uint(seed), dim, deg, None, None, None, None,
                                         ctypes.byref(nidx), None, None, None, None)
    D = deg + 1
    monoptr = np.zeros(dim + 1, np.int32)
    nvr = np.zeros(nmono, np.int32)
    idxptr = np.zeros(nmono + 1, np.int32)
    idx = np.zeros(nidx.value, np.int32)
    cst = np.zeros((2, dim, D))
    cff = np.zeros((2, nmono, D))
    sol = np.zeros((2, dim, D))
    start = np.zeros((2, dim, D))
    r.ref_dbl2_cyclic_row_system(ctypes.c_uint(seed), dim, deg, pi(monoptr), pi(nvr), pi(idxptr),
                                 pi(idx), ctypes.byref(nidx), pd(cst), pd(cff), pd(sol), pd(start))
    return RowSystem(dim, deg, monoptr, nvr, idxptr, idx, cst, cff), sol, start


def _row_newton(lib, fn, K, sysm, start, nbsteps):
    dim, D = sysm.dim, sysm.deg + 1
    x = np.ascontiguousarray(start, np.float64).copy()
    wdeg, tail = np.zeros(nbsteps, np.int32), np.zeros(nbsteps, np.int32)
    up, bs = np.zeros(nbsteps, np.int32), np.zeros(nbsteps, np.int32)
    hist = np.zeros((nbsteps, K, dim, D))
    Q, R = np.zeros((K, dim, dim)), np.zeros((K, dim, dim))
    idx = sysm.idx if sysm.idx.size else np.zeros(1, np.int32)
    exp = None if sysm.exp is None else pi(sysm.exp)
    steps = getattr(lib, fn)(K, dim, sysm.deg, nbsteps, pi(sysm.monoptr), pi(sysm.nvr),
                             pi(sysm.idxptr), pi(idx), exp, pd(sysm.cst), pd(sysm.cff), pd(x),
                             pi(wdeg), pi(tail), pi(up), pi(bs), pd(hist), pd(Q), pd(R))
    return dict(x=x, steps=steps, wdeg=wdeg[:steps], tail=tail[:steps], up=up[:steps],
                bs=bs[:steps], history=hist[:steps], Q=Q, R=R)


def oracle_row_newton(K, sysm, start, nbsteps):
    return _row_newton(oracle(), "onw_row_newton", K, sysm, start, nbsteps)


def ref_row_newton(K, sysm, start, nbsteps):
    return _row_newton(ref(), "ref_row_newton", K, sysm, start, nbsteps)


class ColumnSystem:
    """A real system in the column representation of the column Newton
    method: nbrcol columns of dim monomials, monomial m = c*dim + j (equation
    j of column c) with nvr[m] variables idx[idxptr[m]:idxptr[m+1]] (0: none);
    nbrcol = 1 is a monomial system x^A_i = mb_i (unit coefficients, exps
    parallel to idx); nbrcol > 1 has cff [K][nbrcol*dim][D] and the constant
    -1 in the last equation; mb [K][dim][D] is the right-hand side."""

    def __init__(self, dim, deg, nbrcol, nvr, idx, mb, cff=None, exps=None):
        self.dim, self.deg, self.nbrcol = dim, deg, nbrcol
        self.nvr = np.ascontiguousarray(nvr, np.int32)
        self.idxptr = np.concatenate([[0], np.cumsum(self.nvr)]).astype(np.int32)
        self.idx = np.ascontiguousarray(idx, np.int32)
        self.mb = np.ascontiguousarray(mb, np.float64)
        self.cff = None if cff is None else np.ascontiguousarray(cff, np.float64)
        self.exps = None if exps is None else np.ascontiguousarray(exps, np.int32)
        assert self.nvr.size == nbrcol * dim and self.idx.size == self.idxptr[-1]

    def with_limbs(self, K):
        """the same system with K limbs (limbs beyond the given ones zero)"""
        def lk(a):
            if a is None:
                return None
            out = np.zeros((K,) + a.shape[1:])
            out[:min(K, a.shape[0])] = a[:K]
            return out
        return ColumnSystem(self.dim, self.deg, self.nbrcol, self.nvr, self.idx, lk(self.mb),
                            lk(self.cff), self.exps)


def monomial_columns(rowsA):
    """nvr, idx, exps of the monomial system with exponent matrix rowsA:
    row i's nonzero entries in increasing column order (the layout of
    make_monomial_system, Newton/unimodular_matrices.cpp:150-164)"""
    A = np.asarray(rowsA)
    nvr = (A != 0).sum(axis=1).astype(np.int32)
    idx = np.concatenate([np.nonzero(A[i])[0] for i in range(A.shape[0])]).astype(np.int32)
    exps = np.concatenate([A[i][A[i] != 0] for i in range(A.shape[0])]).astype(np.int32)
    return nvr, idx, exps


def cyclic_columns(dim, nbrcol):
    """nvr, idx of the cyclic columns: equation j < dim-1 of column c has the
    j+1 variables (c + k) mod dim, the last equation only in column 0 with
    all dim variables (make_cyclic_variables/_columns,
    Newton/cyclic_columns.cpp:10-23)"""
    nvr, idx = [], []
    for c in range(nbrcol):
        for j in range(dim):
            n = j + 1 if j < dim - 1 else (dim if c == 0 else 0)
            nvr.append(n)
            idx.extend((c + k) % dim for k in range(n))
    return np.array(nvr, np.int32), np.array(idx, np.int32)


def ref_monomial_system(seed, dim, size, nbritr):
    """rowsA [dim][dim] of the reference's make_monomial_system after
    srand(seed) (needs oracle/_ref)"""
    r = ref()
    A = np.zeros((dim, dim), np.int32)
    nvr = np.zeros(dim, np.int32)
    idx = np.zeros(dim * dim, np.int32)
    exp = np.zeros(dim * dim, np.int32)
    r.ref_monomial_system(ctypes.c_uint(seed), dim, size, nbritr, pi(A), pi(nvr), pi(idx), pi(exp))
    return A


def ref_dbl2_column_setup(seed, dim, deg, nbrcol, rowsA=None, nvr=None, idx=None):
    """the reference's dd column test setup (needs oracle/_ref): returns the
    ColumnSystem (nbrcol = 1: the monomial system of rowsA; 2: the caller's
    two columns; > 2: the cyclic columns), the test solution and the start"""
    r = ref()
    D = deg + 1
    nv = np.zeros(nbrcol * dim, np.int32)
    ix = np.zeros(nbrcol * dim * dim, np.int32)
    if nbrcol == 2:
        ptr = np.concatenate([[0], np.cumsum(nvr)])
        for m in range(2 * dim):
            nv[m] = nvr[m]
            ix[m * dim:m * dim + nvr[m]] = idx[ptr[m]:ptr[m + 1]]
    A = np.ascontiguousarray(rowsA if rowsA is not None else np.zeros((dim, dim)), np.int32)
    cff = np.zeros((2, nbrcol * dim, D))
    sol, mb, start = np.zeros((2, dim, D)), np.zeros((2, dim, D)), np.zeros((2, dim, D))
    r.ref_dbl2_column_setup(ctypes.c_uint(seed), dim, deg, nbrcol, pi(A), pi(nv), pi(ix), pd(cff),
                            pd(sol), pd(mb), pd(start))
    if nbrcol == 1:
        nvr_, idx_, exps = monomial_columns(A)
        csys = ColumnSystem(dim, deg, 1, nvr_, idx_, mb, exps=exps)
    else:
        idx_ = np.concatenate([ix[m * dim:m * dim + nv[m]] for m in range(nbrcol * dim)])
        csys = ColumnSystem(dim, deg, nbrcol, nv, idx_, mb, cff=cff)
    return csys, sol, start


def _column_newton(lib, fn, K, csys, start, nbsteps):
    dim, D = csys.dim, csys.deg + 1
    x = np.ascontiguousarray(start, np.float64).copy()
    wdeg, tail = np.zeros(nbsteps, np.int32), np.zeros(nbsteps, np.int32)
    up, bs = np.zeros(nbsteps, np.int32), np.zeros(nbsteps, np.int32)
    hist = np.zeros((nbsteps, K, dim, D))
    Q, R = np.zeros((K, dim, dim)), np.zeros((K, dim, dim))
    idx = csys.idx if csys.idx.size else np.zeros(1, np.int32)
    exps = None if csys.exps is None else pi(csys.exps)
    cff = None if csys.cff is None else pd(csys.cff)
    steps = getattr(lib, fn)(K, dim, csys.nbrcol, csys.deg, nbsteps, pi(csys.nvr),
                             pi(csys.idxptr), pi(idx), exps, cff, pd(csys.mb), pd(x), pi(wdeg),
                             pi(tail), pi(up), pi(bs), pd(hist), pd(Q), pd(R))
    return dict(x=x, steps=steps, wdeg=wdeg[:steps], tail=tail[:steps], up=up[:steps],
                bs=bs[:steps], history=hist[:steps], Q=Q, R=R)


def oracle_column_newton(K, csys, start, nbsteps):
    return _column_newton(oracle(), "onw_column_newton", K, csys, start, nbsteps)


def ref_column_newton(K, csys, start, nbsteps):
    return _column_newton(ref(), "ref_column_newton", K, csys, start, nbsteps)


def _houseqr(lib, fn, K, A):
    K_, m, n = A.shape
    Q, R = np.zeros((K, m, m)), np.zeros((K, m, n))
    getattr(lib, fn)(K, m, n, pd(np.ascontiguousarray(A)), pd(Q), pd(R))
    return Q, R


def oracle_houseqr(K, A):
    return _houseqr(oracle(), "onw_houseqr", K, A)


def ref_houseqr(K, A):
    return _houseqr(ref(), "ref_houseqr", K, A)


def _qrbs_solve(lib, fn, K, mat, rhs, tailidx, Q=None, R=None, zeroQ=1, noqr=0, sol=None):
    K_, degp1, dim, _ = mat.shape
    rhs = np.ascontiguousarray(rhs).copy()
    sol = np.zeros((K, degp1, dim)) if sol is None else sol.copy()
    Q = np.zeros((K, dim, dim)) if Q is None else Q.copy()
    R = np.zeros((K, dim, dim)) if R is None else R.copy()
    z, nq = ctypes.c_int(zeroQ), ctypes.c_int(noqr)
    up, bs, nt = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(tailidx)
    getattr(lib, fn)(K, dim, degp1, tailidx, pd(np.ascontiguousarray(mat)), pd(rhs), pd(sol),
                     pd(Q), pd(R), ctypes.byref(z), ctypes.byref(nq), ctypes.byref(up),
                     ctypes.byref(bs), ctypes.byref(nt))
    return dict(rhs=rhs, sol=sol, Q=Q, R=R, zeroQ=z.value, noqr=nq.value, up=up.value,
                bs=bs.value, newtail=nt.value)


def oracle_qrbs_solve(K, mat, rhs, tailidx, **kw):
    return _qrbs_solve(oracle(), "onw_qrbs_solve", K, mat, rhs, tailidx, **kw)


def ref_qrbs_solve(K, mat, rhs, tailidx, **kw):
    return _qrbs_solve(ref(), "ref_qrbs_solve", K, mat, rhs, tailidx, **kw)


def _residue(lib, fn, K, mat, rhs, sol, tailidx):
    K_, degp1, dim, _ = mat.shape
    rv = np.zeros((K, degp1, dim))
    rm = np.zeros(K)
    getattr(lib, fn)(K, dim, degp1, tailidx, pd(np.ascontiguousarray(mat)),
                     pd(np.ascontiguousarray(rhs)), pd(np.ascontiguousarray(sol)), pd(rv), pd(rm))
    return rv, rm


def oracle_linear_residue(K, mat, rhs, sol, tailidx):
    return _residue(oracle(), "onw_linear_residue", K, mat, rhs, sol, tailidx)


def ref_linear_residue(K, mat, rhs, sol, tailidx):
    return _residue(ref(), "ref_linear_residue", K, mat, rhs, sol, tailidx)


def random_supports(rng, dim, nper, nvar):
    """nper monomials of nvar distinct sorted variables in each of dim polynomials"""
    nmono = dim * nper
    monoptr = np.arange(0, nmono + 1, nper, dtype=np.int32)
    nvr = np.full(nmono, nvar, np.int32)
    idx = np.concatenate([np.sort(rng.choice(dim, nvar, replace=False))
                          for _ in range(nmono)]).astype(np.int32)
    idxptr = np.concatenate([[0], np.cumsum(nvr)]).astype(np.int32)
    return monoptr, nvr, idxptr, idx


def poly_value(K, sysm, x):
    """values [K][dim][D] of every polynomial at the series x (oracle)"""
    o = oracle()
    dim, D = sysm.dim, sysm.deg + 1
    val = np.zeros((K, dim, D))
    for i in range(dim):
        m0, m1 = sysm.monoptr[i], sysm.monoptr[i + 1]
        nv = np.ascontiguousarray(sysm.nvr[m0:m1])
        ip = np.ascontiguousarray(sysm.idxptr[m0:m1])
        cst = np.ascontiguousarray(sysm.cst[:, i])
        cff = np.ascontiguousarray(sysm.cff[:, m0:m1])
        out = np.zeros((K, dim + 1, D))
        idx = sysm.idx if sysm.idx.size else np.zeros(1, np.int32)
        exp = None if sysm.exp is None else pi(sysm.exp)
        o.osr_poly_evaldiff_exp(K, 0, dim, int(m1 - m0), sysm.deg, pi(nv), pi(ip), pi(idx), exp,
                                pd(cst), pd(cff), pd(np.ascontiguousarray(x)), pd(out))
        val[:, i] = out[:, dim]
    return val


def random_exponents(rng, sysm, values=(1, 2)):
    """exponents drawn uniformly from values, one per variable occurrence"""
    return rng.choice(np.array(values, np.int32), size=len(sysm.idx)).astype(np.int32)


def planted_row_system(rng, K, dim, deg, nper=8, nvar=3, supports=None, exp=None):
    """random K-limb coefficient series and an exponential-series root
    x_i(t) = c_i exp(r_i t) (c_i in [0.5, 1], r_i in [-1, 1]); the constants
    are set to minus the value of the other terms at that root.  Returns the
    system, the root [K][dim][D] and the start series (leading coefficients
    of the root plus 1e-10, higher coefficients zero)."""
    D = deg + 1
    monoptr, nvr, idxptr, idx = supports or random_supports(rng, dim, nper, nvar)
    nmono = int(monoptr[-1])
    cff = rand_md(rng, K, (nmono, D))
    c = rng.uniform(0.5, 1.0, dim)
    r = rng.uniform(-1.0, 1.0, dim)
    fact = np.cumprod(np.concatenate([[1.0], np.arange(1, D)]))
    root = np.zeros((K, dim, D))
    root[0] = c[:, None] * r[:, None] ** np.arange(D)[None, :] / fact[None, :]
    sysm = RowSystem(dim, deg, monoptr, nvr, idxptr, idx, np.zeros((K, dim, D)), cff, exp)
    sysm.cst = np.ascontiguousarray(-poly_value(K, sysm, root))
    start = np.zeros((K, dim, D))
    start[0, :, 0] = root[0, :, 0] + 1e-10
    return sysm, root, start


def unimodular_exponents(rng, dim, maxmul=2, nops=None):
    """a random unimodular matrix of nonnegative exponents: the identity
    after nops row operations row_a += m row_b (m in 1..maxmul), the
    construction of make_unimodular_matrix (Newton/unimodular_matrices.cpp:
    53-71) with Python's generator"""
    A = np.eye(dim, dtype=np.int64)
    for _ in range(dim if nops is None else nops):
        a, b = rng.choice(dim, 2, replace=False)
        A[a] += rng.integers(1, maxmul + 1) * A[b]
    return A.astype(np.int32)


def planted_column_system(rng, K, dim, deg, nbrcol, rowsA=None):
    """a column system with a planted root: nbrcol = 1 the monomial system of
    rowsA (default: unimodular_exponents), else the cyclic columns with
    random constant coefficients (as make_real2_coefficients sets only the
    leading coefficient); mb = the value of the system at the root
    x_i(t) = c_i exp(r_i t), computed by the row evaluator of the oracle
    (an independent code path).  Returns the system, the root and the
    start (the root's leading coefficients plus 1e-10)."""
    D = deg + 1
    c = rng.uniform(0.5, 1.0, dim)
    r = rng.uniform(-1.0, 1.0, dim)
    fact = np.cumprod(np.concatenate([[1.0], np.arange(1, D)]))
    root = np.zeros((K, dim, D))
    root[0] = c[:, None] * r[:, None] ** np.arange(D)[None, :] / fact[None, :]
    if nbrcol == 1:
        A = unimodular_exponents(rng, dim) if rowsA is None else np.asarray(rowsA, np.int32)
        nvr, idx, exps = monomial_columns(A)
        rows = RowSystem(dim, deg, np.arange(dim + 1), nvr,
                         np.concatenate([[0], np.cumsum(nvr)]), idx, np.zeros((K, dim, D)),
                         np.concatenate([np.ones((1, dim, 1)), np.zeros((1, dim, deg))], axis=2)
                         * (np.arange(K) == 0)[:, None, None], exps)
        mb = poly_value(K, rows, root)
        return ColumnSystem(dim, deg, 1, nvr, idx, mb, exps=exps), root, _start(root)
    nvr, idx = cyclic_columns(dim, nbrcol)
    cff = np.zeros((K, nbrcol * dim, D))
    cff[:, :, 0] = rand_md(rng, K, (nbrcol * dim,))
    ptr = np.concatenate([[0], np.cumsum(nvr)])
    monoptr, rnvr, ridx, rcff = [0], [], [], []
    for j in range(dim):
        for col in range(nbrcol):
            m = col * dim + j
            if nvr[m]:
                rnvr.append(nvr[m])
                ridx.extend(idx[ptr[m]:ptr[m + 1]])
                rcff.append(cff[:, m])
        monoptr.append(len(rnvr))
    cst = np.zeros((K, dim, D))
    cst[0, dim - 1, 0] = -1.0
    rows = RowSystem(dim, deg, monoptr, rnvr, np.concatenate([[0], np.cumsum(rnvr)]), ridx, cst,
                     np.stack(rcff, axis=1))
    mb = poly_value(K, rows, root)
    return ColumnSystem(dim, deg, nbrcol, nvr, idx, mb, cff=cff), root, _start(root)


def _start(root):
    start = np.zeros_like(root)
    start[0, :, 0] = root[0, :, 0] + 1e-10
    return start


# ------------------------------------------------------------ blocked QR --

def _cx_qr(lib, fn, K, A, szt, nbt):
    P, m, n = A.shape
    Q, R = np.zeros((P, m, m)), np.zeros((P, m, n))
    getattr(lib, fn)(K, m, n, szt, nbt, pd(np.ascontiguousarray(A).copy()), pd(Q), pd(R))
    return Q, R


def oracle_blocked_houseqr(K, A, szt, nbt):
    return _cx_qr(oracle(), "obq_blocked_houseqr", K, A, szt, nbt)


def ref_blocked_houseqr(K, A, szt, nbt):
    return _cx_qr(ref(), "ref_cmplx_blocked_houseqr", K, A, szt, nbt)


def oracle_real_blocked_houseqr(K, A, szt, nbt):
    """real A [K][m][n] -> Q [K][m][m], R [K][m][n]"""
    return _cx_qr(oracle(), "obq_real_blocked_houseqr", K, A, szt, nbt)


def ref_real_blocked_houseqr(K, A, szt, nbt):
    return _cx_qr(ref(), "ref_real_blocked_houseqr", K, A, szt, nbt)


def _cx_tiled(lib, fn, K, U, b, szt, nbt):
    U, b = np.ascontiguousarray(U).copy(), np.ascontiguousarray(b).copy()
    x = np.zeros_like(b)
    getattr(lib, fn)(K, b.shape[1], szt, nbt, pd(U), pd(b), pd(x))
    return x, U, b


def oracle_upper_tiled_solver(K, U, b, szt, nbt):
    return _cx_tiled(oracle(), "obq_upper_tiled_solver", K, U, b, szt, nbt)


def ref_upper_tiled_solver(K, U, b, szt, nbt):
    return _cx_tiled(ref(), "ref_cmplx_upper_tiled_solver", K, U, b, szt, nbt)


def oracle_qhb(K, Q, b):
    y = np.zeros_like(b)
    oracle().obq_qhb(K, b.shape[1], pd(np.ascontiguousarray(Q)), pd(np.ascontiguousarray(b)), pd(y))
    return y


def oracle_matvec(K, A, x):
    y = np.zeros((A.shape[0], A.shape[1]))
    oracle().obq_matvec(K, A.shape[1], A.shape[2], pd(np.ascontiguousarray(A)),
                        pd(np.ascontiguousarray(x)), pd(y))
    return y


def md_diff(K, a, b):
    """leading limb of a - b computed in K-limb arithmetic, elementwise, for
    real [K][...] or complex [2K][...] planar arrays"""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    P = a.shape[0]
    n = int(np.prod(a.shape[1:]))
    out = np.zeros((P // K,) + a.shape[1:])
    for h in range(P // K):
        aa = np.ascontiguousarray(a[h * K:(h + 1) * K].reshape(K, n))
        bb = np.ascontiguousarray(b[h * K:(h + 1) * K].reshape(K, n))
        c = np.zeros((K, n))
        oracle().omd_batch(1, K, n, pd(aa), pd(bb), pd(c))
        out[h] = c[0].reshape(a.shape[1:])
    return out


# --------------------------------------------------- complex power-series Newton --

def _cx_houseqr(lib, fn, K, A):
    P, m, n = A.shape
    Q, R = np.zeros((P, m, m)), np.zeros((P, m, n))
    getattr(lib, fn)(K, m, n, pd(np.ascontiguousarray(A)), pd(Q), pd(R))
    return Q, R


def oracle_cx_houseqr(K, A):
    return _cx_houseqr(oracle(), "onw_cx_houseqr", K, A)


def ref_cx_houseqr(K, A):
    return _cx_houseqr(ref(), "ref_cx_houseqr", K, A)


def _cx_qrbs_solve(lib, fn, K, mat, rhs, tailidx, Q=None, R=None, zeroQ=1, noqr=0, sol=None):
    P, degp1, dim, _ = mat.shape
    rhs = np.ascontiguousarray(rhs).copy()
    sol = np.zeros((P, degp1, dim)) if sol is None else sol.copy()
    Q = np.zeros((P, dim, dim)) if Q is None else Q.copy()
    R = np.zeros((P, dim, dim)) if R is None else R.copy()
    z, nq = ctypes.c_int(zeroQ), ctypes.c_int(noqr)
    up, bs, nt = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(tailidx)
    getattr(lib, fn)(K, dim, degp1, tailidx, pd(np.ascontiguousarray(mat)), pd(rhs), pd(sol),
                     pd(Q), pd(R), ctypes.byref(z), ctypes.byref(nq), ctypes.byref(up),
                     ctypes.byref(bs), ctypes.byref(nt))
    return dict(rhs=rhs, sol=sol, Q=Q, R=R, zeroQ=z.value, noqr=nq.value, up=up.value,
                bs=bs.value, newtail=nt.value)


def oracle_cx_qrbs_solve(K, mat, rhs, tailidx, **kw):
    return _cx_qrbs_solve(oracle(), "onw_cx_qrbs_solve", K, mat, rhs, tailidx, **kw)


def ref_cx_qrbs_solve(K, mat, rhs, tailidx, **kw):
    return _cx_qrbs_solve(ref(), "ref_cx_qrbs_solve", K, mat, rhs, tailidx, **kw)


def _cx_residue(lib, fn, K, mat, rhs, sol, tailidx):
    P, degp1, dim, _ = mat.shape
    rv = np.zeros((P, degp1, dim))
    rm = np.zeros(K)
    getattr(lib, fn)(K, dim, degp1, tailidx, pd(np.ascontiguousarray(mat)),
                     pd(np.ascontiguousarray(rhs)), pd(np.ascontiguousarray(sol)), pd(rv), pd(rm))
    return rv, rm


def oracle_cx_linear_residue(K, mat, rhs, sol, tailidx):
    return _cx_residue(oracle(), "onw_cx_linear_residue", K, mat, rhs, sol, tailidx)


def ref_cx_linear_residue(K, mat, rhs, sol, tailidx):
    return _cx_residue(ref(), "ref_cx_linear_residue", K, mat, rhs, sol, tailidx)


def _cx_row_newton(lib, fn, K, sysm, start, nbsteps):
    P, dim, D = start.shape
    x = np.ascontiguousarray(start).copy()
    wdeg, tail, up, bs = (np.zeros(nbsteps, np.int32) for _ in range(4))
    hist = np.zeros((nbsteps, P, dim, D))
    Q, R = np.zeros((P, dim, dim)), np.zeros((P, dim, dim))
    idx = sysm.idx if sysm.idx.size else np.zeros(1, np.int32)
    steps = getattr(lib, fn)(K, dim, sysm.deg, nbsteps, pi(sysm.monoptr), pi(sysm.nvr),
                             pi(sysm.idxptr), pi(idx), pd(sysm.cst), pd(sysm.cff), pd(x),
                             pi(wdeg), pi(tail), pi(up), pi(bs), pd(hist), pd(Q), pd(R))
    return dict(x=x, steps=steps, wdeg=wdeg[:steps], tail=tail[:steps], up=up[:steps],
                bs=bs[:steps], history=hist[:steps], Q=Q, R=R)


def oracle_cx_row_newton(K, sysm, start, nbsteps):
    return _cx_row_newton(oracle(), "onw_cx_row_newton", K, sysm, start, nbsteps)


def ref_cx_row_newton(K, sysm, start, nbsteps):
    return _cx_row_newton(ref(), "ref_cx_row_newton", K, sysm, start, nbsteps)


def cx_poly_value(K, sysm, x):
    """complex values [2K][dim][D] of every polynomial at the series x [2K][dim][D]"""
    o = oracle()
    dim, D = sysm.dim, sysm.deg + 1
    val = np.zeros((2 * K, dim, D))
    for i in range(dim):
        m0, m1 = sysm.monoptr[i], sysm.monoptr[i + 1]
        nv = np.ascontiguousarray(sysm.nvr[m0:m1])
        ip = np.ascontiguousarray(sysm.idxptr[m0:m1])
        cst = np.ascontiguousarray(sysm.cst[:, i])
        cff = np.ascontiguousarray(sysm.cff[:, m0:m1])
        out = np.zeros((2 * K, dim + 1, D))
        idx = sysm.idx if sysm.idx.size else np.zeros(1, np.int32)
        o.osr_poly_evaldiff_exp(K, 1, dim, int(m1 - m0), sysm.deg, pi(nv), pi(ip), pi(idx), None,
                                pd(cst), pd(cff), pd(np.ascontiguousarray(x)), pd(out))
        val[:, i] = out[:, dim]
    return val


def planted_cx_row_system(rng, K, dim, deg, nper=8, nvar=3, supports=None):
    """the complex twin of planted_row_system: complex K-limb coefficient
    series, the root x_i(t) = c_i exp(r_i t) with complex c_i (|c_i| in
    [0.5, 1]) and r_i, constants set so that it is a root; start = the
    leading coefficients plus 1e-10"""
    D = deg + 1
    monoptr, nvr, idxptr, idx = supports or random_supports(rng, dim, nper, nvar)
    nmono = int(monoptr[-1])
    cff = rand_cmd(rng, K, (nmono, D))
    c = rng.uniform(0.5, 1.0, dim) * np.exp(2j * np.pi * rng.uniform(0, 1, dim))
    r = rng.uniform(-1.0, 1.0, dim) + 1j * rng.uniform(-1.0, 1.0, dim)
    fact = np.cumprod(np.concatenate([[1.0], np.arange(1, D)]))
    z = c[:, None] * r[:, None] ** np.arange(D)[None, :] / fact[None, :]
    root = np.zeros((2 * K, dim, D))
    root[0], root[K] = z.real, z.imag
    sysm = RowSystem(dim, deg, monoptr, nvr, idxptr, idx, np.zeros((2 * K, dim, D)), cff, None)
    sysm.cst = np.ascontiguousarray(-cx_poly_value(K, sysm, root))
    start = np.zeros((2 * K, dim, D))
    start[0, :, 0] = root[0, :, 0] + 1e-10
    start[K, :, 0] = root[K, :, 0] + 1e-10
    return sysm, root, start
