"""Host side of the linearized power-series Newton method on the GPU.

Mirrors the reference's host interface for this path:
  houseqr        CPU_dbl2_factors_houseqr (Matrices/dbl2_factorizations.cpp:723)
  qrbs_solve     CPU_dbl2_qrbs_solve (Newton/dbl2_bals_host.cpp:479) with its
                 zeroQ / noqr / upidx / bsidx / newtail flags
  linear_residue CPU_dbl2_linear_residue (Newton/dbl2_bals_host.cpp:548)
  row_newton     dbl2_row_newton_qrstep (Newton/dbl2_newton_method.cpp:429)
                 iterated with staggered degrees as test_dbl2_row_newton
                 (:1447-1494); exponents > 1 by the common-factor trick of
                 CPU_dbl2_evaluate_monomials (Newton/dbl2_systems_host.cpp:12-111)
  column_newton  dbl2_column_newton_qrstep (Newton/dbl2_newton_method.cpp:259)
                 iterated as test_dbl2_column_newton (:1036-1287): monomial
                 systems (CPU_dbl2_evaluate_monomials) and column systems
                 (CPU_dbl2_evaluate_columns, :232-329)
for K = 1, 2, 4 limbs (dbl, dbl2, dbl4), batched: every array carries a batch
index after the limb index.  With cplx=True the complex twins
(CPU_cmplx2_factors_houseqr, CPU_cmplx2_qrbs_solve, CPU_cmplx2_linear_residue,
cmplx2_row_newton_qrstep, Newton/cmplx2_newton_method.cpp:565) run on
complex arrays of 2K planes (K real, then K imaginary), K = 2, 4 (8 for
the QR, the solve and the residual).  Arrays are numpy float64 in the limb-planar
layout of include/phcmd.h; the work runs on the GPU through libphcmd.so.
"""
import ctypes

import numpy as np

from ._lib import DeviceArray, P_int, check, load


def _ip(a):
    return np.ascontiguousarray(a, np.int32)


def houseqr(K, A, cplx=False):
    """Q, R of every n x n matrix of A [P][nbatch][n][n] (P = K, or 2K complex)"""
    A = np.ascontiguousarray(A, np.float64)
    _, nb, n, _ = A.shape
    dA = DeviceArray.from_host(A)
    dQ, dR = DeviceArray(A.shape), DeviceArray(A.shape)
    fn = load().phc_cmplx_houseqr_batch if cplx else load().phc_houseqr_batch
    check(fn(K, n, nb, dA.ptr, dQ.ptr, dR.ptr, None))
    return dQ.to_host(), dR.to_host()


def qrbs_solve(K, mat, rhs, tailidx, sol=None, Q=None, R=None, zeroQ=1, noqr=0, cplx=False):
    """block-Toeplitz solve of every batch entry: mat [P][nb][degp1][n][n],
    rhs [P][nb][degp1][n] (P = K, or 2K complex).  Returns dict(rhs, sol,
    Q, R, zeroQ, noqr, up, bs, newtail) with the flags as int arrays [nb]."""
    mat = np.ascontiguousarray(mat, np.float64)
    P, nb, degp1, n, _ = mat.shape
    vshape, qshape = (P, nb, degp1, n), (P, nb, n, n)
    d = {k: DeviceArray.from_host(np.zeros(s) if v is None else np.ascontiguousarray(v, np.float64))
         for k, v, s in (("rhs", rhs, vshape), ("sol", sol, vshape), ("Q", Q, qshape),
                         ("R", R, qshape))}
    flags = np.zeros((nb, 5), np.int32)
    flags[:, 0], flags[:, 1], flags[:, 4] = zeroQ, noqr, tailidx
    dm, df = DeviceArray.from_host(mat), DeviceArray.from_host(flags)
    fn = load().phc_cmplx_qrbs_solve_batch if cplx else load().phc_qrbs_solve_batch
    check(fn(K, n, degp1, tailidx, nb, dm.ptr, d["rhs"].ptr, d["sol"].ptr, d["Q"].ptr,
             d["R"].ptr, df.ptr, None))
    f = df.to_host()
    out = {k: v.to_host() for k, v in d.items()}
    out.update(zeroQ=f[:, 0], noqr=f[:, 1], up=f[:, 2], bs=f[:, 3], newtail=f[:, 4])
    return out


def linear_residue(K, mat, rhs, sol, tailidx, cplx=False):
    """residual vectors [P][nb][degp1][n] and the largest entries [K][nb]"""
    mat = np.ascontiguousarray(mat, np.float64)
    P, nb, degp1, n, _ = mat.shape
    dm = DeviceArray.from_host(mat)
    db, dx = DeviceArray.from_host(np.ascontiguousarray(rhs)), DeviceArray.from_host(np.ascontiguousarray(sol))
    dr = DeviceArray.zeros((P, nb, degp1, n))
    dmax = DeviceArray((K, nb))
    fn = load().phc_cmplx_linear_residue_batch if cplx else load().phc_linear_residue_batch
    check(fn(K, n, degp1, tailidx, nb, dm.ptr, db.ptr, dx.ptr, dr.ptr, dmax.ptr, None))
    return dr.to_host(), dmax.to_host()


class RowSystem:
    """Supports of a polynomial system on the device (phc_rowsys); exp: the
    exponent of every variable occurrence, parallel to idx (None: all 1),
    evaluated with the common-factor trick of CPU_dbl2_evaluate_monomials
    (Newton/dbl2_systems_host.cpp:12-111)."""

    def __init__(self, dim, monoptr, nvr, idxptr, idx, exp=None):
        self.dim = dim
        self.monoptr, self.nvr = _ip(monoptr), _ip(nvr)
        self.idxptr, self.idx = _ip(idxptr), _ip(idx if len(idx) else [0])
        self.exp = None if exp is None else _ip(exp if len(exp) else [1])
        self.nmono = int(self.monoptr[-1])
        h = ctypes.c_void_p()
        check(load().phc_rowsys_create_exp(
            ctypes.byref(h), dim, self.monoptr.ctypes.data_as(P_int),
            self.nvr.ctypes.data_as(P_int), self.idxptr.ctypes.data_as(P_int),
            self.idx.ctypes.data_as(P_int),
            None if self.exp is None else self.exp.ctypes.data_as(P_int)))
        self.handle = h

    def lds_bytes(self, K, deg):
        b = ctypes.c_size_t()
        check(load().phc_rowsys_lds_bytes(self.handle, K, deg, ctypes.byref(b)))
        return b.value

    def __del__(self):
        try:
            if self.handle:
                load().phc_rowsys_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def row_newton_device(self, K, deg, nbsteps, nsys, d_cst, cst_shared, d_cff, cff_shared, d_x,
                          logs=None, stream=None, cplx=False):
        """device-resident run; logs: dict of DeviceArrays (steps, wdeg, tail,
        up, bs, history, Q, R) or None"""
        g = (lambda k: logs[k].ptr if logs and k in logs else None)
        fn = load().phc_cmplx_row_newton if cplx else load().phc_row_newton
        check(fn(self.handle, K, deg, nbsteps, nsys, d_cst.ptr, int(cst_shared), d_cff.ptr,
                 int(cff_shared), d_x.ptr, g("steps"), g("wdeg"), g("tail"), g("up"), g("bs"),
                 g("history"), g("Q"), g("R"), stream))

    def row_newton(self, K, deg, nbsteps, cst, cff, x, history=False, cplx=False):
        """Staggered row Newton on a batch of systems with these supports.

        cst [P][nc][dim][deg+1], cff [P][nf][nmono][deg+1] (nc, nf = 1 for
        coefficients shared by the batch), x [P][nsys][dim][deg+1], P = K
        (or 2K complex).  Returns dict(x, steps, wdeg, tail, up, bs, Q,
        R[, history])."""
        x = np.ascontiguousarray(x, np.float64)
        P, nsys, dim, D = x.shape
        cst, cff = np.ascontiguousarray(cst, np.float64), np.ascontiguousarray(cff, np.float64)
        dc, df, dx = DeviceArray.from_host(cst), DeviceArray.from_host(cff), DeviceArray.from_host(x)
        logs = {k: DeviceArray.zeros((nsys, nbsteps), np.int32) for k in ("wdeg", "tail", "up", "bs")}
        logs["steps"] = DeviceArray.zeros((nsys,), np.int32)
        logs["Q"] = DeviceArray.zeros((P, nsys, dim, dim))
        logs["R"] = DeviceArray.zeros((P, nsys, dim, dim))
        if history:
            logs["history"] = DeviceArray.zeros((nbsteps, P, nsys, dim, D))
        self.row_newton_device(K, D - 1, nbsteps, nsys, dc, cst.shape[1] == 1 and nsys > 1, df,
                               cff.shape[1] == 1 and nsys > 1, dx, logs, cplx=cplx)
        check(load().phc_synchronize())
        out = {k: v.to_host() for k, v in logs.items()}
        out["x"] = dx.to_host()
        return out


class ColumnSystem(RowSystem):
    """A system in the column representation on the device (phc_colsys_create):
    nbrcol columns of dim monomials, monomial m = c*dim + j (equation j of
    column c) with nvr[m] variables (0: none), their indices idx
    concatenated in the order of m.  nbrcol = 1 is the monomial system
    x^A_i = mb_i (unit coefficients, exps parallel to idx); nbrcol > 1 has a
    coefficient series per monomial and the constant -1 in the last
    equation (CPU_dbl2_evaluate_columns, Newton/dbl2_systems_host.cpp:232)."""

    def __init__(self, dim, nbrcol, nvr, idx, exps=None):
        self.dim, self.nbrcol = dim, nbrcol
        self.nvr, self.idx = _ip(nvr), _ip(idx if len(idx) else [0])
        self.exp = None if exps is None else _ip(exps if len(exps) else [1])
        h = ctypes.c_void_p()
        check(load().phc_colsys_create(
            ctypes.byref(h), dim, nbrcol, self.nvr.ctypes.data_as(P_int),
            self.idx.ctypes.data_as(P_int),
            None if self.exp is None else self.exp.ctypes.data_as(P_int)))
        self.handle = h

    @classmethod
    def monomials(cls, rowsA):
        """the monomial system of the exponent matrix rowsA (row i: the
        nonzero entries in increasing column order, as make_monomial_system
        lists them, Newton/unimodular_matrices.cpp:150-164)"""
        A = np.asarray(rowsA, np.int64)
        if A.ndim != 2 or A.shape[0] != A.shape[1] or (A < 0).any():
            raise ValueError("rowsA must be a square matrix of nonnegative exponents")
        nvr = (A != 0).sum(axis=1)
        idx = np.concatenate([np.nonzero(r)[0] for r in A])
        exps = np.concatenate([r[r != 0] for r in A])
        return cls(A.shape[0], 1, nvr, idx, exps)

    def row_newton_device(self, *args, **kw):
        raise TypeError("a column system runs column_newton_device")

    def column_newton_device(self, K, deg, nbsteps, nsys, d_cff, cff_shared, d_mb, mb_shared, d_x,
                             logs=None, stream=None):
        """device-resident run (d_cff None for a monomial system); logs as
        RowSystem.row_newton_device"""
        g = (lambda k: logs[k].ptr if logs and k in logs else None)
        check(load().phc_column_newton(self.handle, K, deg, nbsteps, nsys,
                                       None if d_cff is None else d_cff.ptr, int(cff_shared),
                                       d_mb.ptr, int(mb_shared), d_x.ptr, g("steps"), g("wdeg"),
                                       g("tail"), g("up"), g("bs"), g("history"), g("Q"), g("R"),
                                       stream))

    def column_newton(self, K, deg, nbsteps, mb, x, cff=None, history=False):
        """Staggered column Newton on a batch of systems with this structure.

        mb [K][nm][dim][deg+1] (nm = 1: shared), cff [K][nf][nbrcol*dim][deg+1]
        (nf = 1: shared; None for a monomial system), x [K][nsys][dim][deg+1].
        Returns dict(x, steps, wdeg, tail, up, bs, Q, R[, history])."""
        x = np.ascontiguousarray(x, np.float64)
        P, nsys, dim, D = x.shape
        mb = np.ascontiguousarray(mb, np.float64)
        dm, dx = DeviceArray.from_host(mb), DeviceArray.from_host(x)
        df = None
        if cff is not None and self.nbrcol > 1:
            cff = np.ascontiguousarray(cff, np.float64)
            df = DeviceArray.from_host(cff)
        logs = {k: DeviceArray.zeros((nsys, nbsteps), np.int32) for k in ("wdeg", "tail", "up", "bs")}
        logs["steps"] = DeviceArray.zeros((nsys,), np.int32)
        logs["Q"] = DeviceArray.zeros((P, nsys, dim, dim))
        logs["R"] = DeviceArray.zeros((P, nsys, dim, dim))
        if history:
            logs["history"] = DeviceArray.zeros((nbsteps, P, nsys, dim, D))
        self.column_newton_device(K, D - 1, nbsteps, nsys, df,
                                  df is not None and cff.shape[1] == 1 and nsys > 1, dm,
                                  mb.shape[1] == 1 and nsys > 1, dx, logs)
        check(load().phc_synchronize())
        out = {k: v.to_host() for k, v in logs.items()}
        out["x"] = dx.to_host()
        return out
