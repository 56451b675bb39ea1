"""Host side of the complex and real blocked Householder QR and tiled back
substitution on the GPU (K = 2 double-double, 4 quad-double, 8 octo-double).

Mirrors CPU_cmplx4_blocked_houseqr (Matrices/dbl4_baqr_host.cpp:1206) and
CPU_cmplx4_upper_tiled_solver (Matrices/dbl4_tabs_host.cpp:345): the same
arguments (nrows, ncols, szt, nbt, the limb planes of A, Q, R; dim, szt,
nbt, U, b, x), with the limb planes stacked into one complex limb-planar
array [2K][rows][cols] (K real planes, then K imaginary planes).
"""
import ctypes

import numpy as np

from ._lib import DeviceArray, check, load


class BlockedQR:
    """Workspace of the blocked QR for nrows x ncols matrices in tiles of szt."""

    def __init__(self, K, nrows, ncols, szt):
        self.K, self.nrows, self.ncols, self.szt = K, nrows, ncols, szt
        self.nbt = ncols // szt
        h = ctypes.c_void_p()
        check(load().phc_baqr_create(ctypes.byref(h), K, nrows, ncols, szt))
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                load().phc_baqr_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def set_tuning(self, gemm_block=0, split_target=0):
        """launch configuration of the plan's matrix products
        (phc_baqr_set_tuning; 0, 0: the defaults)"""
        check(load().phc_baqr_set_tuning(self.handle, int(gemm_block), int(split_target)))

    def tuning(self):
        """(gemm_block, split_target) in use"""
        b, s = ctypes.c_int(), ctypes.c_int()
        check(load().phc_baqr_get_tuning(self.handle, ctypes.byref(b), ctypes.byref(s)))
        return b.value, s.value

    def factor_device(self, dA, dQ, dR, stream=None):
        check(load().phc_cmplx_blocked_houseqr(self.handle, dA.ptr, dQ.ptr, dR.ptr, stream))

    def factor_real_device(self, dA, dQ, dR, stream=None):
        """the real QR: dA [K][m][n], dQ [K][m][m], dR [K][m][n]"""
        check(load().phc_real_blocked_houseqr(self.handle, dA.ptr, dQ.ptr, dR.ptr, stream))

    def factor_real(self, A):
        """Q [K][m][m], R [K][m][n] of the real A [K][m][n]"""
        K, m, n = self.K, self.nrows, self.ncols
        dA = DeviceArray.from_host(np.ascontiguousarray(A, np.float64))
        dQ, dR = DeviceArray((K, m, m)), DeviceArray((K, m, n))
        self.factor_real_device(dA, dQ, dR)
        check(load().phc_synchronize())
        return dQ.to_host(), dR.to_host()

    def qtb_device(self, dQ, db, dy, stream=None):
        """y = Q^T b for the real Q (phc_real_qtb)"""
        check(load().phc_real_qtb(self.handle, dQ.ptr, db.ptr, dy.ptr, stream))

    def qhb_device(self, dQ, db, dy, stream=None):
        """y = Q^H b (GPU_cmplx4_bals_qhb, Newton/dbl4_bals_kernels.cu:602)"""
        check(load().phc_cmplx_qhb(self.handle, dQ.ptr, db.ptr, dy.ptr, stream))

    def factor(self, A):
        """Q [2K][m][m], R [2K][m][n] of A [2K][m][n]"""
        K, m, n = self.K, self.nrows, self.ncols
        dA = DeviceArray.from_host(np.ascontiguousarray(A, np.float64))
        dQ, dR = DeviceArray((2 * K, m, m)), DeviceArray((2 * K, m, n))
        self.factor_device(dA, dQ, dR)
        check(load().phc_synchronize())
        return dQ.to_host(), dR.to_host()


def upper_tiled_solver_device(K, dim, szt, nbt, dU, db, dx, stream=None):
    check(load().phc_cmplx_upper_tiled_solver(K, dim, szt, nbt, dU.ptr, db.ptr, dx.ptr, stream))


def upper_tiled_solver(K, U, b, szt, nbt):
    """x with U x = b; returns (x, U with inverted diagonal tiles, updated b)"""
    dim = szt * nbt
    dU = DeviceArray.from_host(np.ascontiguousarray(U, np.float64))
    db = DeviceArray.from_host(np.ascontiguousarray(b, np.float64))
    dx = DeviceArray.zeros((2 * K, dim))
    upper_tiled_solver_device(K, dim, szt, nbt, dU, db, dx)
    return dx.to_host(), dU.to_host(), db.to_host()


def gemm_device(K, conjA, conjB, M, N, Kd, A, B, C, accumulate=False, stream=None):
    """A, B, C: tuples (DeviceArray, element offset, s0, s1) of planar complex
    operands (plane stride = the array's plane size)"""
    def op(t):
        d, off, s0, s1 = t
        plane = d.nbytes // 8 // d.shape[0]
        return (ctypes.c_void_p(d.ptr.value + 8 * off), plane, s0, s1)
    a, b, c = op(A), op(B), op(C)
    check(load().phc_cmplx_gemm(K, int(conjA), int(conjB), M, N, Kd, *a, *b, *c, int(accumulate),
                                stream))


def qh_times(K, dQ, db, n, dy, stream=None):
    """y = Q^H b for Q [2K][n][n] and b [2K][n] on the device"""
    gemm_device(K, True, False, n, 1, n, (dQ, 0, 1, n), (db, 0, 1, 0), (dy, 0, 1, 0), False, stream)
