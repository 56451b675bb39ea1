"""Host mirror of the reference's series entry points over the C-ABI.

Names follow the reference's host functions; data are numpy arrays in the
limb-planar layout of include/phcmd.h ([P][...] with P = K or 2K planes).

  product(K, cplx, x, y)        GPU_dbl2_product  (Convolutions/dbl2_convolutions_kernels.cu:372)
  PolyPlan.evaldiff(inputs)     GPU_dbl2_poly_evaldiff (Polynomials/dbl2_polynomials_kernels.cu:1577)
"""
import ctypes

import numpy as np

from ._lib import DeviceArray, check, load, ptr_d, ptr_i


def md_batch(op, K, a, b=None):
    """Elementwise op (0 add, 1 sub, 2 mul, 3 div, 4 sqrt) on [K][n] arrays."""
    lib = load()
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[1]
    da = DeviceArray.from_host(a)
    db = DeviceArray.from_host(np.ascontiguousarray(b, dtype=np.float64)) if b is not None else da
    dc = DeviceArray((K, n))
    check(lib.phc_md_batch(op, K, n, da.ptr, db.ptr, dc.ptr, None))
    check(lib.phc_synchronize())
    return dc.to_host()


def product(K, cplx, x, y):
    """Truncated products of power series.

    x, y: [P][nbatch][D] (or [P][D] for one product); returns z of the same shape.
    """
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    single = x.ndim == 2
    if single:
        x, y = x[:, None, :], y[:, None, :]
    P, nb, D = x.shape
    assert P == (2 * K if cplx else K) and y.shape == x.shape
    dx, dy = DeviceArray.from_host(x), DeviceArray.from_host(y)
    dz = DeviceArray(x.shape)
    check(lib.phc_series_product(K, int(bool(cplx)), D - 1, nb, dx.ptr, dy.ptr, dz.ptr, None))
    check(lib.phc_synchronize())
    z = dz.to_host()
    return z[:, 0, :] if single else z


class PolyPlan:
    """Device-resident polynomial p = cst + sum_i cff_i prod_{j in idx_i} x_j.

    supports: list of variable-index lists (distinct indices per monomial);
    cst: [P][D]; cff: [P][nbr][D].
    """

    def __init__(self, K, cplx, dim, deg, supports, cst, cff):
        lib = load()
        self.K, self.cplx, self.dim, self.deg = K, int(bool(cplx)), dim, deg
        self.P = 2 * K if cplx else K
        nvr = np.array([len(s) for s in supports], dtype=np.int32)
        idx = np.array([j for s in supports for j in s] or [0], dtype=np.int32)
        cst = np.ascontiguousarray(cst, dtype=np.float64)
        cff = np.ascontiguousarray(cff, dtype=np.float64)
        assert cst.shape == (self.P, deg + 1)
        assert cff.shape == (self.P, len(supports), deg + 1)
        h = ctypes.c_void_p()
        check(lib.phc_poly_create(ctypes.byref(h), K, self.cplx, dim, len(supports), deg,
                                  ptr_i(nvr), ptr_i(idx), ptr_d(cst), ptr_d(cff)))
        self.handle = h

    def evaldiff_device(self, nbatch, d_in, d_out, stream=None):
        check(load().phc_poly_evaldiff(self.handle, nbatch, d_in.ptr, d_out.ptr, stream))

    def evaldiff(self, inputs):
        """inputs [P][nbatch][dim][D] -> outputs [P][nbatch][dim+1][D]."""
        inputs = np.ascontiguousarray(inputs, dtype=np.float64)
        P, nb, dim, D = inputs.shape
        assert P == self.P and dim == self.dim and D == self.deg + 1
        d_in = DeviceArray.from_host(inputs)
        d_out = DeviceArray((P, nb, dim + 1, D))
        self.evaldiff_device(nb, d_in, d_out)
        check(load().phc_synchronize())
        return d_out.to_host()

    def __del__(self):
        try:
            if self.handle:
                load().phc_poly_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
