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
