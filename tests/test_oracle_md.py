"""The oracle's multiple-double arithmetic against the reference's golden
vectors (bit-exact), the compiled reference (when built), and closed forms."""
import ctypes
import os
from fractions import Fraction

import numpy as np
import pytest

from oraclelib import GOLDEN, oracle, pd, rand_md, ref

OPS = ["add", "sub", "mul", "div", "sqrt"]


def run_op(lib_fn, op, K, a, b):
    n = a.shape[1]
    c = np.zeros((K, n))
    lib_fn(op, K, n, pd(np.ascontiguousarray(a)), pd(np.ascontiguousarray(b)), pd(c))
    return c


@pytest.mark.parametrize("K", [2, 4, 8])
@pytest.mark.parametrize("op", range(5))
def test_md_golden_bitexact(K, op):
    g = np.load(os.path.join(GOLDEN, "md_ops.npz"))
    a, b = g[f"a{K}"], g[f"b{K}"]
    if op == 4:
        a = np.abs(a) * np.sign(a[0])
    c = run_op(oracle().omd_batch, op, K, a, b)
    assert np.array_equal(c, g[f"{OPS[op]}{K}"])


@pytest.mark.parametrize("K", [2, 4, 8])
def test_md_random_vs_compiled_reference(K):
    r = ref()
    if r is None:
        pytest.skip("compiled reference not available")
    rng = np.random.default_rng(7 + K)
    a = rand_md(rng, K, (500,), -1e3, 1e3)
    b = rand_md(rng, K, (500,), -1e-3, 1e-3)
    for op in range(5):
        aa = np.abs(a) * np.sign(a[0]) if op == 4 else a
        assert np.array_equal(run_op(oracle().omd_batch, op, K, aa, b),
                              run_op(r.ref_md_batch, op, K, aa, b))


def frac(v):
    return sum(Fraction(float(x)) for x in v)


@pytest.mark.parametrize("K,tol", [(2, 1e-30), (4, 1e-62), (8, 1e-124)])
def test_sqrt2_squared(K, tol):
    """sqrt(2)^2 = 2 (Norms/test_double_doubles.cpp:25-94 checks sqrt(2))."""
    o = oracle()
    two = np.zeros(K); two[0] = 2.0
    s = np.zeros(K); sq = np.zeros(K)
    o.omd_sqrt(K, pd(two), pd(s))
    o.omd_mul(K, pd(s), pd(s), pd(sq))
    assert abs(float(frac(sq) - 2)) < tol
    # and the limbs of sqrt(2) are the exact expansion: (sum s)^2 - 2 tiny
    assert abs(float(frac(s) ** 2 - 2)) < tol


def test_error_free_transformations():
    o = oracle()
    o.omd_two_sum.restype = ctypes.c_double
    o.omd_two_prod.restype = ctypes.c_double
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b = rng.standard_normal(2) * 10.0 ** rng.integers(-20, 20, 2)
        e = ctypes.c_double()
        s = o.omd_two_sum(ctypes.c_double(a), ctypes.c_double(b), ctypes.byref(e))
        assert Fraction(s) + Fraction(e.value) == Fraction(a) + Fraction(b)
        p = o.omd_two_prod(ctypes.c_double(a), ctypes.c_double(b), ctypes.byref(e))
        assert Fraction(p) + Fraction(e.value) == Fraction(a) * Fraction(b)


@pytest.mark.parametrize("K,tol", [(2, 1e-31), (4, 1e-63), (8, 1e-126)])
def test_mul_div_accuracy_exact_rationals(K, tol):
    """a*b and a/b agree with exact rational arithmetic to about eps^K."""
    rng = np.random.default_rng(11)
    a = rand_md(rng, K, (50,), 0.5, 2.0)
    b = rand_md(rng, K, (50,), 0.5, 2.0)
    m = run_op(oracle().omd_batch, 2, K, a, b)
    d = run_op(oracle().omd_batch, 3, K, a, b)
    for j in range(50):
        fa, fb = frac(a[:, j]), frac(b[:, j])
        assert abs(float(frac(m[:, j]) - fa * fb)) < tol * 8
        assert abs(float(frac(d[:, j]) - fa / fb)) < tol * 8
