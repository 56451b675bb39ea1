"""The oracle's series products and polynomial evaluation/differentiation:
bit-exact against the reference's golden vectors and the compiled reference,
and checked against closed forms (exp(t)exp(-t) = 1, exact convolutions,
finite differences)."""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oraclelib import GOLDEN, oracle, pd, pi, rand_cmd, rand_md, ref


def product(lib, K, cplx, x, y):
    z = np.zeros_like(x)
    lib.osr_product(K, cplx, x.shape[1] - 1, pd(x), pd(y), pd(z)) if lib is oracle() else \
        lib.ref_product(K, cplx, x.shape[1] - 1, pd(x), pd(y), pd(z))
    return z


@pytest.mark.parametrize("K", [2, 4, 8])
@pytest.mark.parametrize("cplx", [0, 1])
def test_product_golden(K, cplx):
    g = np.load(os.path.join(GOLDEN, "products.npz"))
    x, y = g[f"x{K}_{cplx}"], g[f"y{K}_{cplx}"]
    assert np.array_equal(product(oracle(), K, cplx, x, y), g[f"z{K}_{cplx}"])


def poly_oracle(K, cplx, nvr, idx, cst, cff, inp):
    P, dim, D = inp.shape
    ptr = np.concatenate([[0], np.cumsum(nvr)]).astype(np.int32)
    out = np.zeros((P, dim + 1, D))
    oracle().osr_poly_evaldiff(K, cplx, dim, len(nvr), D - 1, pi(nvr), pi(ptr), pi(idx),
                               pd(cst), pd(cff), pd(inp), pd(out))
    return out


@pytest.mark.parametrize("K", [2, 4, 8])
@pytest.mark.parametrize("cplx", [0, 1])
def test_poly_golden(K, cplx):
    g = np.load(os.path.join(GOLDEN, "poly.npz"))
    t = f"{K}_{cplx}"
    out = poly_oracle(K, cplx, g["nvr"], g["idx"], g["cst" + t], g["cff" + t], g["inp" + t])
    assert np.array_equal(out, g["out" + t])


@pytest.mark.parametrize("K", [2, 4])
@pytest.mark.parametrize("cplx", [0, 1])
def test_poly_random_vs_compiled_reference(K, cplx):
    r = ref()
    if r is None:
        pytest.skip("compiled reference not available")
    rng = np.random.default_rng(5)
    dim, nbr, deg = 7, 12, 9
    nvr = rng.integers(1, 7, nbr).astype(np.int32)
    idx = np.concatenate([np.sort(rng.choice(dim, v, replace=False)) for v in nvr]).astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum(nvr)]).astype(np.int32)
    gen = (lambda s: rand_cmd(rng, K, s)) if cplx else (lambda s: rand_md(rng, K, s))
    cst, cff, inp = gen((deg + 1,)), gen((nbr, deg + 1)), gen((dim, deg + 1))
    o2 = np.zeros((cst.shape[0], dim + 1, deg + 1))
    r.ref_poly_evaldiff(K, cplx, dim, nbr, deg, pi(nvr), pi(ptr), pi(idx), pd(cst), pd(cff),
                        pd(inp), pd(o2))
    assert np.array_equal(poly_oracle(K, cplx, nvr, idx, cst, cff, inp), o2)


def exp_series(K, deg, sign):
    """exp(sign*t) with coefficients sign^k/k! computed in K-limb arithmetic"""
    o = oracle()
    s = np.zeros((K, deg + 1))
    c = np.zeros(K); c[0] = 1.0
    s[:, 0] = c
    for k in range(1, deg + 1):
        kk = np.zeros(K); kk[0] = float(k)
        o.omd_div(K, pd(c), pd(kk), pd(c))
        s[:, k] = c * (sign ** k)
    return s


@pytest.mark.parametrize("K,tol", [(2, 1e-26), (4, 1e-58), (8, 1e-120)])
def test_exp_times_expm(K, tol):
    """exp(t)*exp(-t) = 1 (Convolutions/dbl2_convolutions_testers.cpp:364-434)."""
    deg = 15
    z = product(oracle(), K, 0, exp_series(K, deg, 1), exp_series(K, deg, -1))
    total = z.sum(axis=0)
    assert abs(total[0] - 1.0) < tol
    assert np.all(np.abs(total[1:]) < tol)


def test_product_exact_convolution_small():
    """D <= 4: compare against the exact rational convolution."""
    rng = np.random.default_rng(9)
    for K in (2, 4):
        for D in (1, 2, 3, 4):
            x, y = rand_md(rng, K, (D,)), rand_md(rng, K, (D,))
            z = product(oracle(), K, 0, x, y)
            for k in range(D):
                ex = sum(sum(map(Fraction, x[:, i])) * sum(map(Fraction, y[:, k - i]))
                         for i in range(k + 1))
                assert abs(float(sum(map(Fraction, z[:, k])) - ex)) < 4 * (k + 1) * 2.0 ** (-52 * K)


def test_poly_gradient_by_central_differences():
    """Derivatives of the dd oracle agree with central differences of the
    value (error O(h^2) + rounding)."""
    rng = np.random.default_rng(4)
    K, dim, deg = 2, 5, 4
    supports = [[0, 1, 2], [1, 3], [4], [0, 2, 3, 4]]
    nvr = np.array([len(s) for s in supports], np.int32)
    idx = np.array([j for s in supports for j in s], np.int32)
    cst, cff, inp = rand_md(rng, K, (deg + 1,)), rand_md(rng, K, (len(supports), deg + 1)), \
        rand_md(rng, K, (dim, deg + 1))
    base = poly_oracle(K, 0, nvr, idx, cst, cff, inp)
    h = 1e-7
    for j in range(dim):
        ip, im = inp.copy(), inp.copy()
        ip[0, j, 0] += h
        im[0, j, 0] -= h
        vp = poly_oracle(K, 0, nvr, idx, cst, cff, ip)[:, dim, 0].sum()
        vm = poly_oracle(K, 0, nvr, idx, cst, cff, im)[:, dim, 0].sum()
        assert math.isclose((vp - vm) / (2 * h), base[:, j, 0].sum(), rel_tol=1e-6, abs_tol=1e-8)
