"""GPU parity of the arithmetic, series product and polynomial eval/diff
kernels against the oracle and the reference's golden vectors.  The device
code follows the reference's operation order with fused two_prod, so every
result must be bit-identical."""
import os

import numpy as np
import pytest

from oraclelib import GOLDEN, oracle, pd, pi, rand_cmd, rand_md

pytestmark = pytest.mark.gpu

OPS = ["add", "sub", "mul", "div", "sqrt"]


@pytest.mark.parametrize("K", [2, 4, 8])
@pytest.mark.parametrize("op", range(5))
def test_md_batch_golden(K, op):
    from phcpack_amd.series import md_batch
    g = np.load(os.path.join(GOLDEN, "md_ops.npz"))
    a, b = g[f"a{K}"], g[f"b{K}"]
    if op == 4:
        a = np.ascontiguousarray(np.abs(a) * np.sign(a[0]))
    c = md_batch(op, K, a, b)
    assert np.array_equal(c, g[f"{OPS[op]}{K}"]), (K, op)


@pytest.mark.parametrize("K", [2, 4, 8])
def test_md_batch_random_vs_oracle(K):
    from phcpack_amd.series import md_batch
    rng = np.random.default_rng(21 + K)
    n = 4099
    a = rand_md(rng, K, (n,), -100, 100)
    b = rand_md(rng, K, (n,), -1, 1)
    for op in range(5):
        aa = np.ascontiguousarray(np.abs(a) * np.sign(a[0])) if op == 4 else a
        c = np.zeros((K, n))
        oracle().omd_batch(op, K, n, pd(aa), pd(b), pd(c))
        assert np.array_equal(md_batch(op, K, aa, b), c), (K, op)


@pytest.mark.parametrize("K", [1, 2, 4, 8])
@pytest.mark.parametrize("cplx", [0, 1])
def test_product_batch_vs_oracle(K, cplx):
    from phcpack_amd.series import product
    rng = np.random.default_rng(100 + K + cplx)
    for deg in (0, 8, 15, 31):
        nb = 37
        gen = (lambda s: rand_cmd(rng, K, s)) if cplx else (lambda s: rand_md(rng, K, s))
        x, y = gen((nb, deg + 1)), gen((nb, deg + 1))
        z = product(K, cplx, x, y)
        for b in range(nb):
            xb, yb = np.ascontiguousarray(x[:, b]), np.ascontiguousarray(y[:, b])
            zb = np.zeros_like(xb)
            oracle().osr_product(K, cplx, deg, pd(xb), pd(yb), pd(zb))
            assert np.array_equal(z[:, b], zb), (K, cplx, deg, b)


@pytest.mark.parametrize("K", [2, 4, 8])
@pytest.mark.parametrize("cplx", [0, 1])
def test_product_golden(K, cplx):
    from phcpack_amd.series import product
    g = np.load(os.path.join(GOLDEN, "products.npz"))
    assert np.array_equal(product(K, cplx, g[f"x{K}_{cplx}"], g[f"y{K}_{cplx}"]),
                          g[f"z{K}_{cplx}"])


@pytest.mark.parametrize("K", [2, 4, 8])
@pytest.mark.parametrize("cplx", [0, 1])
def test_poly_evaldiff_golden(K, cplx):
    """Config-1 shape: 8 variables, 8 monomials of 3 variables, degree 15."""
    from phcpack_amd.series import PolyPlan
    g = np.load(os.path.join(GOLDEN, "poly.npz"))
    t = f"{K}_{cplx}"
    nvr, idx = g["nvr"], g["idx"]
    ptr = np.concatenate([[0], np.cumsum(nvr)])
    supports = [list(idx[ptr[i]:ptr[i + 1]]) for i in range(len(nvr))]
    plan = PolyPlan(K, cplx, 8, 15, supports, g["cst" + t], g["cff" + t])
    out = plan.evaldiff(g["inp" + t][:, None])
    assert np.array_equal(out[:, 0], g["out" + t])


def oracle_poly(K, cplx, nvr, idx, cst, cff, inp):
    P, dim, D = inp.shape
    ptr = np.concatenate([[0], np.cumsum(nvr)]).astype(np.int32)
    out = np.zeros((P, dim + 1, D))
    oracle().osr_poly_evaldiff(K, cplx, dim, len(nvr), D - 1, pi(nvr), pi(ptr), pi(idx),
                               pd(cst), pd(cff), pd(np.ascontiguousarray(inp)), pd(out))
    return out


@pytest.mark.parametrize("K", [1, 2, 4, 8])
@pytest.mark.parametrize("cplx", [0, 1])
def test_poly_evaldiff_batch_vs_oracle(K, cplx):
    """Random supports with 1..6 variables per monomial, a batch of inputs."""
    from phcpack_amd.series import PolyPlan
    rng = np.random.default_rng(300 + K * 2 + cplx)
    dim, nbr, deg, nb = 7, 10, 8, 19
    nvr = rng.integers(1, 7, nbr).astype(np.int32)
    supports = [sorted(rng.choice(dim, v, replace=False).tolist()) for v in nvr]
    idx = np.array([j for s in supports for j in s], np.int32)
    gen = (lambda s: rand_cmd(rng, K, s)) if cplx else (lambda s: rand_md(rng, K, s))
    cst, cff, inp = gen((deg + 1,)), gen((nbr, deg + 1)), gen((nb, dim, deg + 1))
    plan = PolyPlan(K, cplx, dim, deg, supports, cst, cff)
    out = plan.evaldiff(inp)
    for b in range(nb):
        assert np.array_equal(out[:, b], oracle_poly(K, cplx, nvr, idx, cst, cff, inp[:, b])), b
