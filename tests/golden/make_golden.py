"""Writes the golden vectors of the arithmetic and series layers from the
compiled reference host code (oracle/_ref/libphcref.so, built by
`make -C oracle` from the reference sources) into tests/golden/*.npz.

Inputs are seeded; outputs are whatever the reference functions return:
  md_ops.npz     ddf_/qdf_/odf_ add, sub, mul, div, sqrt  (Norms/*_functions.cpp)
  products.npz   CPU_{dbl,cmplx}{2,4,8}_product            (Convolutions/*_convolutions_host.cpp)
  poly.npz       CPU_{dbl,cmplx}{2,4,8}_poly_evaldiff      (Polynomials/*_polynomials_host.cpp)
                 at the shape 8 variables, 8 monomials of 3 variables, degree 15.
  newton.npz     the reference's dd row-Newton test system (cyclic supports,
                 Newton/test_dbl2_newton.cpp:191-212, seed 7, dimension 8,
                 degree 15) and the series after each of its 8 staggered steps
                 (CPU_dbl2_poly_evaldiff + CPU_dbl2_qrbs_solve driven as in
                 Newton/dbl2_newton_method.cpp:1447-1494), with the tail
                 indices, skip counters and the last Q and R.
  columns.npz    the reference's dd column-Newton test systems
                 (Newton/test_dbl2_newton.cpp:93-189, dbl2_column_setup,
                 Newton/dbl2_newton_method.cpp:766-821): a monomial system
                 of a random unimodular exponent matrix (make_monomial_system,
                 seed 11, dimension 6, 10 row operations of multipliers below
                 3), the two-column lower/upper triangular system of
                 dimension 7 and the cyclic columns of dimension 6 with 6
                 columns, degree 15, and the series after each of their 12
                 staggered steps (CPU_dbl2_evaluate_monomials / _columns and
                 CPU_dbl2_qrbs_solve driven as test_dbl2_column_newton,
                 :1036-1287), with the tail indices, skip counters, Q and R.
Run: python tests/golden/make_golden.py [newton|columns]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oraclelib import pd, pi, rand_cmd, rand_md, ref  # noqa: E402

SEED = 20261202


def newton():
    from oraclelib import ref_cyclic_row_system, ref_row_newton
    sysm, sol, start = ref_cyclic_row_system(7, 8, 15)
    run = ref_row_newton(2, sysm, start, 8)
    np.savez_compressed(os.path.join(HERE, "newton.npz"), monoptr=sysm.monoptr, nvr=sysm.nvr,
                        idxptr=sysm.idxptr, idx=sysm.idx, cst=sysm.cst, cff=sysm.cff,
                        testsol=sol, start=start, history=run["history"], wdeg=run["wdeg"],
                        tail=run["tail"], up=run["up"], bs=run["bs"], Q=run["Q"], R=run["R"])


COLUMN_CASES = ("mono", "two", "cyclic")


def column_cases():
    """the three dd column systems of columns.npz (needs oracle/_ref)"""
    from oraclelib import monomial_columns, ref_dbl2_column_setup, ref_monomial_system
    A = ref_monomial_system(11, 6, 3, 10)
    yield "mono", ref_dbl2_column_setup(3, 6, 15, 1, rowsA=A)
    L = np.tril(np.ones((7, 7), np.int32))
    nvr, idx, _ = monomial_columns(L)
    ptr = np.concatenate([[0], np.cumsum(nvr)])
    # column 1 holds the rows of L in reverse order (test_dbl2_newton.cpp:154-160)
    nvr2 = np.concatenate([nvr, nvr[::-1]])
    idx2 = np.concatenate([idx] + [idx[ptr[i]:ptr[i + 1]] for i in range(6, -1, -1)])
    yield "two", ref_dbl2_column_setup(5, 7, 15, 2, nvr=nvr2, idx=idx2)
    yield "cyclic", ref_dbl2_column_setup(7, 6, 15, 6)


def columns():
    from oraclelib import ref_column_newton
    out = {}
    for name, (csys, sol, start) in column_cases():
        run = ref_column_newton(2, csys, start, 12)
        out[name + "_meta"] = np.array([csys.dim, csys.deg, csys.nbrcol], np.int32)
        out[name + "_nvr"], out[name + "_idx"] = csys.nvr, csys.idx
        if csys.exps is not None:
            out[name + "_exps"] = csys.exps
        if csys.cff is not None:
            out[name + "_cff"] = csys.cff
        out[name + "_mb"], out[name + "_testsol"], out[name + "_start"] = csys.mb, sol, start
        for k in ("history", "wdeg", "tail", "up", "bs", "Q", "R"):
            out[name + "_" + k] = run[k]
    np.savez_compressed(os.path.join(HERE, "columns.npz"), **out)


def main():
    r = ref()
    if r is None:
        raise SystemExit("oracle/_ref/libphcref.so not built (needs the reference checkout)")
    if sys.argv[1:] == ["newton"]:
        newton()
        return
    if sys.argv[1:] == ["columns"]:
        columns()
        return
    rng = np.random.default_rng(SEED)
    out = {}
    for K in (2, 4, 8):
        n = 64
        a = rand_md(rng, K, (n,), -10.0, 10.0)
        b = rand_md(rng, K, (n,), -10.0, 10.0)
        out[f"a{K}"], out[f"b{K}"] = a, b
        for op, name in enumerate(["add", "sub", "mul", "div", "sqrt"]):
            aa = np.abs(a) * np.sign(a[0]) if op == 4 else a
            aa = np.ascontiguousarray(aa)
            c = np.zeros((K, n))
            r.ref_md_batch(op, K, n, pd(aa), pd(b), pd(c))
            out[f"{name}{K}"] = c
    np.savez_compressed(os.path.join(HERE, "md_ops.npz"), **out)

    out = {}
    deg = 15
    for K in (2, 4, 8):
        for cplx in (0, 1):
            x = rand_cmd(rng, K, (deg + 1,)) if cplx else rand_md(rng, K, (deg + 1,))
            y = rand_cmd(rng, K, (deg + 1,)) if cplx else rand_md(rng, K, (deg + 1,))
            z = np.zeros_like(x)
            r.ref_product(K, cplx, deg, pd(x), pd(y), pd(z))
            out[f"x{K}_{cplx}"], out[f"y{K}_{cplx}"], out[f"z{K}_{cplx}"] = x, y, z
    np.savez_compressed(os.path.join(HERE, "products.npz"), **out)

    out = {}
    dim, nbr, deg = 8, 8, 15
    nvr = np.full(nbr, 3, dtype=np.int32)
    idx = np.concatenate([np.sort(rng.choice(dim, 3, replace=False)) for _ in range(nbr)]).astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum(nvr)]).astype(np.int32)
    out["nvr"], out["idx"] = nvr, idx
    for K in (2, 4, 8):
        for cplx in (0, 1):
            gen = (lambda s: rand_cmd(rng, K, s)) if cplx else (lambda s: rand_md(rng, K, s))
            cst, cff, inp = gen((deg + 1,)), gen((nbr, deg + 1)), gen((dim, deg + 1))
            P = cst.shape[0]
            o = np.zeros((P, dim + 1, deg + 1))
            r.ref_poly_evaldiff(K, cplx, dim, nbr, deg, pi(nvr), pi(ptr), pi(idx), pd(cst),
                                pd(cff), pd(inp), pd(o))
            tag = f"{K}_{cplx}"
            out["cst" + tag], out["cff" + tag], out["inp" + tag], out["out" + tag] = cst, cff, inp, o
    np.savez_compressed(os.path.join(HERE, "poly.npz"), **out)
    newton()
    columns()
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
