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
Run: python tests/golden/make_golden.py [newton]
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


def main():
    r = ref()
    if r is None:
        raise SystemExit("oracle/_ref/libphcref.so not built (needs the reference checkout)")
    if sys.argv[1:] == ["newton"]:
        newton()
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
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
