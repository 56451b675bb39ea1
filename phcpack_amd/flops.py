"""Algorithmic work of the tracker and refinement kernels.

Counts complex multiple-double operations per phase exactly from the
compiled program and the per-path work counters the tracker returns, then
converts them to FP64 floating-point operations with two tables:

* build: the FP64 instructions the error-free transformations of
  phcpack_amd/csrc/md.hpp issue (a fused multiply-add counts 2, a division
  1); two_prod is a multiply plus an FMA (3), two_sum 6, quick_two_sum 3.
* reference: the cost factors PHCpack reports with
  (Matrices/write_dbl2_qrtimeflops.cpp:49-58, write_dbl4_qrtimeflops.cpp:50-59,
  write_dbl8_qrtimeflops.cpp:50-59): dd add/mul/div 20/23/70,
  qd 89/336/893, od 270/1742/5126 (Dekker products).
"""

import numpy as np

TWO_SUM, QTS, TWO_PROD = 6, 3, 3


def _add_k(K):
    if K == 1:
        return 1
    if K == 2:
        return 2 * TWO_SUM + 1 + QTS + 1 + QTS
    ts = K + K * (K - 1) // 2
    return ts * TWO_SUM + K + K * QTS + (K - 1) * QTS


def _mul_k(K):
    if K == 1:
        return 1
    if K == 2:
        return TWO_PROD + 4 + QTS
    f = (K - 1) + (K - 2)  # plain products of level K
    for L in range(K - 1, -1, -1):
        tail = K - 1 - L
        f += TWO_PROD + tail * TWO_SUM + 1
        f += L * (TWO_PROD + tail * TWO_SUM + 1 + TWO_SUM + tail * TWO_SUM + 1)
    return f + K * QTS + (K - 1) * QTS


def _mul_d_k(K):
    if K == 1:
        return 1
    if K == 2:
        return TWO_PROD + 2 + QTS
    ts = (K - 1) * K // 2
    return K * TWO_PROD + ts * TWO_SUM + K + K * QTS + (K - 1) * QTS


def _div_k(K):
    if K == 1:
        return 1
    if K == 2:
        return 3 + 2 * (_mul_d_k(2) + _add_k(2)) + QTS + (TWO_SUM + 1 + QTS)
    return (K + 1) + K * (_mul_d_k(K) + _add_k(K)) + K * QTS + (K - 1) * QTS


BUILD = {K: dict(add=_add_k(K), mul=_mul_k(K), mul_d=_mul_d_k(K), div=_div_k(K))
         for K in (1, 2, 4, 8)}
REFERENCE = {1: dict(add=1, mul=1, mul_d=1, div=1),
             2: dict(add=20, mul=23, mul_d=23, div=70),
             4: dict(add=89, mul=336, mul_d=336, div=893),
             8: dict(add=270, mul=1742, mul_d=1742, div=5126)}


def complex_costs(K, table=BUILD):
    t = table[K]
    return dict(cmul=4 * t["mul"] + 2 * t["add"], cmac=4 * t["mul"] + 4 * t["add"],
                cadd=2 * t["add"], scale_d=2 * t["mul_d"],
                cdiv=6 * t["mul"] + 3 * t["add"] + 2 * t["div"])


def _groups(prog):
    """monomial groups of the mixed residual: runs of terms of one node in a
    segment"""
    nseg = len(prog.seg_ptr) - 1
    node = prog.pairs.reshape(-1, 4)[:, 3]
    g = 0
    for sgi in range(nseg):
        prev = None
        for t in range(prog.seg_ptr[sgi], prog.seg_ptr[sgi + 1]):
            m = node[prog.term_pair[t]]
            if m != prev:
                g += 1
            prev = m
    return g


def _sigma_flops(n):
    """FP64 flops of one largest-singular-value estimate of an n x n complex
    matrix: the upper triangle of the Gram matrix (n(n+1)/2 entries of n
    complex multiply-adds, 8 flops each), of six squarings whose operands
    are scaled as they are read (12 flops per multiply-add), traces, the
    column and the Rayleigh quotient"""
    nu = n * (n + 1) // 2
    return 8 * nu * n + 6 * 12 * nu * n + 7 * n + 10 * n * n + 6 * n + 2


def _phase_ops(prog):
    """complex operation counts of each phase of the tracker kernel:
    dict phase -> dict op -> count (cmac also stands for cmsub); "fp64"
    counts plain FP64 flops of the step-control quantities"""
    n, D, G = prog.n, prog.deg + 1, prog.G
    nprod = prog.nprod
    npair, nterm = len(prog.pairs) // 4, len(prog.term_pair)
    njpair, njterm = len(prog.jpairs) // 2, len(prog.jterms)
    nscale = int(((prog.jterms.view(np.uint32) >> 28) != 1).sum()) if njterm else 0
    ngroup = _groups(prog)
    ndyn = int((np.asarray(prog.hsig) < 0).sum()) if prog.hsig is not None else 0
    nhterm = len(prog.hterms) // 3 if prog.hterms is not None else 0

    def add(*ds):
        out = {}
        for d in ds:
            for k, v in d.items():
                out[k] = out.get(k, 0) + v
        return out

    def residual(k):
        return dict(cmul=npair, cmac=npair * min(prog.wdeg, k) + nterm, cadd=n * (G - 1))

    products0 = dict(cmul=nprod)
    ev = add(products0, residual(0))
    # mixed residual: node moduli, coefficient of every term times its
    # weight, group moduli, segment sums, the mean (FP64)
    radial = dict(fp64=4 * n + nprod + 8 * nterm + 6 * ngroup + n * (G - 1) + 7 * n + 1)
    jac = dict(cmul=njpair, cmac=njterm, scale_d=nscale)
    gj = dict(cdiv=n, cmul=2 * n * (n - 1), cmac=n * (n - 1) ** 2)
    apply = dict(cmac=n * n, cadd=n * (G - 1))
    newton = add(jac, gj, apply, dict(cadd=n), ev, radial)
    orders = {}
    for k in range(1, D):
        orders = add(orders, dict(cmul=nprod, cmac=nprod * k), residual(k), apply,
                     dict(cmul=nprod, cmac=nprod, cadd=nprod))
    # curvature step: sigma_min of the inverse Jacobian, the Hessians of the
    # equations of degree > 2 and their singular values, Fabry's ratio and
    # the Pade error estimates
    step = dict(fp64=_sigma_flops(n) + ndyn * _sigma_flops(n) + 14 * nhterm + 12 * n + 60)
    pred = add(ev, jac, gj, orders, step)
    feval = add(ev, dict(scale_d=n * (D - 1), cadd=n * (D - 1)), radial)
    return dict(predictor=pred, newton=newton, feval=feval, eval=ev)


def tracker_flops(prog, K, work, table=BUILD):
    """FP64 flops of one tracking launch given the work counters [8][npaths]
    (predictor series, Newton iterations, predicted-residual evaluations,
    corrector calls, the three step counts, Newton runs whose first
    iteration evaluates the point: every other Newton iteration forms the
    Jacobian at the point evaluated at the end of the previous iteration or
    by the feedback)."""
    ph = _phase_ops(prog)
    c = complex_costs(K, table)
    c = dict(c, fp64=1)

    def cost(d):
        return sum(c[k] * v for k, v in d.items())

    tot = [int(x.sum()) for x in work]
    npred, nnewt, nfeval, nfirst = tot[0], tot[1], tot[2], tot[7]
    return (npred * cost(ph["predictor"]) + nnewt * cost(ph["newton"])
            + nfeval * cost(ph["feval"]) + nfirst * cost(ph["eval"]))


def refine_flops(prog0, K, npaths, iters, table=BUILD):
    """FP64 flops of a refinement launch of `iters` Newton iterations on
    every path (evaluation, Jacobian, inverse, update) plus the final
    residual evaluation (upper bound: converged paths stop early)."""
    ph = _phase_ops(prog0)
    c = dict(complex_costs(K, table), fp64=1)
    cost = lambda d: sum(c[k] * v for k, v in d.items())  # noqa: E731
    it = cost(ph["newton"]) - ph["newton"].get("fp64", 0)   # no mixed residual here
    return npaths * (iters * it + cost(ph["eval"]))


def complex_mac(K, table=BUILD):
    """FP64 flops of one complex multiply-add (4 real products, 4 real sums)"""
    t = table[K]
    return 4 * t["mul"] + 4 * t["add"]


def baqr_work(m, n, szt):
    """complex multiply-adds executed by the GPU blocked QR of an m x n matrix
    in tiles of szt columns (panel dots and updates, Gram matrix, W = Y S,
    the two thin products of each trailing R update and of each Q update)"""
    macs = 0
    nbt = n // szt
    for k in range(nbt):
        rowoff = k * szt
        rowdim = m - rowoff
        for L in range(szt):
            colidx = rowoff + L
            nr, nc = m - colidx, szt - L
            macs += 2 * nr * nc + nr          # dots, update, |x|^2
        macs += szt * szt * rowdim            # G = Y^H Y
        macs += szt * szt * szt // 6          # S recurrence
        macs += rowdim * szt * szt            # W = Y S
        if k < nbt - 1:
            coldim = n - (k + 1) * szt
            macs += 2 * szt * coldim * rowdim  # W^H C and Y T1
        macs += 2 * m * szt * rowdim          # Q W and T2 Y^H
    return macs


def baqr_canonical(n):
    """(real multiplications, real additions) of Householder QR of a complex
    n x n matrix with Q formed explicitly, LAPACK's count: (16/3) n^3 each"""
    c = 16.0 / 3.0 * n ** 3
    return c, c
