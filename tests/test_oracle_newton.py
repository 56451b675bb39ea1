"""The CPU oracle of the linearized power-series Newton method
(oracle/newton.c) against the reference's own host code compiled in
oracle/_ref (bit-exact), against committed fixtures of the reference's dd
row-Newton test system, and against what the mathematics fixes (QR
identities, planted Toeplitz and Newton solutions)."""
import ctypes
import os

import numpy as np
import pytest

from oraclelib import (GOLDEN, RowSystem, oracle, oracle_houseqr, oracle_linear_residue,
                       oracle_qrbs_solve, oracle_row_newton, pd, planted_row_system, rand_md,
                       ref, ref_cyclic_row_system, ref_houseqr, ref_linear_residue,
                       ref_qrbs_solve, ref_row_newton)

needs_ref = pytest.mark.skipif(ref() is None, reason="oracle/_ref not built (no reference checkout)")
KS = (1, 2, 4, 8)


def md_matmul(K, A, B):
    """C = A*B in K-limb arithmetic (oracle omd ops), planar [K][m][n]"""
    o = oracle()
    m, k = A.shape[1], A.shape[2]
    n = B.shape[2]
    C = np.zeros((K, m, n))
    acc, t = np.zeros(K), np.zeros(K)
    a, b = np.zeros(K), np.zeros(K)
    for i in range(m):
        for j in range(n):
            acc[:] = 0.0
            for L in range(k):
                a[:], b[:] = A[:, i, L], B[:, L, j]
                o.omd_mul(K, pd(a), pd(b), pd(t))
                o.omd_add(K, pd(acc), pd(t), pd(acc))
            C[:, i, j] = acc
    return C


def lead(x):
    """sum of the limbs (double rounding of a K-limb number)"""
    return x.sum(axis=0)


# ------------------------------------------------------ golden fixtures --

def test_cyclic_row_newton_matches_fixture():
    g = np.load(os.path.join(GOLDEN, "newton.npz"))
    sysm = RowSystem(8, 15, g["monoptr"], g["nvr"], g["idxptr"], g["idx"], g["cst"], g["cff"])
    out = oracle_row_newton(2, sysm, g["start"], 8)
    assert out["steps"] == len(g["tail"])
    for k in ("wdeg", "tail", "up", "bs", "history", "Q", "R"):
        assert np.array_equal(out[k], g[k]), k
    # the test solution is planted: the series agrees with it to dd precision
    err = np.abs(lead(out["x"]) - lead(g["testsol"])).max()
    assert err < 1e-20


@needs_ref
def test_cyclic_system_generator_matches_fixture():
    g = np.load(os.path.join(GOLDEN, "newton.npz"))
    sysm, sol, start = ref_cyclic_row_system(7, 8, 15)
    for k, v in (("cst", sysm.cst), ("cff", sysm.cff), ("testsol", sol), ("start", start),
                 ("idx", sysm.idx), ("nvr", sysm.nvr)):
        assert np.array_equal(v, g[k]), k


# ------------------------------------------------- reference, bit-exact --

@needs_ref
@pytest.mark.parametrize("K", KS)
@pytest.mark.parametrize("shape", [(8, 8), (9, 7), (1, 1), (5, 1)])
def test_houseqr_matches_reference(K, shape):
    rng = np.random.default_rng(100 + K)
    A = rand_md(rng, K, shape)
    q1, r1 = ref_houseqr(K, A)
    q2, r2 = oracle_houseqr(K, A)
    assert np.array_equal(q1, q2) and np.array_equal(r1, r2)


@needs_ref
@pytest.mark.parametrize("K", KS)
@pytest.mark.parametrize("tailidx", [1, 3])
def test_qrbs_solve_matches_reference(K, tailidx):
    rng = np.random.default_rng(200 + K + tailidx)
    dim, degp1 = 6, 7
    mat = rand_md(rng, K, (degp1, dim, dim))
    rhs = rand_md(rng, K, (degp1, dim))
    a = ref_qrbs_solve(K, mat, rhs, tailidx)
    b = oracle_qrbs_solve(K, mat, rhs, tailidx)
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    # second call with the factorization kept (zeroQ off) and a zero leading
    # right-hand side: the head is skipped, the tail still runs
    rhs2 = rhs.copy()
    rhs2[:, 0] = 0.0
    rhs2[:, 2] = 0.0
    kw = dict(Q=a["Q"], R=a["R"], zeroQ=0, noqr=0)
    a2 = ref_qrbs_solve(K, mat, rhs2, tailidx, **kw)
    b2 = oracle_qrbs_solve(K, mat, rhs2, tailidx, **kw)
    for k in a2:
        assert np.array_equal(np.asarray(a2[k]), np.asarray(b2[k])), k
    assert a2["noqr"] == 1


@needs_ref
@pytest.mark.parametrize("K", KS)
def test_linear_residue_matches_reference(K):
    rng = np.random.default_rng(300 + K)
    dim, degp1 = 5, 6
    mat = rand_md(rng, K, (degp1, dim, dim))
    rhs = rand_md(rng, K, (degp1, dim))
    sol = rand_md(rng, K, (degp1, dim))
    for tail in (0, 2):
        v1, m1 = ref_linear_residue(K, mat, rhs, sol, tail)
        v2, m2 = oracle_linear_residue(K, mat, rhs, sol, tail)
        assert np.array_equal(v1, v2) and np.array_equal(m1, m2)


@needs_ref
@pytest.mark.parametrize("K", KS)
def test_row_newton_matches_reference(K):
    rng = np.random.default_rng(400 + K)
    sysm, root, start = planted_row_system(rng, K, 8, 15)
    a = ref_row_newton(K, sysm, start, 12)
    b = oracle_row_newton(K, sysm, start, 12)
    assert a["steps"] == b["steps"]
    for k in ("wdeg", "tail", "up", "bs", "history", "x", "Q", "R"):
        assert np.array_equal(a[k], b[k]), k


@needs_ref
def test_row_newton_cyclic_seeds_match_reference():
    for seed in (1, 2, 3):
        sysm, sol, start = ref_cyclic_row_system(seed, 6, 10)
        a = ref_row_newton(2, sysm, start, 8)
        b = oracle_row_newton(2, sysm, start, 8)
        for k in ("tail", "up", "bs", "history"):
            assert np.array_equal(a[k], b[k]), (seed, k)


# ------------------------------------------------------ mathematics ------

@pytest.mark.parametrize("K", KS)
def test_houseqr_identities(K):
    """Q^T Q = I and Q R = A to the working precision; R upper triangular"""
    rng = np.random.default_rng(500 + K)
    A = rand_md(rng, K, (7, 5))
    Q, R = oracle_houseqr(K, A)
    eps = (2.0 ** -52) ** K * 64
    QR = md_matmul(K, Q, R)
    # differences computed in K limbs
    QT = np.ascontiguousarray(np.transpose(Q, (0, 2, 1)))
    QTQ = md_matmul(K, QT, Q)
    I = np.zeros_like(QTQ)
    I[0] = np.eye(7)
    o = oracle()
    d, t = np.zeros(K), np.zeros(K)
    worst_qr, worst_qq = 0.0, 0.0
    for i in range(7):
        for j in range(5):
            a, b = QR[:, i, j].copy(), A[:, i, j].copy()
            o.omd_sub(K, pd(a), pd(b), pd(d))
            worst_qr = max(worst_qr, abs(d[0]))
        for j in range(7):
            a, b = QTQ[:, i, j].copy(), I[:, i, j].copy()
            o.omd_sub(K, pd(a), pd(b), pd(t))
            worst_qq = max(worst_qq, abs(t[0]))
    assert worst_qr < eps and worst_qq < eps
    # the updates leave rounding residues, not zeros, below the diagonal
    assert np.abs(np.tril(R[0], -1)).max() < eps


@pytest.mark.parametrize("K", (2, 4, 8))
def test_toeplitz_solve_recovers_planted_series(K):
    """b_k = sum_{j<=k} A_j x_{k-j} for a planted x: the solve returns x"""
    rng = np.random.default_rng(600 + K)
    dim, degp1 = 5, 6
    mat = rand_md(rng, K, (degp1, dim, dim))
    for k in range(dim):
        mat[0, 0, k, k] += 4.0  # well conditioned leading matrix
    x = rand_md(rng, K, (degp1, dim))
    o = oracle()
    rhs = np.zeros((K, degp1, dim))
    for k in range(degp1):
        acc = np.zeros((K, dim, 1))
        for j in range(k + 1):
            prod = md_matmul(K, np.ascontiguousarray(mat[:, j]),
                             np.ascontiguousarray(x[:, k - j, :, None]))
            for i in range(dim):
                a, b = acc[:, i, 0].copy(), prod[:, i, 0].copy()
                o.omd_add(K, pd(a), pd(b), pd(a))
                acc[:, i, 0] = a
        rhs[:, k] = acc[:, :, 0]
    out = oracle_qrbs_solve(K, mat, rhs, 1)
    err = np.abs(lead(out["sol"]) - lead(x)).max()
    assert err < (2.0 ** -52) ** K * 1e6
    assert out["up"] == 0 and out["bs"] == 0 and out["newtail"] == 1


@pytest.mark.parametrize("K", (4, 8))
def test_row_newton_converges_to_planted_root(K):
    """A higher precision reaches the planted exponential-series root more
    closely: the error of the K-limb run is far below the dd run's."""
    errs = {}
    for k in (2, K):
        rng = np.random.default_rng(700)
        sysm, root, start = planted_row_system(rng, k, 6, 8)
        out = oracle_row_newton(k, sysm, start, 12)
        errs[k] = np.abs(lead(out["x"]) - lead(root)).max()
    assert errs[2] < 1e-16
    assert errs[K] < errs[2] * 1e-10 or errs[K] == 0.0


# ------------------------------------------ exponents > 1 (common factors) --

def c1_system(rng, K, dim=8, deg=15, nper=8, nvar=3):
    """the row-Newton configuration with exponents: dim polynomials of nper
    monomials of nvar distinct variables, exponents uniform in {1, 2},
    coefficient series of degree deg with coefficients uniform in [-1, 1],
    the constant series random too; start: a random constant series"""
    from oraclelib import random_exponents, random_supports
    D = deg + 1
    monoptr, nvr, idxptr, idx = random_supports(rng, dim, nper, nvar)
    cff = rand_md(rng, K, (int(monoptr[-1]), D))
    cst = rand_md(rng, K, (dim, D))
    sysm = RowSystem(dim, deg, monoptr, nvr, idxptr, idx, cst, cff)
    sysm.exp = random_exponents(rng, sysm)
    start = np.zeros((K, dim, D))
    start[0, :, 0] = rng.uniform(-1.0, 1.0, dim)
    return sysm, start


@needs_ref
@pytest.mark.parametrize("K", KS)
def test_row_newton_with_exponents_matches_reference(K):
    """8 staggered steps at the configuration's shape (8 x 8, three
    variables per monomial, exponents in {1, 2}, degree 15): the oracle's
    common-factor evaluation and Newton steps against the reference's
    CPU_*_product, CPU_*_speel and *_mul_*_d composed in the driver"""
    rng = np.random.default_rng(800 + K)
    sysm, start = c1_system(rng, K)
    assert set(np.unique(sysm.exp)) == {1, 2}
    a = ref_row_newton(K, sysm, start, 8)
    b = oracle_row_newton(K, sysm, start, 8)
    assert a["steps"] == b["steps"] == 8
    for k in ("wdeg", "tail", "up", "bs", "history", "x", "Q", "R"):
        assert np.array_equal(a[k], b[k]), k


@needs_ref
@pytest.mark.parametrize("nvar", [1, 2, 4])
def test_row_newton_with_exponents_other_sizes_match_reference(nvar):
    """monomials of one, two and four variables (the special cases of the
    speel products) with exponents up to 3"""
    from oraclelib import random_exponents
    rng = np.random.default_rng(810 + nvar)
    sysm, start = c1_system(rng, 2, dim=5, deg=6, nper=4, nvar=nvar)
    sysm.exp = random_exponents(rng, sysm, (1, 2, 3))
    a = ref_row_newton(2, sysm, start, 6)
    b = oracle_row_newton(2, sysm, start, 6)
    for k in ("wdeg", "tail", "up", "bs", "history", "x"):
        assert np.array_equal(a[k], b[k]), k


def test_exponents_equal_repeated_variables():
    """x_j^2 written with the exponent and written as the variable repeated
    give the same value and gradient (the repeated variable's derivative
    terms add up to 2 x_j ...), to double-double rounding"""
    from oraclelib import pi, random_exponents, random_supports
    rng = np.random.default_rng(820)
    K, dim, deg = 2, 10, 7   # dim >= the largest total degree (work rows)
    D = deg + 1
    monoptr, nvr, idxptr, idx = random_supports(rng, dim, 5, 3)
    nmono = int(monoptr[-1])
    exp = rng.choice(np.array([1, 2, 3], np.int32), size=len(idx)).astype(np.int32)
    cff = rand_md(rng, K, (nmono, D))
    cst = rand_md(rng, K, (1, D))[:, 0]
    x = rand_md(rng, K, (dim, D))
    o = oracle()
    # polynomial made of the first 5 monomials
    nb = 5
    nv, ip = np.ascontiguousarray(nvr[:nb]), np.ascontiguousarray(idxptr[:nb])
    out_e = np.zeros((K, dim + 1, D))
    o.osr_poly_evaldiff_exp(K, 0, dim, nb, deg, pi(nv), pi(ip), pi(idx), pi(exp),
                            pd(np.ascontiguousarray(cst)), pd(np.ascontiguousarray(cff[:, :nb])),
                            pd(x), pd(out_e))
    # the same monomials with every variable repeated e times
    idx_r, nvr_r = [], []
    for m in range(nb):
        vs = [int(idx[j]) for j in range(idxptr[m], idxptr[m + 1]) for _ in range(exp[j])]
        idx_r += vs
        nvr_r.append(len(vs))
    idx_r = np.array(idx_r, np.int32)
    nvr_r = np.array(nvr_r, np.int32)
    ip_r = np.concatenate([[0], np.cumsum(nvr_r)[:-1]]).astype(np.int32)
    out_r = np.zeros((K, dim + 1, D))
    o.osr_poly_evaldiff(K, 0, dim, nb, deg, pi(nvr_r), pi(ip_r), pi(idx_r),
                        pd(np.ascontiguousarray(cst)), pd(np.ascontiguousarray(cff[:, :nb])),
                        pd(x), pd(out_r))
    a, b = out_e[0] + out_e[1], out_r[0] + out_r[1]
    assert np.abs(a - b).max() <= 1e-28 * max(1.0, np.abs(b).max())
    assert np.abs(a).max() > 1e-3


@pytest.mark.parametrize("K", (2, 4))
def test_row_newton_with_exponents_converges_to_planted_root(K):
    """with exponents {1, 2} the staggered Newton method still reaches a
    planted exponential-series root"""
    from oraclelib import random_exponents, random_supports
    rng = np.random.default_rng(830)
    sup = random_supports(rng, 6, 6, 3)
    exp = rng.choice(np.array([1, 2], np.int32), size=len(sup[3])).astype(np.int32)
    sysm, root, start = planted_row_system(np.random.default_rng(831), K, 6, 8, supports=sup,
                                           exp=exp)
    out = oracle_row_newton(K, sysm, start, 12)
    assert np.abs(lead(out["x"]) - lead(root)).max() < {2: 1e-20, 4: 1e-40}[K]


# ------------------------------------------------------------ complex

CKS = (2, 4, 8)


@needs_ref
@pytest.mark.parametrize("K", CKS)
@pytest.mark.parametrize("shape", [(8, 8), (9, 7), (1, 1), (5, 1)])
def test_cx_houseqr_matches_reference(K, shape):
    """CPU_cmplx{2,4,8}_factors_houseqr restated: bit-exact"""
    from oraclelib import oracle_cx_houseqr, rand_cmd, ref_cx_houseqr
    rng = np.random.default_rng(10 * K + shape[0])
    A = rand_cmd(rng, K, shape)
    q1, r1 = ref_cx_houseqr(K, A)
    q2, r2 = oracle_cx_houseqr(K, A)
    assert np.array_equal(q1, q2) and np.array_equal(r1, r2)


@needs_ref
@pytest.mark.parametrize("K", CKS)
@pytest.mark.parametrize("tailidx", [1, 3])
def test_cx_qrbs_solve_matches_reference(K, tailidx):
    """CPU_cmplx{2,4,8}_qrbs_solve restated (head, tail, skip rules): a
    zero block right-hand side and a zero solution block exercise the skips"""
    from oraclelib import oracle_cx_qrbs_solve, rand_cmd, ref_cx_qrbs_solve
    rng = np.random.default_rng(40 + K + tailidx)
    dim, degp1 = 6, 7
    mat = rand_cmd(rng, K, (degp1, dim, dim))
    rhs = rand_cmd(rng, K, (degp1, dim))
    rhs[:, 4] = 0.0
    a = ref_cx_qrbs_solve(K, mat, rhs, tailidx)
    b = oracle_cx_qrbs_solve(K, mat, rhs, tailidx)
    for k in ("rhs", "sol", "Q", "R", "zeroQ", "noqr", "up", "bs", "newtail"):
        assert np.array_equal(a[k], b[k]), k
    # higher blocks zero and a zero block of the right-hand side: its
    # solution block is skipped (bs), and then the update with it (up)
    mz = mat.copy()
    mz[:, 1:] = 0.0
    rz = rhs.copy()
    rz[:, 3] = 0.0
    az = ref_cx_qrbs_solve(K, mz, rz, tailidx)
    bz = oracle_cx_qrbs_solve(K, mz, rz, tailidx)
    for k in ("rhs", "sol", "Q", "R", "zeroQ", "noqr", "up", "bs", "newtail"):
        assert np.array_equal(az[k], bz[k]), k
    assert az["bs"] >= 1 and az["up"] >= 1
    # a second call reusing Q, R with a zero leading right-hand side (noqr)
    rhs2 = rand_cmd(rng, K, (degp1, dim))
    rhs2[:, 0] = 0.0
    kw = dict(Q=a["Q"], R=a["R"], zeroQ=a["zeroQ"], noqr=a["noqr"])
    a2 = ref_cx_qrbs_solve(K, mat, rhs2, tailidx, **kw)
    b2 = oracle_cx_qrbs_solve(K, mat, rhs2, tailidx, **kw)
    for k in ("rhs", "sol", "zeroQ", "noqr", "up", "bs", "newtail"):
        assert np.array_equal(a2[k], b2[k]), k
    assert a2["noqr"] == 1


@needs_ref
@pytest.mark.parametrize("K", CKS)
def test_cx_linear_residue_matches_reference(K):
    from oraclelib import oracle_cx_linear_residue, rand_cmd, ref_cx_linear_residue
    rng = np.random.default_rng(70 + K)
    dim, degp1 = 5, 6
    mat = rand_cmd(rng, K, (degp1, dim, dim))
    rhs = rand_cmd(rng, K, (degp1, dim))
    sol = rand_cmd(rng, K, (degp1, dim))
    for tail in (0, 2):
        r1, m1 = ref_cx_linear_residue(K, mat, rhs, sol, tail)
        r2, m2 = oracle_cx_linear_residue(K, mat, rhs, sol, tail)
        assert np.array_equal(r1, r2) and np.array_equal(m1, m2)


@needs_ref
@pytest.mark.parametrize("K", (2, 4))
def test_cx_row_newton_matches_reference(K):
    """the staggered complex row Newton (test_cmplx{2,4}_row_newton's host
    branch on the reference's functions) restated: every step bit-exact"""
    from oraclelib import oracle_cx_row_newton, planted_cx_row_system, ref_cx_row_newton
    rng = np.random.default_rng(90 + K)
    sysm, root, start = planted_cx_row_system(rng, K, 6, 12, nper=4, nvar=3)
    a = ref_cx_row_newton(K, sysm, start, 12)
    b = oracle_cx_row_newton(K, sysm, start, 12)
    assert a["steps"] == b["steps"]
    for k in ("x", "wdeg", "tail", "up", "bs", "history", "Q", "R"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("K", (2, 4))
def test_cx_row_newton_converges_to_planted_root(K):
    """the complex row Newton from the perturbed leading coefficients
    reaches the planted exponential-series root to working precision"""
    from oraclelib import oracle_cx_row_newton, planted_cx_row_system
    rng = np.random.default_rng(120 + K)
    sysm, root, start = planted_cx_row_system(rng, K, 6, 12, nper=4, nvar=3)
    out = oracle_cx_row_newton(K, sysm, start, 16)
    err = np.abs(out["x"][:K].sum(axis=0) - root[:K].sum(axis=0)).max() + \
        np.abs(out["x"][K:].sum(axis=0) - root[K:].sum(axis=0)).max()
    assert err < 1e-12, err
    assert out["tail"][-1] >= 12


@pytest.mark.parametrize("K", (2, 4))
def test_cx_houseqr_identities(K):
    """Q^H Q = I and Q R = A in double from the limb sums; R upper triangular"""
    from oraclelib import oracle_cx_houseqr, rand_cmd
    rng = np.random.default_rng(5 + K)
    A = rand_cmd(rng, K, (9, 9))
    Q, R = oracle_cx_houseqr(K, A)
    q = Q[:K].sum(axis=0) + 1j * Q[K:].sum(axis=0)
    r = R[:K].sum(axis=0) + 1j * R[K:].sum(axis=0)
    a = A[:K].sum(axis=0) + 1j * A[K:].sum(axis=0)
    assert np.abs(q.conj().T @ q - np.eye(9)).max() < 1e-13
    assert np.abs(q @ r - a).max() < 1e-13
    assert np.abs(np.tril(r, -1)).max() < 1e-13
