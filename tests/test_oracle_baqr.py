"""The CPU oracle of the complex blocked Householder QR and the tiled back
substitution (oracle/baqr.c) against the reference's host code compiled in
oracle/_ref (bit-exact), and against what the mathematics fixes."""
import numpy as np
import pytest

from oraclelib import (oracle_blocked_houseqr, oracle_matvec, oracle_qhb,
                       oracle_upper_tiled_solver, rand_cmd, ref, ref_blocked_houseqr,
                       ref_upper_tiled_solver)

needs_ref = pytest.mark.skipif(ref() is None, reason="oracle/_ref not built (no reference checkout)")


def cplx(x, K):
    """complex128 value (limb sums) of [2K][...] planes"""
    return x[:K].sum(axis=0) + 1j * x[K:].sum(axis=0)


@needs_ref
@pytest.mark.parametrize("K", (2, 4, 8))
@pytest.mark.parametrize("shape", [(12, 12, 4, 3), (16, 8, 4, 2), (9, 9, 3, 3)])
def test_blocked_houseqr_matches_reference(K, shape):
    m, n, szt, nbt = shape
    rng = np.random.default_rng(m * 100 + n + K)
    A = rand_cmd(rng, K, (m, n))
    q1, r1 = ref_blocked_houseqr(K, A, szt, nbt)
    q2, r2 = oracle_blocked_houseqr(K, A, szt, nbt)
    assert np.array_equal(q1, q2) and np.array_equal(r1, r2)


@needs_ref
@pytest.mark.parametrize("K", (2, 4, 8))
def test_upper_tiled_solver_matches_reference(K):
    rng = np.random.default_rng(77 + K)
    szt, nbt = 4, 3
    dim = szt * nbt
    U = rand_cmd(rng, K, (dim, dim))
    for p in range(2 * K):
        U[p] = np.triu(U[p])
    for i in range(dim):
        U[0, i, i] += 3.0
    b = rand_cmd(rng, K, (dim,))
    x1, U1, b1 = ref_upper_tiled_solver(K, U, b, szt, nbt)
    x2, U2, b2 = oracle_upper_tiled_solver(K, U, b, szt, nbt)
    assert np.array_equal(x1, x2) and np.array_equal(U1, U2) and np.array_equal(b1, b2)


@pytest.mark.parametrize("K", (2, 4))
def test_blocked_qr_least_squares_recovers_planted_solution(K):
    """A x = b with b = A x0: Q R = A (Q unitary), x = R^{-1} Q^H b = x0"""
    rng = np.random.default_rng(5 + K)
    szt, nbt = 4, 4
    n = szt * nbt
    A = rand_cmd(rng, K, (n, n))
    x0 = rand_cmd(rng, K, (n,))
    b = oracle_matvec(K, A, x0)
    Q, R = oracle_blocked_houseqr(K, A, szt, nbt)
    Qc, Rc, Ac = cplx(Q, K), cplx(R, K), cplx(A, K)
    assert np.abs(Qc.conj().T @ Qc - np.eye(n)).max() < 1e-14
    assert np.abs(Qc @ Rc - Ac).max() < 1e-14
    y = oracle_qhb(K, Q, b)
    x, _, _ = oracle_upper_tiled_solver(K, R, y, szt, nbt)
    err = np.abs(cplx(x, K) - cplx(x0, K)).max()
    # in K limbs: far below double precision
    d = x - x0
    lead = np.abs(d[0]).max()
    assert err < 1e-14 and lead < (2.0 ** -52) ** (K - 1) * 1e-8


@needs_ref
@pytest.mark.parametrize("K", (2, 4, 8))
@pytest.mark.parametrize("shape", [(12, 12, 4, 3), (16, 8, 4, 2), (9, 9, 3, 3)])
def test_real_blocked_houseqr_matches_reference(K, shape):
    """CPU_dbl{2,4,8}_blocked_houseqr restated: bit-exact"""
    from oraclelib import oracle_real_blocked_houseqr, rand_md, ref_real_blocked_houseqr
    m, n, szt, nbt = shape
    rng = np.random.default_rng(m * 100 + n + K + 5)
    A = rand_md(rng, K, (m, n))
    # a column whose leading entry is positive and one with a zero entry
    # below the diagonal of the first tile: both Householder branches
    A[:, 0, 1] = np.abs(A[:, 0, 1]) * np.sign(A[0, 0, 1])
    q1, r1 = ref_real_blocked_houseqr(K, A, szt, nbt)
    q2, r2 = oracle_real_blocked_houseqr(K, A, szt, nbt)
    assert np.array_equal(q1, q2) and np.array_equal(r1, r2)


@pytest.mark.parametrize("K", (2, 4, 8))
def test_real_blocked_houseqr_properties(K):
    """Q^T Q = I, Q R = A, R upper triangular with a positive diagonal (the
    real Householder vector v0 = x0 - |x| or -sigma/(x0 + |x|) maps x to
    |x| e1), in double from the limb sums; both signs of x0 occur"""
    from oraclelib import oracle_real_blocked_houseqr, rand_md
    m, n, szt, nbt = 24, 16, 4, 4
    rng = np.random.default_rng(3 + K)
    A = rand_md(rng, K, (m, n))
    Q, R = oracle_real_blocked_houseqr(K, A, szt, nbt)
    q, r, a = Q.sum(axis=0), R.sum(axis=0), A.sum(axis=0)
    assert np.abs(q.T @ q - np.eye(m)).max() < 1e-13
    assert np.abs(q @ r - a).max() < 1e-13
    assert np.abs(np.tril(r, -1)).max() < 1e-13
    assert np.all(np.diag(r)[:n] > 0)
    assert (a[:, 0] > 0).any() and (a[0, :] < 0).any()
    # the same factors as numpy's QR up to the signs of the columns of Q
    qn, rn = np.linalg.qr(a, mode="complete")
    s = np.sign(np.diag(rn)[:n])
    assert np.abs(np.abs(rn[:n]) - np.abs(r[:n])).max() < 1e-12
    assert np.abs(qn[:, :n] * s - q[:, :n]).max() < 1e-12
