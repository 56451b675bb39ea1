"""GPU complex blocked Householder QR and tiled back substitution
(phcpack_amd/csrc/baqr.hip) against the CPU oracle (oracle/baqr.c, itself
bit-exact with the reference's host code).  The GPU splits the long sums
across lanes, so the results agree to a few units in the last limb, not
bit for bit: the tolerances below are in units of eps^K (eps = 2^-52)
times the dimension."""
import numpy as np
import pytest

from oraclelib import (md_diff, oracle_blocked_houseqr, oracle_matvec, oracle_upper_tiled_solver,
                       rand_cmd)

pytestmark = pytest.mark.gpu


def ulps(K):
    return (2.0 ** -52) ** K


# Bound on the difference from the oracle: the GPU adds the long sums of a
# column (length <= m) in tree order, the oracle sequentially, so an entry
# of Q differs by at most a few units of eps^K times the sum's length; 8 m
# units for Q and 8 m^1.5 for R (whose entries gather the column sums of
# m rows).  Measured on the MI355X (max over the entries, in units of
# eps^K): Q 114 / R 418 at K = 2, 40 x 40 (bounds 320 / 2024); Q 33 / R 108
# at K = 2, 64 x 64; Q 2.7 / R 188 at K = 2, 256 x 128; Q 1.1 / R 16 at
# K = 4, 128 x 128; below 0.01 at K = 8.
QR_UNITS = 8


@pytest.mark.parametrize("K,shape", [(K, s) for K in (2, 4)
                                     for s in [(32, 32, 8), (48, 32, 8), (64, 64, 16), (40, 40, 40)]]
                         + [(8, (32, 32, 8)), (8, (40, 24, 8))])
def test_blocked_qr_matches_oracle(K, shape):
    from phcpack_amd.baqr import BlockedQR
    m, n, szt = shape
    print(f"blocked qr K={K} {shape}", flush=True)
    rng = np.random.default_rng(m + n + szt + K)
    A = rand_cmd(rng, K, (m, n))
    Q, R = BlockedQR(K, m, n, szt).factor(A)
    Qo, Ro = oracle_blocked_houseqr(K, A, szt, n // szt)
    tol = QR_UNITS * m * ulps(K)
    dq, dr = np.abs(md_diff(K, Q, Qo)).max(), np.abs(md_diff(K, R, Ro)).max()
    print(f"  max |Q - Qo| {dq:.2e} = {dq / ulps(K):.1f} u  |R - Ro| {dr:.2e} = {dr / ulps(K):.1f} u"
          f"  tol {tol:.2e}", flush=True)
    assert dq < tol and dr < tol * np.sqrt(m)


@pytest.mark.parametrize("K,shape", [(2, (256, 128, 128)), (4, (128, 128, 64)),
                                     (4, (128, 128, 128))])
def test_blocked_qr_split_sums_match_oracle(K, shape):
    """sizes where the Gram, W and Q-update products split their k range over
    workgroups (k >= 128) and the panel spans up to 128 columns (the last:
    the quad-double tile of 128 columns of the 1024 x 1024 solve)"""
    test_blocked_qr_matches_oracle(K, shape)


@pytest.mark.parametrize("K", (2, 4, 8))
def test_upper_tiled_solver_matches_oracle(K):
    from phcpack_amd.baqr import upper_tiled_solver
    print(f"tiled solver K={K}", flush=True)
    rng = np.random.default_rng(9 + K)
    szt, nbt = 8, 4
    dim = szt * nbt
    U = rand_cmd(rng, K, (dim, dim))
    for p in range(2 * K):
        U[p] = np.triu(U[p])
    for i in range(dim):
        U[0, i, i] += 4.0
    b = rand_cmd(rng, K, (dim,))
    x, U1, b1 = upper_tiled_solver(K, U, b, szt, nbt)
    xo, Uo, bo = oracle_upper_tiled_solver(K, U, b, szt, nbt)
    tol = 64 * dim * ulps(K)
    dx = np.abs(md_diff(K, x, xo)).max()
    print(f"  max |x - xo| {dx:.2e} = {dx / ulps(K):.1f} u  tol {tol:.2e}", flush=True)
    assert dx < tol
    assert np.abs(md_diff(K, U1, Uo)).max() < tol
    assert np.abs(md_diff(K, b1, bo)).max() < tol


@pytest.mark.parametrize("K,n,szt", [(2, 96, 32), (4, 96, 32), (2, 256, 64)])
def test_least_squares_recovers_planted_solution(K, n, szt):
    """b = A x0 in the oracle; QR, Q^H b (split k range at n = 256) and the
    tiled solve on the GPU give x0; Q^H b agrees with the unsplit product"""
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.baqr import BlockedQR, qh_times, upper_tiled_solver_device
    print(f"least squares K={K} n={n} szt={szt}", flush=True)
    rng = np.random.default_rng(21 + K + n)
    A = rand_cmd(rng, K, (n, n))
    x0 = rand_cmd(rng, K, (n,))
    b = oracle_matvec(K, A, x0)
    qr = BlockedQR(K, n, n, szt)
    dA = DeviceArray.from_host(A)
    dQ, dR = DeviceArray((2 * K, n, n)), DeviceArray((2 * K, n, n))
    qr.factor_device(dA, dQ, dR)
    db, dy, dx = DeviceArray.from_host(b), DeviceArray.zeros((2 * K, n)), DeviceArray.zeros((2 * K, n))
    qr.qhb_device(dQ, db, dy)
    dy1 = DeviceArray.zeros((2 * K, n))
    qh_times(K, dQ, db, n, dy1)
    check(load().phc_synchronize())
    assert np.abs(md_diff(K, dy.to_host(), dy1.to_host())).max() < 64 * n * ulps(K)
    upper_tiled_solver_device(K, n, szt, n // szt, dR, dy, dx)
    check(load().phc_synchronize())
    err = np.abs(md_diff(K, dx.to_host(), x0)).max()
    print(f"  error {err:.2e}", flush=True)
    assert err < 1e6 * ulps(K)


def test_octo_double_gemm_matches_oracle():
    """the complex octo-double GEMM (phc_cmplx_gemm, K = 8: the octo-double
    checks of the quad-double least-squares solve) against the oracle's
    octo-double matrix-vector product: every k term summed in order, so the
    two differ only by the order of the four real products of a complex
    multiply-add: within 4 n eps^8"""
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.baqr import gemm_device
    K, m, n = 8, 40, 72
    rng = np.random.default_rng(88)
    A = rand_cmd(rng, K, (m, n))
    x = rand_cmd(rng, K, (n,))
    dA, dx = DeviceArray.from_host(A), DeviceArray.from_host(x)
    dy = DeviceArray.zeros((2 * K, m))
    print("octo-double gemm", flush=True)
    gemm_device(K, False, False, m, 1, n, (dA, 0, n, 1), (dx, 0, 1, 0), (dy, 0, 1, 0))
    check(load().phc_synchronize())
    y = oracle_matvec(K, A, x)
    err = np.abs(md_diff(K, dy.to_host(), y)).max()
    print(f"  max |y - yo| {err:.2e}", flush=True)
    assert err < 4 * n * ulps(K)


@pytest.mark.parametrize("K,shape", [(2, (96, 64, 32)), (4, (128, 128, 32)), (8, (64, 32, 16))])
def test_tuned_configurations_match_oracle(K, shape):
    """every register block and k split the autotuner may choose: the same
    tolerance against the oracle; the register block alone does not change
    a bit (each output's k sum keeps its order), only the split does"""
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.autotune import QR_BLOCKS, tune_qr
    from phcpack_amd.baqr import BlockedQR
    m, n, szt = shape
    rng = np.random.default_rng(77 + K)
    A = rand_cmd(rng, K, (m, n))
    Qo, Ro = oracle_blocked_houseqr(K, A, szt, n // szt)
    qr = BlockedQR(K, m, n, szt)
    dA = DeviceArray.from_host(np.ascontiguousarray(A))
    dQ, dR = DeviceArray((2 * K, m, m)), DeviceArray((2 * K, m, n))
    tol = QR_UNITS * m * ulps(K)
    first = {}
    for split in (64, 1024, 8192):
        for blk in (1, 2, 3, 4):
            qr.set_tuning(blk, split)
            assert qr.tuning() == (blk, split)
            qr.factor_device(dA, dQ, dR)
            check(load().phc_synchronize())
            Q, R = dQ.to_host(), dR.to_host()
            dq, dr = np.abs(md_diff(K, Q, Qo)).max(), np.abs(md_diff(K, R, Ro)).max()
            print(f"  K={K} {QR_BLOCKS[blk]} split {split}: {dq / ulps(K):.1f} u / {dr / ulps(K):.1f} u",
                  flush=True)
            assert dq < tol and dr < tol * np.sqrt(m)
            if split in first:
                assert np.array_equal(Q, first[split][0]) and np.array_equal(R, first[split][1])
            else:
                first[split] = (Q, R)
    qr.set_tuning(0, 0)
    assert qr.tuning() == ((4, 1024) if K == 2 else (1, 1024))
    best, table = tune_qr(K, m, n, dA, szts=[szt, n], blocks=(1, 4), splits=(1024,), reps=1)
    assert len(table) == 4 and best in table and table[best] == min(table.values())


@pytest.mark.parametrize("K,shape", [(K, s) for K in (2, 4) for s in [(32, 32, 8), (48, 32, 8),
                                                                      (64, 64, 16), (40, 40, 40)]]
                         + [(8, (32, 32, 8)), (2, (256, 128, 128))])
def test_real_blocked_qr_matches_oracle(K, shape):
    """the real QR (phc_real_blocked_houseqr) against the oracle's
    restatement of CPU_dbl{2,4,8}_blocked_houseqr (bit-exact with the
    reference): the tolerance of the complex QR (measured: at most 3.3 units
    of eps^K for Q and 20 for R in dd); R's diagonal positive"""
    from oraclelib import oracle_real_blocked_houseqr, rand_md
    from phcpack_amd.baqr import BlockedQR
    m, n, szt = shape
    print(f"  real blocked qr K={K} {shape}", flush=True)
    rng = np.random.default_rng(31 * m + n + K)
    A = rand_md(rng, K, (m, n))
    Q, R = BlockedQR(K, m, n, szt).factor_real(A)
    Qo, Ro = oracle_real_blocked_houseqr(K, A, szt, n // szt)
    dq, dr = np.abs(md_diff(K, Q, Qo)).max(), np.abs(md_diff(K, R, Ro)).max()
    tol = QR_UNITS * m * ulps(K)
    print(f"  max |Q - Qo| {dq / ulps(K):.1f} u  |R - Ro| {dr / ulps(K):.1f} u", flush=True)
    assert dq < tol and dr < tol * np.sqrt(m)
    # positive diagonal, except a last column with no row below its
    # diagonal (sigma = 0: no reflection, as in the reference)
    assert np.all(np.diag(R[0])[:min(n, m - 1)] > 0)


def test_real_least_squares_recovers_planted_solution():
    """A x0 = b with x0 planted: y = Q^T b (phc_real_qtb), then R x = y on
    the host in double from the limb sums, close to x0"""
    from oraclelib import rand_md
    from phcpack_amd._lib import DeviceArray, check, load
    from phcpack_amd.baqr import BlockedQR
    K, n, szt = 4, 96, 32
    rng = np.random.default_rng(5)
    A = rand_md(rng, K, (n, n))
    x0 = rng.uniform(-1, 1, n)
    a = A.sum(axis=0)
    b = np.zeros((K, n))
    b[0] = a @ x0
    qr = BlockedQR(K, n, n, szt)
    dA = DeviceArray.from_host(np.ascontiguousarray(A))
    dQ, dR = DeviceArray((K, n, n)), DeviceArray((K, n, n))
    db, dy = DeviceArray.from_host(b), DeviceArray((K, n))
    qr.factor_real_device(dA, dQ, dR)
    qr.qtb_device(dQ, db, dy)
    check(load().phc_synchronize())
    q, r, y = dQ.to_host().sum(axis=0), dR.to_host().sum(axis=0), dy.to_host().sum(axis=0)
    assert np.abs(q.T @ b[0] - y).max() < 1e-12
    x = np.linalg.solve(np.triu(r), y)
    assert np.abs(x - x0).max() < 1e-10
