"""The column Newton method of the oracle (onw_column_newton): monomial
systems x^A = mb and column systems (cyclic columns, two triangular columns).

Pinned by
  * bit-exact equality with the reference's own host code compiled in
    oracle/_ref (CPU_dbl*_evaluate_monomials / _columns, the linearization,
    CPU_dbl*_qrbs_solve, driven as test_dbl2_column_newton) at K = 1, 2, 4, 8,
    on the reference's generators and on planted systems;
  * the committed fixture of the reference's dd runs (tests/golden/columns.npz),
    which needs no reference checkout;
  * convergence to a planted root whose right-hand side the oracle's row
    evaluator computed (a wrong product order, sign or exponent scaling in
    the column evaluation would leave a residual and stall the tail index)."""
import os

import numpy as np
import pytest

from oraclelib import (ColumnSystem, oracle_column_newton, planted_column_system, ref,
                       ref_column_newton)

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("x", "wdeg", "tail", "up", "bs", "history", "Q", "R")
needs_ref = pytest.mark.skipif(ref() is None, reason="oracle/_ref not built (no reference checkout)")


def golden_case(name):
    z = np.load(os.path.join(HERE, "golden", "columns.npz"))
    dim, deg, nbrcol = (int(v) for v in z[name + "_meta"])
    g = (lambda k: z[name + "_" + k] if name + "_" + k in z else None)
    csys = ColumnSystem(dim, deg, nbrcol, g("nvr"), g("idx"), g("mb"), cff=g("cff"),
                        exps=g("exps"))
    return csys, g


@pytest.mark.parametrize("name", ["mono", "two", "cyclic"])
def test_column_newton_matches_reference_fixture(name):
    csys, g = golden_case(name)
    got = oracle_column_newton(2, csys, g("start"), 12)
    assert got["steps"] == len(g("tail"))
    for k in ("history", "wdeg", "tail", "up", "bs", "Q", "R"):
        assert np.array_equal(got[k], g(k)), k


def test_fixture_runs_reach_the_test_solution():
    """the fixture runs approach the reference's exponential test solution:
    the leading coefficients to dd precision, the two triangular columns and
    the monomial system over the whole series (the cyclic columns amplify the
    error about 16 times per degree, so only their leading part is held)"""
    for name, deg_ok, tol in (("two", 16, 1e-29), ("mono", 16, 1e-22), ("cyclic", 6, 1e-23)):
        csys, g = golden_case(name)
        x = g("history")[-1]
        err = np.abs((x[0] - g("testsol")[0]) + (x[1] - g("testsol")[1]))[:, :deg_ok].max()
        assert err < tol, (name, err)
        assert np.abs((x[0] - g("testsol")[0]) + (x[1] - g("testsol")[1]))[:, 0].max() < 1e-28


@needs_ref
@pytest.mark.parametrize("K", [1, 2, 4, 8])
@pytest.mark.parametrize("case", ["mono", "two", "cyclic"])
def test_column_newton_matches_compiled_reference_on_its_systems(K, case):
    from golden.make_golden import column_cases
    csys, sol, start = dict(column_cases())[case]
    c = csys.with_limbs(K)
    st = np.zeros((K,) + start.shape[1:])
    st[:min(K, 2)] = start[:K]
    got, want = oracle_column_newton(K, c, st, 12), ref_column_newton(K, c, st, 12)
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k


@needs_ref
@pytest.mark.parametrize("K", [1, 2, 4, 8])
@pytest.mark.parametrize("dim,nbrcol", [(5, 1), (8, 1), (3, 2), (6, 3), (7, 7)])
def test_column_newton_matches_compiled_reference_on_planted_systems(K, dim, nbrcol):
    rng = np.random.default_rng(100 * dim + nbrcol + K)
    csys, root, start = planted_column_system(rng, K, dim, 11, nbrcol)
    got, want = oracle_column_newton(K, csys, start, 10), ref_column_newton(K, csys, start, 10)
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.parametrize("K,tol", [(1, 1e-9), (2, 1e-24), (4, 1e-24)])
@pytest.mark.parametrize("nbrcol", [1, 4])
def test_column_newton_converges_to_planted_root(K, tol, nbrcol):
    """mb is the row evaluator's value at the root (an independent code
    path): the staggered Newton reaches the full degree and the root (the
    root's coefficients are doubles, so K > 2 is held to the dd level of
    the start perturbation's recovery)"""
    rng = np.random.default_rng(7 + nbrcol)
    A = np.tril(np.ones((6, 6), np.int32)) + np.diag(np.arange(6) % 2).astype(np.int32)
    csys, root, start = planted_column_system(rng, K, 6, 15, nbrcol, rowsA=A)
    run = oracle_column_newton(K, csys, start, 12)
    assert run["tail"][-1] >= csys.deg
    err = np.abs(run["x"].sum(axis=0) - root.sum(axis=0)).max()
    assert err < tol, err
