"""The C++ program compiler (phc_program_compile, used by callers without
Python such as the gpumanypaths_dd entry) against compile_homotopy, the
Python host compiler: the same arrays, integer for integer and bit for bit,
on dense, cubic, quartic, rank-deficient and the reference's Demo systems,
with merged and cancelling terms; the Hessian singular values (a Jacobi
SVD in C++, LAPACK's in numpy) to rounding.  Double-double coefficients
are checked by evaluating the compiled program with the oracle against
exact rational arithmetic."""
import os
from fractions import Fraction

import numpy as np
import pytest

from oraclelib import oracle_eval
from phcpack_amd.formats import parse_system
from phcpack_amd.homotopy import (Homotopy, System, compile_homotopy, compile_homotopy_native,
                                  compile_terms, dense_system, total_degree_homotopy)

DEMO = os.path.join(os.path.dirname(__file__), "golden", "demo")
INT_FIELDS = ("nodes", "lev_ptr", "pairs", "seg_ptr", "term_pair", "jpairs", "jent_ptr",
              "jterms", "hent_ptr", "hterms", "hpairs")


def assert_same(a, b):
    for f in ("n", "deg", "ncomp", "wdeg", "G", "nslots", "nfactor"):
        assert getattr(a, f) == getattr(b, f), f
    for f in INT_FIELDS:
        x = np.asarray(getattr(a, f)).reshape(-1).astype(np.int64)
        y = np.asarray(getattr(b, f)).reshape(-1).astype(np.int64)
        assert x.shape == y.shape and np.array_equal(x, y), f
    for f in ("term_coef", "weights"):
        x, y = np.asarray(getattr(a, f)), np.asarray(getattr(b, f))
        assert x.shape == y.shape and np.array_equal(x.view(np.int64), y.view(np.int64)), f
    np.testing.assert_allclose(b.hsig, a.hsig, rtol=1e-13, atol=0)


@pytest.mark.parametrize("n,d,rankdef", [(3, 2, 0), (12, 2, 0), (5, 3, 0), (8, 3, 0), (4, 4, 0),
                                         (3, 6, 0), (14, 2, 2), (6, 3, 1)])
@pytest.mark.parametrize("deg", [0, 3, 8])
def test_dense_systems(n, d, rankdef, deg):
    rng = np.random.default_rng(1000 * n + 10 * d + rankdef)
    hom, _ = total_degree_homotopy(dense_system(n, d, rng, rank_deficient=rankdef), rng)
    assert_same(compile_homotopy(hom, deg), compile_homotopy_native(hom, deg))


@pytest.mark.parametrize("name", ["katsura5", "cyclic5", "redeco5", "mickey", "utbikker"])
def test_demo_systems(name):
    with open(os.path.join(DEMO, name)) as f:
        system, _ = parse_system(f.read())
    hom, _ = total_degree_homotopy(system, np.random.default_rng(7))
    for lanes in (64, 32):
        assert_same(compile_homotopy(hom, 8, lanes), compile_homotopy_native(hom, 8, lanes))


def test_merged_cancelled_and_signed_zero_terms():
    """repeated monomials are added in the order they come, a sum that
    cancels stays a (zero) term, -0.0 becomes +0.0, and the relaxation
    power 2 gives weights of degree 2"""
    eqs = [
        [(1.5 + 0.5j, (2, 0, 0)), (-1.5 - 0.5j, (2, 0, 0)), (0.25j, (0, 1, 1)), (1.0, (0, 0, 0)),
         (0.1 + 0.2j, (0, 1, 1)), (0.3 - 0.1j, (0, 1, 1))],
        [(complex(-0.0, -0.0), (1, 1, 0)), (2.0, (0, 2, 0)), (-1.0, (0, 0, 0)), (0.7, (1, 0, 0))],
        [(1.0, (0, 0, 3)), (0.5, (1, 1, 1)), (-0.5, (1, 1, 1)), (0.25, (1, 1, 1)), (-2.0, (0, 0, 0))],
    ]
    p = System(3, eqs)
    q = System(3, [[(1.0, (2, 0, 0)), (-1.0, (0, 0, 0))], [(1.0, (0, 2, 0)), (-1.0, (0, 0, 0))],
                   [(1.0, (0, 0, 3)), (-1.0, (0, 0, 0))]])
    for k in (1, 2):
        hom = Homotopy(p, q, complex(0.6, 0.8), k)
        for deg in (0, 4, 8):
            assert_same(compile_homotopy(hom, deg), compile_homotopy_native(hom, deg))


def test_hessian_singular_value_closed_form():
    """with a linear target, the Hessians of gamma*(1-t)*q + t*p at t = 0 are
    gamma times those of q; |gamma| = 1, so the singular values are those
    of q's Hessians: diag(2, -8, 0) -> 8, diag(0, 6, 0) -> 6, and the
    Hessian of y z, with ones off the diagonal -> 1"""
    n = 3
    q = System(n, [[(1.0, (2, 0, 0)), (-4.0, (0, 2, 0)), (-1.0, (0, 0, 0))],
                   [(3.0, (0, 2, 0)), (-1.0, (0, 0, 0))],
                   [(1.0, (0, 1, 1)), (-1.0, (0, 0, 0))]])
    p = System(n, [[(1.0, (1, 0, 0))], [(1.0, (0, 1, 0))], [(1.0, (0, 0, 1))]])
    g = complex(0.28, 0.96)
    prog = compile_homotopy_native(Homotopy(p, q, g, 1), 8)
    # Hessians diag(2, -8, 0), diag(0, 6, 0), [[0,0,0],[0,0,1],[0,1,0]]
    np.testing.assert_allclose(prog.hsig, [8.0, 6.0, 1.0], rtol=1e-15)


def _exact_value(comps, weights, x, t, n):
    """H(x, t) in exact complex rational arithmetic; x complex rationals"""
    def mul(a, b):
        return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])

    def add(a, b):
        return (a[0] + b[0], a[1] + b[1])

    out = [(Fraction(0), Fraction(0))] * n
    for (eqs, exps, cf), w in zip(comps, weights):
        wt = (Fraction(0), Fraction(0))
        tp = (Fraction(1), Fraction(0))
        for l in range(3):
            wt = add(wt, mul((Fraction(w[l].real), Fraction(w[l].imag)), tp))
            tp = mul(tp, (Fraction(t), Fraction(0)))
        for k in range(len(eqs)):
            c = (Fraction(cf[0][k]) + Fraction(cf[1][k]), Fraction(cf[2][k]) + Fraction(cf[3][k]))
            v = mul(wt, c)
            for j in range(n):
                for _ in range(exps[k][j]):
                    v = mul(v, x[j])
            out[eqs[k]] = add(out[eqs[k]], v)
    return out


def test_double_double_coefficients():
    """a coefficient c_hi + c_lo becomes two terms of its monomial: the
    oracle's double-double value of the compiled program equals the exact
    value of the double-double system to ~1e-31, where dropping the low
    parts would leave errors of ~1e-17"""
    rng = np.random.default_rng(11)
    n = 3
    exps = [e for e in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (0, 1, 1),
                        (0, 0, 2))]
    comps = []
    for _ in range(2):
        eqs, ex, cf = [], [], []
        for i in range(n):
            for e in exps:
                eqs.append(i)
                ex.append(e)
                hi = rng.uniform(-1, 1, 2)
                lo = hi * rng.uniform(-1, 1, 2) * 2.0 ** -53
                cf.append((hi[0], lo[0], hi[1], lo[1]))
        cf = np.array(cf).T.copy()  # [4][m]
        comps.append((np.array(eqs), np.array(ex), cf))
    weights = np.array([[0, 1, 0], [0.6 + 0.8j, -0.6 - 0.8j, 0]], np.complex128)
    prog = compile_terms(n, 0, weights, comps)
    nlo = sum(int(np.count_nonzero((c[1] != 0) | (c[3] != 0))) for _, _, c in comps)
    prog1 = compile_terms(n, 0, weights, [(e, x, c[0] + 1j * c[2]) for e, x, c in comps])
    assert prog.term_pair.size == prog1.term_pair.size + nlo
    # a double-double point: hi + lo with lo ~ 2^-53 hi
    xh = rng.uniform(-1, 1, (2, n))
    xl = xh * rng.uniform(-1, 1, (2, n)) * 2.0 ** -53
    planes = np.stack([xh[0], xl[0], xh[1], xl[1]])
    t = 0.375
    val, _ = oracle_eval(prog, planes, 2, t)
    val1, _ = oracle_eval(prog1, planes, 2, t)
    xq = [(Fraction(xh[0, j]) + Fraction(xl[0, j]), Fraction(xh[1, j]) + Fraction(xl[1, j]))
          for j in range(n)]
    exact = _exact_value(comps, weights, xq, t, n)
    for i in range(n):
        got = (Fraction(val[0, i]) + Fraction(val[1, i]), Fraction(val[2, i]) + Fraction(val[3, i]))
        got1 = (Fraction(val1[0, i]) + Fraction(val1[1, i]), Fraction(val1[2, i]) + Fraction(val1[3, i]))
        err = max(abs(float(got[q] - exact[i][q])) for q in range(2))
        err1 = max(abs(float(got1[q] - exact[i][q])) for q in range(2))
        assert err < 1e-29, err
        assert err1 > 1e-20, err1  # the low parts matter at this precision


def test_bad_input_is_refused():
    from phcpack_amd._lib import PhcError
    w = np.array([[0, 1, 0]], np.complex128)
    with pytest.raises(PhcError):  # exponent above 15
        compile_terms(2, 8, w, [(np.array([0, 1]), np.array([[16, 0], [0, 1]]), np.ones(2, complex))])
    with pytest.raises(PhcError):  # equation index out of range
        compile_terms(2, 8, w, [(np.array([0, 2]), np.array([[1, 0], [0, 1]]), np.ones(2, complex))])
    with pytest.raises(PhcError):  # every weight zero
        compile_terms(2, 8, np.zeros((1, 3), complex),
                      [(np.array([0, 1]), np.array([[1, 0], [0, 1]]), np.ones(2, complex))])
