"""Pins the tracker oracle.

1. The compiled program's value and Jacobian of H(x,t) agree with the
   reference-order polynomial evaluator (osr_poly_evaldiff, bit-exact with
   CPU_*_poly_evaldiff) applied to every component, monomials with exponents
   written as repeated variables (the product rule then sums the occurrences).
2. The Taylor series of the solution curve makes H(x(s), t0+s) vanish to
   order deg when H is evaluated with coefficient series by the same
   reference-order evaluator.
3. Closed forms: x_i^2 = a_i (all sign patterns of sqrt(a)), and the Bezout
   count of distinct endpoints for dense random quadratics.
"""
import itertools

import numpy as np
import pytest

from oraclelib import oracle, oracle_eval, oracle_series, oracle_track, pd, pi
from phcpack_amd.homotopy import (System, compile_homotopy, dense_system, total_degree_homotopy)
from phcpack_amd.tracker import TrackParams, planes_to_points, points_to_planes


def cplx_to_md(z, K):
    """complex128 array -> [2K][...] planes"""
    z = np.asarray(z, dtype=np.complex128)
    out = np.zeros((2 * K,) + z.shape)
    out[0], out[K] = z.real, z.imag
    return out


def supports_of(sysm, i):
    sup, cf = [], []
    for c, e in sysm.eqs[i]:
        if sum(e) == 0:
            continue
        sup.append([j for j in range(sysm.n) for _ in range(e[j])])
        cf.append(c)
    cst = sum(c for c, e in sysm.eqs[i] if sum(e) == 0)
    return sup, cf, cst


def ref_eval_component(sysm, i, K, x_series, coef_scale_series):
    """value and gradient of equation i at series x (planes [2K][n][D]),
    coefficients multiplied by a scalar series [2K][D]; reference order."""
    sup, cf, cst = supports_of(sysm, i)
    n, D = sysm.n, x_series.shape[2]
    o = oracle()
    P = 2 * K
    cffs = np.zeros((P, max(len(sup), 1), D))
    for m, c in enumerate(cf):
        cm = cplx_to_md(c, K)
        for k in range(D):
            prod = np.zeros(P)
            o.omd_cmul(K, pd(np.ascontiguousarray(cm)), pd(np.ascontiguousarray(coef_scale_series[:, k])), pd(prod))
            cffs[:, m, k] = prod
    cstc = np.zeros((P, D))
    cm = cplx_to_md(cst, K)
    for k in range(D):
        prod = np.zeros(P)
        o.omd_cmul(K, pd(np.ascontiguousarray(cm)), pd(np.ascontiguousarray(coef_scale_series[:, k])), pd(prod))
        cstc[:, k] = prod
    nvr = np.array([len(s) for s in sup], np.int32)
    idx = np.array([j for s in sup for j in s] or [0], np.int32)
    ptr = np.concatenate([[0], np.cumsum(nvr)]).astype(np.int32)
    out = np.zeros((P, n + 1, D))
    o.osr_poly_evaldiff(K, 1, n, len(sup), D - 1, pi(nvr), pi(ptr), pi(idx), pd(cstc),
                        pd(np.ascontiguousarray(cffs)), pd(np.ascontiguousarray(x_series)), pd(out))
    return out


def weight_series(w, t, D, K):
    """Taylor coefficients of w(t0+s) as complex K planes [2K][D]"""
    W = np.zeros(D, np.complex128)
    W[0] = w[0] + w[1] * t + w[2] * t * t
    if D > 1:
        W[1] = w[1] + 2 * w[2] * t
    if D > 2:
        W[2] = w[2]
    return cplx_to_md(W, K)


@pytest.mark.parametrize("degree", [2, 3])
def test_program_eval_matches_reference_evaluator(degree):
    rng = np.random.default_rng(degree)
    n, K = 4, 2
    p = dense_system(n, degree, rng)
    hom, _ = total_degree_homotopy(p, rng)
    prog = compile_homotopy(hom, 0)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    t = 0.37
    val, jac = oracle_eval(prog, cplx_to_md(x, K), K, t)
    v = val[:K].sum(0) + 1j * val[K:].sum(0)
    J = (jac[:K].sum(0) + 1j * jac[K:].sum(0)).reshape(n, n)
    vr = np.zeros(n, np.complex128)
    Jr = np.zeros((n, n), np.complex128)
    xs = cplx_to_md(x, K)[:, :, None]
    for w, sysm in hom.components:
        ws = weight_series(w, t, 1, K)
        for i in range(n):
            out = ref_eval_component(sysm, i, K, xs, ws)
            z = out[:K].sum(0) + 1j * out[K:].sum(0)
            vr[i] += z[n, 0]
            Jr[i] += z[:n, 0]
    assert np.allclose(v, vr, rtol=0, atol=1e-13 * np.abs(vr).max())
    assert np.allclose(J, Jr, rtol=0, atol=1e-13 * np.abs(Jr).max())
    # independent complex128 evaluation of the value
    assert np.allclose(v, hom.eval(x, t), atol=1e-12)


@pytest.mark.parametrize("degree,k", [(2, 1), (3, 1), (2, 2)])
def test_series_solves_homotopy_to_order_deg(degree, k):
    """H(x(s), t0+s) = O(s^{deg+1}) with the reference-order evaluator on
    coefficient series (a planted-solution check as in dbl2_row_setup,
    Newton/dbl2_newton_method.cpp:823-991)."""
    from phcpack_amd.tracker import Tracker  # noqa: F401  (import check only)
    rng = np.random.default_rng(10 + degree + k)
    n, K, deg = 3, 2, 8
    p = dense_system(n, degree, rng)
    hom, sols = total_degree_homotopy(p, rng, k=k)
    prog = compile_homotopy(hom, deg)
    # a point on the path at t0: track the first path to t0 with the oracle
    t0 = 0.0
    x0 = sols[1]
    s, ok = oracle_series(prog, cplx_to_md(x0, K), K, t0)
    assert ok
    D = deg + 1
    total = np.zeros((n, D), np.complex128)
    for w, sysm in hom.components:
        ws = weight_series(w, t0, D, K)
        for i in range(n):
            out = ref_eval_component(sysm, i, K, s, ws)
            total[i] += out[:K, n].sum(0) + 1j * out[K:, n].sum(0)
    scale = np.abs(s[:K].sum(0) + 1j * s[K:].sum(0)).max() + 1
    # order 0 is H at the start root, which is given in double precision
    assert np.abs(total[:, 0]).max() < 1e-14
    # orders 1..deg are solved in double-double
    assert np.abs(total[:, 1:]).max() < 1e-27 * scale ** degree
    # and the series is not trivially zero
    assert np.abs(s[0, :, 1:]).max() > 1e-3


def test_decoupled_closed_form_oracle():
    rng = np.random.default_rng(11)
    n = 3
    a = np.exp(2j * np.pi * rng.uniform(size=n)) * rng.uniform(0.5, 2.0, n)
    eqs = []
    for i in range(n):
        e = [0] * n
        e[i] = 2
        eqs.append([(1.0 + 0j, tuple(e)), (complex(-a[i]), (0,) * n)])
    hom, sols = total_degree_homotopy(System(n, eqs), rng)
    prog = compile_homotopy(hom, 8)
    out = oracle_track(prog, points_to_planes(sols), 2, TrackParams())
    assert np.all(out["status"] == 1)
    x = planes_to_points(out["endp"], 2)
    want = np.array([np.array(s) * np.sqrt(a) for s in itertools.product([1, -1], repeat=n)])
    for xi in x:
        assert np.min(np.abs(want - xi).max(axis=1)) < 1e-14


@pytest.mark.parametrize("n", [3, 4])
def test_bezout_count_distinct_endpoints(n):
    rng = np.random.default_rng(20 + n)
    p = dense_system(n, 2, rng)
    hom, sols = total_degree_homotopy(p, rng)
    prog = compile_homotopy(hom, 8)
    out = oracle_track(prog, points_to_planes(sols), 2, TrackParams())
    assert np.all(out["status"] == 1)
    x = planes_to_points(out["endp"], 2)
    assert len(x) == 2 ** n
    dmin = min(np.abs(x[i] - x[j]).max() for i, j in itertools.combinations(range(len(x)), 2))
    assert dmin > 1e-6
    assert max(np.abs(p.eval(xi)).max() for xi in x) < 1e-12


def test_fabry_pole_step_closed_form():
    """For x(t) with a pole at distance r the Fabry ratio recovers r: the
    scalar homotopy (1-t) x - gamma... built as x*(r - t) = r*c gives
    x = r*c/(r-t), coefficients c*(1/r)^k, ratio r (Track/dbl_fabry_host.cpp:9-64)."""
    r, c = 1.7, 0.8 + 0.3j
    # H = w_p(t)*p + w_q(t)*q with p = x - 0 ... encode (r - t) x - r c = 0:
    # component A: weight (r - t), system x ; component B: weight 1, system -r c
    from phcpack_amd.homotopy import Homotopy
    hom = Homotopy(System(1, [[(1.0 + 0j, (1,))]]), System(1, [[(complex(-r * c), (0,))]]), 1.0)
    hom.components = [(np.array([r, -1.0, 0.0], np.complex128), hom.target),
                      (np.array([1.0, 0.0, 0.0], np.complex128), hom.start)]
    prog = compile_homotopy(hom, 8)
    s, ok = oracle_series(prog, cplx_to_md(np.array([c]), 2), 2, 0.0)
    assert ok
    z = s[:2, 0].sum(0) + 1j * s[2:, 0].sum(0)
    assert np.allclose(z, c * (1.0 / r) ** np.arange(9), rtol=1e-14)
    assert abs(abs(z[5] / z[6]) - r) < 1e-12


# ------------------------------------------------------------ Pade predictor

def md_coeffs(z, K):
    """complex128 [m] -> [m][2K] (limbs re, then im; exact doubles)"""
    out = np.zeros((len(z), 2 * K))
    out[:, 0], out[:, K] = np.real(z), np.imag(z)
    return out


def md_value(v, K):
    return v[:K].sum() + 1j * v[K:].sum()


@pytest.mark.parametrize("L", [1, 2, 5, 7])
def test_pade_coefficients_match_scipy(L):
    """[L/1] numerator and denominator against scipy.interpolate.pade (a
    library routine that solves the same Toeplitz system): a swapped index
    or sign in the denominator system or in the numerator sum fails."""
    from scipy.interpolate import pade
    from oraclelib import oracle_pade
    rng = np.random.default_rng(30 + L)
    c = rng.normal(size=L + 2) + 1j * rng.normal(size=L + 2)
    cc, x0, val = oracle_pade(1, md_coeffs(c, 1), L, 1, 0.3)
    p, q = pade(c, 1, L)
    num = np.concatenate([[x0[0] + 1j * x0[1]], cc[1:L + 1, 0] + 1j * cc[1:L + 1, 1]])
    assert np.allclose(num, p.coeffs[::-1], rtol=1e-12, atol=1e-13)
    d = cc[L + 1, 0] + 1j * cc[L + 1, 1]
    assert np.allclose([d, 1.0], q.coeffs, rtol=1e-12, atol=1e-13)
    assert abs(md_value(val, 1) - p(0.3) / q(0.3)) < 1e-12 * abs(p(0.3) / q(0.3))


@pytest.mark.parametrize("K", [1, 2])
@pytest.mark.parametrize("L", [1, 4, 6])
def test_pade_reproduces_a_rational_curve(K, L):
    """x(t) = (a + b t)/(1 - t/2) with dyadic a, b: its Taylor coefficients
    are exact doubles, every [L/1] approximant is the function itself, so
    the prediction at s agrees with the exact quotient (Fractions) to the
    working precision while the Taylor polynomial of the same coefficients
    misses by the truncation term."""
    from fractions import Fraction as F
    from oraclelib import oracle_pade
    a, b = 0.75 + 0.5j, 0.25 - 1.0j
    c = np.array([a] + [(a / 2 + b) * 0.5 ** (k - 1) for k in range(1, L + 2)])
    s = 0.3
    cc, x0, val = oracle_pade(K, md_coeffs(c, K), L, 1, s)
    # d = -1/2 exactly, the numerator is a + b s
    assert np.array_equal(cc[L + 1], md_coeffs([-0.5], K)[0])
    assert np.array_equal(cc[1], md_coeffs([b], K)[0])
    assert not np.any(cc[2:L + 1])
    fs = F(s)
    den = 1 - fs / 2
    want_re = (F(a.real) + F(b.real) * fs) / den
    want_im = (F(a.imag) + F(b.imag) * fs) / den
    got_re = sum(F(v) for v in val[:K])
    got_im = sum(F(v) for v in val[K:])
    tol = 1e-15 if K == 1 else 1e-30
    assert abs(float(got_re - want_re)) < tol and abs(float(got_im - want_im)) < tol
    _, _, tay = oracle_pade(K, md_coeffs(c, K), L + 1, 0, s)
    assert abs(md_value(tay, K) - complex(float(want_re), float(want_im))) > 1e-4 * (s / 2) ** (L + 1)


@pytest.mark.parametrize("L,M", [(5, 2), (4, 2), (2, 2), (1, 2), (4, 3)])
def test_pade_higher_denominators_match_scipy(L, M):
    """[L/M] through the oracle's lufac/lusolve restatement: the denominator
    against scipy.interpolate.pade, the numerator against the truncated
    product of denominator and series (numpy convolution).  For L < M the
    reference's Denominator_System puts zeros where the series index falls
    below 1, so c_0 drops out of the denominator system: scipy then gets
    the series with c_0 = 0 (the numerator keeps c_0)."""
    from scipy.interpolate import pade
    from oraclelib import oracle_pade
    rng = np.random.default_rng(40 + 10 * L + M)
    c = rng.normal(size=L + M + 1) + 1j * rng.normal(size=L + M + 1)
    cc, x0, val = oracle_pade(1, md_coeffs(c, 1), L, M, 0.3)
    cmod = c.copy()
    if L < M:
        cmod[0] = 0.0
    _, q = pade(cmod, M, L)
    num = np.concatenate([[x0[0] + 1j * x0[1]], cc[1:L + 1, 0] + 1j * cc[1:L + 1, 1]])
    den = np.concatenate([[1.0], cc[L + 1:L + M + 1, 0] + 1j * cc[L + 1:L + M + 1, 1]])
    assert np.allclose(den, q.coeffs[::-1], rtol=1e-10, atol=1e-11)
    assert np.allclose(num, np.convolve(den, c)[:L + 1], rtol=1e-10, atol=1e-11)
    want = np.polyval(num[::-1], 0.3) / np.polyval(den[::-1], 0.3)
    assert abs(md_value(val, 1) - want) < 1e-12 * abs(want)


@pytest.mark.parametrize("K", [1, 2])
@pytest.mark.parametrize("L", [2, 3, 5])
def test_pade_2_reproduces_a_rational_curve(K, L):
    """x(t) = (a + b t)/((1 - t/2)(1 - t/4)): dyadic Taylor coefficients,
    denominator 1 - 3t/4 + t^2/8; every [L/2] approximant with L >= 2 is
    the function itself (at L = 1 the reference's system leaves out c_0)"""
    from fractions import Fraction as F
    from oraclelib import oracle_pade
    ar, ai, br, bi = F(3, 4), F(1, 2), F(1, 4), F(-1)
    # series of 1/((1-t/2)(1-t/4)) = sum_k (2^{k+1} - 1)/4^k t^k
    g = [F(2 ** (k + 1) - 1, 4 ** k) for k in range(L + 3)]
    cre = [ar * g[0]] + [ar * g[k] + br * g[k - 1] for k in range(1, L + 3)]
    cim = [ai * g[0]] + [ai * g[k] + bi * g[k - 1] for k in range(1, L + 3)]
    c = np.array([float(x) + 1j * float(y) for x, y in zip(cre, cim)])
    assert all(F(float(x)) == x for x in cre + cim)  # exact doubles
    s = 0.3
    cc, x0, val = oracle_pade(K, md_coeffs(c, K), L, 2, s)
    tol = 1e-14 if K == 1 else 1e-29
    # the complex elimination rounds (the 2x2 system of coefficients ~2^-L
    # is mildly conditioned): d_1 = -3/4, d_2 = 1/8 to working precision
    for idx, want in ((L + 1, F(-3, 4)), (L + 2, F(1, 8))):
        assert abs(float(sum(F(v) for v in cc[idx][:K]) - want)) < tol
        assert abs(float(sum(F(v) for v in cc[idx][K:]))) < tol
    fs = F(s)
    den = (1 - fs / 2) * (1 - fs / 4)
    assert abs(float(sum(F(v) for v in val[:K]) - (ar + br * fs) / den)) < tol
    assert abs(float(sum(F(v) for v in val[K:]) - (ai + bi * fs) / den)) < tol


@pytest.mark.parametrize("K", [1, 2])
def test_pade_singular_denominator(K):
    """c_L = 0: the reference's Pade keeps the denominator 1 and takes the
    Taylor numerator when c_{L+1} = 0 too (a polynomial curve), else a zero
    numerator (lufac reports the singular 1x1 system)."""
    from oraclelib import oracle_pade
    L, s = 3, 0.25
    poly = np.array([1.0 + 2j, -0.5 + 1j, 0.25, 0.0, 0.0])
    cc, x0, val = oracle_pade(K, md_coeffs(poly, K), L, 1, s)
    assert abs(md_value(val, K) - np.polyval(poly[::-1], s)) < 1e-15
    assert not np.any(cc[L + 1])
    other = np.array([1.0 + 2j, -0.5 + 1j, 0.25, 0.0, 3.0])
    cc, x0, val = oracle_pade(K, md_coeffs(other, K), L, 1, s)
    assert not np.any(val) and not np.any(x0) and not np.any(cc[1:L + 2])
    # [2/2] of a quadratic: the 2x2 system [[c1 c2] [c2 c3]] = [[c1 c2] [c2 0]]
    # is regular; of a linear series it is singular with a zero right-hand side
    lin = np.array([1.0 + 2j, -0.5 + 1j, 0.0, 0.0, 0.0])
    cc, x0, val = oracle_pade(K, md_coeffs(lin, K), 2, 2, s)
    assert abs(md_value(val, K) - np.polyval(lin[::-1], s)) < 1e-15
    assert not np.any(cc[3:5])


@pytest.mark.parametrize("K", [1, 2])
def test_taylor_predictor_of_lower_degree(K):
    """pade_dendeg 0 with numerator degree L: the Taylor polynomial of
    degree L of the series (Horner against numpy's polyval)"""
    from oraclelib import oracle_pade
    rng = np.random.default_rng(5)
    c = rng.normal(size=9) + 1j * rng.normal(size=9)
    for L in (1, 3, 8):
        _, _, val = oracle_pade(K, md_coeffs(c, K), L, 0, 0.2)
        assert abs(md_value(val, K) - np.polyval(c[:L + 1][::-1], 0.2)) < 1e-14


def rational_path_homotopy(a, b, r):
    """(r - t) x - (a + b t) = 0: x(t) = (a + b t)/(r - t), a [1/1]
    rational curve with a pole at t = r"""
    from phcpack_amd.homotopy import Homotopy
    hom = Homotopy(System(1, [[(1.0 + 0j, (1,))]]), System(1, [[(-1.0 + 0j, (0,))]]), 1.0)
    hom.components = [(np.array([r, -1.0, 0.0], np.complex128), hom.target),
                      (np.array([a, b, 0.0], np.complex128), hom.start)]
    return hom


@pytest.mark.parametrize("K", [1, 2])
def test_pade_tracker_on_a_rational_path(K):
    """On x(t) = (a + b t)/(r - t) the [5/1] predictor is exact: no step is
    rejected and every corrector converges at its first Newton iteration
    (then one extra iteration unless the update or residual is exactly
    zero, as LU_Newton_Steps does), where the Taylor predictor of the same
    series (degree 6, same pole steps: the equation is linear, so there is
    no curvature step) needs more; both end at (a + b)/(r - 1)."""
    a, b, r = 0.6 + 0.2j, -0.3 + 0.9j, 1.25
    prog = compile_homotopy(rational_path_homotopy(a, b, r), 8)
    start = points_to_planes(np.array([[a / r]]), 1)
    want = (a + b) / (r - 1)
    tol = 1e-13 if K == 1 else 1e-26
    outs = {}
    for name, prm in (("pade", TrackParams(pade_numdeg=5, pade_dendeg=1, endsteps=0)),
                      ("taylor", TrackParams(pade_numdeg=0, pade_dendeg=0, endsteps=0))):
        o = oracle_track(prog, start, K, prm)
        assert o["status"][0] == 1 and o["nrej"][0] == 0
        x = planes_to_points(o["endp"], K)[0, 0]
        assert abs(x - want) < tol * abs(want), name
        outs[name] = o
    wp, wt = outs["pade"]["work"][:, 0], outs["taylor"]["work"][:, 0]
    assert outs["pade"]["nsteps"][0] == outs["taylor"]["nsteps"][0]
    assert wp[3] <= wp[1] <= 2 * wp[3]  # converged at once, plus the extra step
    assert wt[1] > wp[1]


def test_interior_failures_select_precision_escalation():
    """the mixed method tracks again in double-double the paths that failed
    inside the interval for numerical reasons (minstep, maxsteps, singular
    Jacobian at t < 0.99), not successes, divergences or failures at t -> 1"""
    from phcpack_amd.tracker import interior_failures
    status = np.array([1, 2, 3, 4, 5, 2, 2, 4, 0])
    tend = np.array([1.0, 0.5, 0.3, 0.98, 0.2, 0.995, 0.989, 1.0, 0.1])
    assert interior_failures(status, tend).tolist() == [1, 2, 3, 6]
    assert interior_failures(status, tend, t_interior=0.999).tolist() == [1, 2, 3, 5, 6]


def test_dd_rescues_an_fp64_interior_failure():
    """paths that FP64 tracking stops inside the interval (here by a step
    budget of 12, which ends most paths at maxsteps) are the ones the rule
    selects, and dd tracking from their start with the default parameters
    finishes every one of them"""
    from phcpack_amd.tracker import interior_failures
    rng = np.random.default_rng(5)
    hom, sols = total_degree_homotopy(dense_system(4, 2, rng), rng)
    prog = compile_homotopy(hom, 8)
    planes = points_to_planes(sols)
    o1 = oracle_track(prog, planes, 1, TrackParams(tolend=1e-13, maxsteps=12))
    idx = interior_failures(o1["status"], o1["tend"])
    assert set(idx.tolist()) == set(np.nonzero((o1["status"] == 3) & (o1["tend"] < 0.99))[0].tolist())
    assert len(idx) > 0
    o2 = oracle_track(prog, planes[:, idx], 2, TrackParams(tolend=1e-28))
    assert np.all(o2["status"] == 1)
