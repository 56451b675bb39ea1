"""Pins the step rule of the tracker oracle (LU_Prediction, Hesse_Pade,
Predictor_Feedback and LU_Newton_Steps of the Ada trackers,
Deformations/Curves/standard_predictor_convolutions.adb:558-1185,
predictor_corrector_loops.adb:16-83) piece by piece against closed forms
and independent numpy computations, and whole step sequences by their
invariants and by recomputing the halving decisions.
"""
import itertools

import numpy as np
import pytest

from oraclelib import (oracle_hessian_sum, oracle_mixed_residual, oracle_pade,
                       oracle_pade_error, oracle_root, oracle_series, oracle_sigma_max,
                       oracle_track_trace)
from phcpack_amd.homotopy import (Homotopy, System, compile_homotopy, dense_system,
                                  total_degree_homotopy)
from phcpack_amd.tracker import TrackParams, points_to_planes

EPS = 2.0 ** -52


def cplx_md(z, K):
    z = np.asarray(z, np.complex128)
    out = np.zeros((2 * K,) + z.shape)
    out[0], out[K] = z.real, z.imag
    return out


def custom_homotopy(comps):
    """a Homotopy whose components are the given (weight [1, t, t^2], System)"""
    hom = Homotopy(comps[0][1], comps[-1][1], 1.0)
    hom.components = [(np.asarray(w, np.complex128), s) for w, s in comps]
    return hom


# ------------------------------------------------------------ the pieces --

def test_root_is_the_kth_root():
    """against the k-th root in 50-digit decimal arithmetic"""
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    rng = np.random.default_rng(0)
    for _ in range(400):
        x = 10.0 ** rng.uniform(-40, 40)
        k = int(rng.integers(1, 17))
        r = oracle_root(x, k)
        want = Decimal(x) ** (Decimal(1) / Decimal(k))
        assert abs(Decimal(r) - want) <= Decimal(2 * EPS) * want
    assert oracle_root(0.0, 8) == 0.0
    assert oracle_root(np.inf, 8) == np.inf
    assert np.isnan(oracle_root(np.nan, 8))


@pytest.mark.parametrize("n", [1, 2, 5, 12, 16])
def test_sigma_max_matches_svd(n):
    """random matrices: the two largest singular values are apart enough
    that (sigma_2/sigma_1)^256 leaves at most 1e-9 of relative error"""
    rng = np.random.default_rng(n)
    for _ in range(5):
        M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        want = np.linalg.svd(M, compute_uv=False)[0]
        assert abs(oracle_sigma_max(M) - want) <= 1e-9 * want
    # the inverse gives sigma_min, as the curvature step uses it
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    want = np.linalg.svd(M, compute_uv=False)[-1]
    assert abs(1.0 / oracle_sigma_max(np.linalg.inv(M)) - want) <= 1e-10 * want


def test_sigma_max_close_and_equal_singular_values():
    """Powers of the Gram matrix separate close singular values slowly: the
    Rayleigh quotient stays between sigma_2 and sigma_1 and within
    1e-3 n of sigma_1; equal ones (a scaled unitary matrix) are exact."""
    rng = np.random.default_rng(7)
    n = 12
    U, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    for gap in (1e-1, 1e-2, 1e-3, 1e-6):
        s = np.linspace(1.0, 0.1, n)
        s[1] = 1.0 - gap
        M = U @ np.diag(s) @ V.conj().T
        est = oracle_sigma_max(M)
        assert s[1] * (1 - 1e-14) <= est <= 1.0 + 1e-14
        assert 1.0 - est <= 1e-3 * n
    assert abs(oracle_sigma_max(3.0 * U) - 3.0) <= 1e-14
    assert oracle_sigma_max(np.zeros((4, 4))) == 0.0
    assert abs(oracle_sigma_max(np.diag([0.5, 2.0, -1.0])) - 2.0) <= 1e-15


def test_mixed_residual_formula():
    """mean_i |f_i| / (|f|_i + 1), |f|_i with the moduli of the coefficients
    of the homotopy (components of a monomial added first) at |x|
    (standard_mixed_residuals.adb:472-484)"""
    rng = np.random.default_rng(3)
    for n, deg, k in ((3, 2, 2), (4, 3, 1)):
        p = dense_system(n, deg, rng)
        hom, _ = total_degree_homotopy(p, rng, k=k)
        prog = compile_homotopy(hom, 8)
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        t = 0.43
        f = hom.eval(x, t)
        rad = np.zeros(n)
        for i in range(n):
            coef = {}
            for w, sysm in hom.components:
                wt = w[0] + w[1] * t + w[2] * t * t
                for c, e in sysm.eqs[i]:
                    coef[e] = coef.get(e, 0) + wt * c
            rad[i] = sum(abs(c) * np.prod(np.abs(x) ** np.array(e)) for e, c in coef.items())
        want = np.mean(np.abs(f) / (rad + 1.0))
        got = oracle_mixed_residual(prog, cplx_md(x, 1)[:, None, :], 1, t)
        assert abs(got - want) <= 1e-12 * want


@pytest.mark.parametrize("deg", [2, 3])
def test_hessian_sum_matches_numpy(deg):
    """sum_i sigma_max(Hessian_i)^2: constants from t = 0 for quadratic
    equations (Cached_Singular_Values), Hessians at (x, t) for cubic ones"""
    rng = np.random.default_rng(4 + deg)
    n = 4
    p = dense_system(n, deg, rng)
    hom, _ = total_degree_homotopy(p, rng, k=2)
    prog = compile_homotopy(hom, 8)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    t = 0.3 if deg == 3 else 0.0

    def hess(i, x, t):
        H = np.zeros((n, n), np.complex128)
        for w, sysm in hom.components:
            wt = w[0] + w[1] * t + w[2] * t * t
            for c, e in sysm.eqs[i]:
                e = np.array(e)
                for a in range(n):
                    for b in range(n):
                        ee = e.copy()
                        m = ee[a]
                        ee[a] -= 1
                        m2 = ee[b]
                        ee[b] -= 1
                        if m and m2:
                            H[a, b] += wt * c * m * m2 * np.prod(x ** ee)
        return H

    want = sum(np.linalg.svd(hess(i, x, t), compute_uv=False)[0] ** 2 for i in range(n))
    got = oracle_hessian_sum(prog, cplx_md(x, 1)[:, None, :], 1, t)
    assert abs(got - want) <= 1e-10 * want
    if deg == 2:
        # the constant Hessians do not depend on the point or on t
        got2 = oracle_hessian_sum(prog, cplx_md(2 * x, 1)[:, None, :], 1, 0.9)
        assert got2 == got


def md_coeffs(z, K):
    out = np.zeros((len(z), 2 * K))
    out[:, 0], out[:, K] = np.real(z), np.imag(z)
    return out


def test_pade_error_estimate():
    """Solution_Error_Estimate (homotopy_pade_approximants.adb:597-622)"""
    rng = np.random.default_rng(5)
    L, M = 5, 1
    # geometric series: the [L/1] approximant is exact, the estimate vanishes
    # and is replaced by 1
    z = 0.7 * (1 / 0.9) ** np.arange(9)
    assert oracle_pade_error(1, md_coeffs(z, 1), L, M)[2] == 1.0
    # random series: c_{L+M+2} + d_1 c_{L+M+1}
    z = rng.standard_normal(9) + 1j * rng.standard_normal(9)
    c, _, est = oracle_pade_error(1, md_coeffs(z, 1), L, M)
    d1 = -z[L + 1] / z[L]
    assert abs(est - (z[L + M + 2] + d1 * z[L + M + 1])) <= 1e-13 * abs(est)
    # [4/2]: c_8 + d_1 c_7 + d_2 c_6 with the denominator of the 2x2 system
    L2, M2 = 4, 2
    c, _, est = oracle_pade_error(1, md_coeffs(z, 1), L2, M2)
    A = np.array([[z[3], z[4]], [z[4], z[5]]])
    b = -np.array([z[5], z[6]])
    sol = np.linalg.solve(A, b)
    d = sol[::-1]
    assert abs(est - (z[8] + d[0] * z[7] + d[1] * z[6])) <= 1e-12 * abs(est)
    # Taylor (M = 0): c_{L+2}
    c, _, est = oracle_pade_error(1, md_coeffs(z, 1), 6, 0)
    assert est == z[8]
    # small high coefficients: the threshold index drops to <= L, where the
    # numerator and the estimate agree: replaced by 1
    z2 = z.copy()
    z2[6:] = 1e-7
    assert oracle_pade_error(1, md_coeffs(z2, 1), L, M)[2] == 1.0
    assert oracle_pade_error(1, md_coeffs(np.zeros(9) + 1e-9, 1), L, M)[2] == 1.0


# --------------------------------------------------- closed-form steps --

def test_pole_step_on_a_pole():
    """x(t) = r c/(r - t): the series is geometric, Fabry's ratio (offset 2)
    is the distance r - t to the pole and the step is pbeta*(r - t) (no
    curvature: the equation is linear)"""
    r, c = 0.15, 0.8 + 0.3j
    hom = custom_homotopy([([r, -1.0, 0.0], System(1, [[(1.0 + 0j, (1,))]])),
                           ([1.0, 0.0, 0.0], System(1, [[(complex(-r * c), (0,))]]))])
    prog = compile_homotopy(hom, 8)
    prm = TrackParams(tolend=1e-13)
    out, steps = oracle_track_trace(prog, points_to_planes(np.array([[c]])), 1, prm)
    assert len(steps) > 5
    for s in steps[:6]:
        dist = r - s["t"]
        assert abs(s["r"] - dist) <= 1e-12 * dist
        assert s["curv"] == np.inf and s["which"] == 0
        assert s["rule"] == s["pole"] == prm.pbeta * s["r"]
    # far from the pole maxstep binds
    hom.components[0] = (np.array([1.7, -1.0, 0.0], np.complex128), hom.components[0][1])
    prog = compile_homotopy(hom, 8)
    out, steps = oracle_track_trace(prog, points_to_planes(np.array([[c * 1.7 / 1.7]])), 1, prm)
    assert steps[0]["which"] == 2 and steps[0]["rule"] == prm.maxstep


def test_fabry_is_zero_component():
    """a component whose path is constant has zero coefficients: Is_Zero
    drops it from the minimum; with every component constant the pole step
    is maxstep (convergence_radius_estimates.adb:28-38,138-163)"""
    a, cq, cp = 0.3 - 0.2j, np.exp(0.4j), np.exp(2.1j)
    eq2 = [(1.0 + 0j, (0, 1)), (complex(-a), (0, 0))]
    hom = custom_homotopy([([1.0, -1.0, 0.0], System(2, [[(1.0 + 0j, (2, 0)), (complex(-cq), (0, 0))], eq2])),
                           ([0.0, 1.0, 0.0], System(2, [[(1.0 + 0j, (2, 0)), (complex(-cp), (0, 0))], eq2]))])
    prog = compile_homotopy(hom, 8)
    x0 = np.array([[np.sqrt(cq), a]])
    s, ok = oracle_series(prog, cplx_md(x0[0], 1), 1, 0.0)
    assert ok and np.all(s[:, 1, 1:] == 0.0)
    z = s[0, 0] + 1j * s[1, 0]
    out, steps = oracle_track_trace(prog, points_to_planes(x0), 1, TrackParams(tolend=1e-13))
    assert abs(steps[0]["r"] - abs(z[5] / z[6])) <= 1e-14 * steps[0]["r"]
    assert out["status"][0] == 1
    hom1 = custom_homotopy([([1.0, -1.0, 0.0], System(1, [[(1.0 + 0j, (1,)), (complex(-a), (0,))]])),
                            ([0.0, 1.0, 0.0], System(1, [[(1.0 + 0j, (1,)), (complex(-a), (0,))]]))])
    out, steps = oracle_track_trace(compile_homotopy(hom1, 8), points_to_planes(np.array([[a]])), 1,
                                    TrackParams(tolend=1e-13))
    assert steps[0]["r"] == -1.0 and steps[0]["pole"] == 0.1 and steps[0]["which"] == 2


def series_sqrt(q, x0):
    """Taylor coefficients of sqrt(q(t)) with constant term x0"""
    D = len(q)
    x = np.zeros(D, np.complex128)
    x[0] = x0
    for k in range(1, D):
        x[k] = (q[k] - sum(x[i] * x[k - i] for i in range(1, k))) / (2 * x0)
    return x


def series_div(a, b):
    D = len(a)
    q = np.zeros(D, np.complex128)
    for k in range(D):
        q[k] = (a[k] - sum(q[i] * b[k - i] for i in range(k))) / b[0]
    return q


def quadratic_1d(seed):
    """H = gamma (1-t)^2 (x^2 - c) + t^2 (x^2 - c'), x(0) = sqrt(c): the
    Taylor series of x(t) = sqrt(N/Dn) at t = 0 computed independently"""
    rng = np.random.default_rng(seed)
    cp = np.exp(2j * np.pi * rng.uniform())
    hom, sols = total_degree_homotopy(System(1, [[(1.0 + 0j, (2,)), (complex(-cp), (0,))]]), rng, k=2)
    c = -hom.start.eqs[0][1][0]
    g = hom.gamma
    D = 9
    Dn = np.zeros(D, np.complex128)
    Dn[:3] = [g, -2 * g + 0, g + 1]
    N = np.zeros(D, np.complex128)
    N[:3] = [g * c, -2 * g * c, g * c + cp]
    x = series_sqrt(series_div(N, Dn), sols[0][0])
    return hom, sols, x, g, c, cp


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_first_step_closed_form(seed):
    """The first step of LU_Prediction on a scalar quadratic: pole step
    0.5 |c5/c6|, curvature step (cbeta eta / |c8 + d1 c7|)^(1/8) with
    eta = 2 sigma_min(J) / |Hessian| = 2 |2 gamma x0| / |2 gamma|,
    d1 = -c6/c5 ([5/1]), and the minimum with maxstep"""
    hom, sols, x, g, c, cp = quadratic_1d(seed)
    prog = compile_homotopy(hom, 8)
    prm = TrackParams(tolend=1e-13)
    out, steps = oracle_track_trace(prog, points_to_planes(sols[:1]), 1, prm)
    s0 = steps[0]
    assert abs(s0["r"] - abs(x[5] / x[6])) <= 1e-12 * s0["r"]
    assert s0["hs2"] == pytest.approx(abs(2 * g) ** 2, rel=1e-15)
    assert s0["sigma1"] == pytest.approx(abs(2 * g * x[0]), rel=1e-13)
    eta = 2 * abs(2 * g * x[0]) / abs(2 * g)
    est = x[8] + (-x[6] / x[5]) * x[7]
    nrm = abs(est) if abs(est.real) + abs(est.imag) >= 1e-5 else 1.0
    curv = (prm.cbeta * eta / nrm) ** (1 / 8)
    assert s0["nrm"] == pytest.approx(nrm, rel=1e-11)
    assert s0["curv"] == pytest.approx(curv, rel=1e-11)
    assert s0["rule"] == min(prm.pbeta * s0["r"], s0["curv"], prm.maxstep)


def test_feedback_halvings_recomputed():
    """The halvings of Predictor_Feedback: from the step of the rule, halve
    while the mixed residual of the [5/1] prediction is >= alpha; recomputed
    with numpy from the independent series of test_first_step_closed_form"""
    hom, sols, x, g, c, cp = quadratic_1d(2)
    prog = compile_homotopy(hom, 8)

    def mixres(step):
        d1 = -x[6] / x[5]
        num = np.array([x[0]] + [x[i] + d1 * x[i - 1] for i in range(1, 6)])
        xp = np.polyval(num[::-1], step) / (1 + d1 * step)
        t = step
        a2 = g * (1 - t) ** 2 + t * t
        a0 = -(g * (1 - t) ** 2 * c + t * t * cp)
        f = a2 * xp * xp + a0
        return abs(f) / (abs(a2) * abs(xp) ** 2 + abs(a0) + 1.0)

    for alpha in (1e-6, 1e-8, 1e-10):
        prm = TrackParams(tolend=1e-13, alpha=alpha)
        out, steps = oracle_track_trace(prog, points_to_planes(sols[:1]), 1, prm)
        s0 = steps[0]
        step, want = s0["step1"], 0
        while mixres(step) >= alpha:
            assert abs(mixres(step) / alpha - 1) > 1e-6   # no tie at the threshold
            step /= 2
            want += 1
            if step < prm.minstep:
                break
        assert s0["nfb"] == want
        if s0["ncr"] == 0:
            assert s0["step"] == step
    assert want >= 1   # the smallest alpha forces halvings


@pytest.mark.parametrize("n,deg,K", [(4, 2, 1), (3, 3, 1), (3, 2, 2)])
def test_step_sequence_invariants(n, deg, K):
    """Every logged step: the rule is Minimum(pole, curvature, maxstep) with
    the reference's tie order; pole = pbeta*r or maxstep; eta and the
    curvature step follow from sigma_min, the Hessian sum and nrm;
    Bounded_Update sets 1 - t; the final step is the first one halved once
    per feedback or corrector failure; accepted steps move t to t + step
    (1 exactly after Bounded_Update); steps never exceed maxstep"""
    rng = np.random.default_rng(40 + n + deg)
    p = dense_system(n, deg, rng)
    hom, sols = total_degree_homotopy(p, rng, k=2)
    prog = compile_homotopy(hom, 8)
    prm = TrackParams(tolend=1e-13 if K == 1 else 1e-28)
    seen_bound = seen_half = 0
    for path in range(min(len(sols), 6)):
        out, steps = oracle_track_trace(prog, points_to_planes(sols, K) if K > 1 else points_to_planes(sols),
                                        K, prm, path=path)
        assert out["status"][path] == 1
        assert len(steps) == out["nsteps"][path]
        assert sum(s["nfb"] + s["ncr"] for s in steps) == out["nrej"][path]
        w = out["work"][:, path]
        assert w[4] + w[5] + w[6] == w[0]
        for s in steps:
            vals = (s["pole"], s["curv"], prm.maxstep)
            i = 0 if vals[0] < vals[1] and vals[0] < vals[2] else (
                2 if vals[0] < vals[1] else (1 if vals[1] < vals[2] else 2))
            assert s["which"] == i and s["rule"] == vals[i]
            assert s["pole"] == (prm.maxstep if s["r"] < 0 else prm.pbeta * s["r"])
            assert s["eta"] == (2.0 * s["sigma1"]) / np.sqrt(s["hs2"])
            assert s["curv"] == oracle_root((prm.cbeta * s["eta"]) / s["nrm"], 8)
            if deg == 2:
                assert s["hs2"] == 4.0 * n
            if s["t"] + s["rule"] >= 1.0 - prm.minstep:
                assert s["step1"] == 1.0 - s["t"]
                seen_bound += 1
            else:
                assert s["step1"] == s["rule"]
            assert s["step1"] <= prm.maxstep
            assert s["step"] == s["step1"] / 2.0 ** (s["nfb"] + s["ncr"])
            seen_half += s["nfb"] + s["ncr"] > 0
            assert s["accepted"] == 1
            if s["nfb"] + s["ncr"] == 0 and s["t"] + s["rule"] >= 1.0 - prm.minstep:
                assert s["t_after"] == 1.0
            else:
                assert s["t_after"] == s["t"] + s["step"]
            # converged corrector plus its extra step
            assert s["nit"] >= 2
        assert steps[-1]["t_after"] == 1.0
    assert seen_bound >= 1
    # at step 0 the cubic's Hessians are those of gamma*(x_i^3 - c_i) at the
    # start root: sigma_i = 6 |gamma x_i| = 6
    if deg == 3:
        out, steps = oracle_track_trace(prog, points_to_planes(sols), 1, prm, path=0)
        assert steps[0]["hs2"] == pytest.approx(36.0 * n, rel=1e-13)
