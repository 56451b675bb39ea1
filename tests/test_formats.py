"""PHCpack's text formats: the reference's own example systems and their
stored solutions (tests/golden/demo, copied from src/Demo) pin the parser
and the solution reader: a misread coefficient, exponent or unknown order
would leave the stored solutions far from zero of the parsed system."""
import os

import numpy as np
import pytest

from phcpack_amd.formats import (ParseError, parse_solutions, parse_system, write_solutions,
                                 write_system)

DEMO = os.path.join(os.path.dirname(__file__), "golden", "demo")
NAMES = ("katsura5", "cyclic5", "redeco5", "mickey", "utbikker")
# number of unknowns, degrees, stored solutions (from the files' own headers)
EXPECT = {
    "katsura5": (6, [2, 2, 2, 2, 2, 1], 32),
    "cyclic5": (5, [1, 2, 3, 4, 5], 7),
    "redeco5": (5, [2, 2, 2, 1, 1], 8),
    "mickey": (2, [2, 2], 4),
    "utbikker": (4, [2, 2, 3, 3], 36),
}


def read(name):
    with open(os.path.join(DEMO, name)) as f:
        return f.read()


def scaled_residual(system, x):
    """max_i |p_i(x)| / (sum_m |c_m| max(1, |x|)^deg_i)"""
    s = max(1.0, float(np.max(np.abs(x))))
    out = 0.0
    for eq in system.eqs:
        d = max(sum(e) for _, e in eq)
        v = abs(sum(c * np.prod(np.power(x, e)) for c, e in eq))
        out = max(out, v / (sum(abs(c) for c, _ in eq) * s ** d))
    return out


@pytest.mark.parametrize("name", NAMES)
def test_demo_system_and_stored_solutions(name):
    text = read(name)
    system, symbols = parse_system(text)
    n, degs, nsol = EXPECT[name]
    assert system.n == n and system.degrees() == degs
    sols = parse_solutions(text, symbols)
    assert len(sols) == nsol
    for s in sols:
        assert s["x"].shape == (n,)
        assert scaled_residual(system, s["x"]) < 1e-13, name
    # the diagnostics line is read too
    assert all("res" in s and "rco" in s for s in sols)


def test_katsura5_unknown_order_and_coefficients():
    system, symbols = parse_system(read("katsura5"))
    assert symbols == ["x", "y", "z", "t", "u", "v"]
    # the last equation: 2*x+2*y+2*z+2*t+2*u+v-1
    last = {e: c for c, e in system.eqs[5]}
    assert last[(1, 0, 0, 0, 0, 0)] == 2 and last[(0, 0, 0, 0, 0, 1)] == 1
    assert last[(0,) * 6] == -1
    first = {e: c for c, e in system.eqs[0]}
    assert first[(0, 0, 0, 0, 0, 2)] == 1 and first[(0, 0, 0, 0, 0, 1)] == -1


def test_syntax_of_the_reference_examples():
    """the examples of standard_complex_polynomials_io.ads: ** and ^,
    rational coefficients, complex coefficients in brackets, products of
    bracketed polynomials"""
    system, sym = parse_system(
        "3\n x**2*y + 1/2*z*y**2 - 2*z + y^3 + x - 1E9/-8.E-6* y + 3;\n"
        " (1.01 + 2.8*i)*x**2*y + z**2*x - 3*x + 2*y*z - 3;\n"
        " (x^2*y + z^2*x - 3*x + 2*y*z - 3)*y**2*(y-1+i);\n")
    assert sym == ["x", "y", "z"]
    p = {e: c for c, e in system.eqs[0]}
    assert p[(2, 1, 0)] == 1 and p[(0, 2, 1)] == 0.5 and p[(0, 3, 0)] == 1
    assert p[(0, 1, 0)] == pytest.approx(1e9 / 8e-6)
    q = {e: c for c, e in system.eqs[1]}
    assert q[(2, 1, 0)] == 1.01 + 2.8j
    # the product: evaluate at a point against the expression
    x = np.array([0.3 + 0.1j, -0.7 + 0.2j, 1.1 - 0.4j])
    a = x[0] ** 2 * x[1] + x[2] ** 2 * x[0] - 3 * x[0] + 2 * x[1] * x[2] - 3
    expect = a * x[1] ** 2 * (x[1] - 1 + 1j)
    assert abs(system.eval(x)[2] - expect) < 1e-13


def test_parse_errors():
    with pytest.raises(ParseError):
        parse_system("2\n x + y;\n x - y\n")  # missing ';'
    with pytest.raises(ParseError):
        parse_system("2\n x + y;\n x/y;\n")  # division by an unknown
    with pytest.raises(ParseError):
        parse_system("2\n x + y + z;\n x - y;\n")  # 3 unknowns for 2 equations
    with pytest.raises(ParseError):
        parse_system("1\n x^y;\n")


def test_solution_list_round_trip():
    text = read("katsura5")
    system, symbols = parse_system(text)
    sols = parse_solutions(text, symbols)
    out = write_solutions(sols, symbols)
    back = parse_solutions(out, symbols)
    assert len(back) == len(sols)
    for a, b in zip(sols, back):
        # 15 significant digits are written
        assert np.max(np.abs(a["x"] - b["x"])) <= 1e-14 * max(1.0, np.max(np.abs(a["x"])))
        assert b["m"] == a["m"]
        for key in ("err", "rco", "res"):
            assert b[key] == pytest.approx(a[key], rel=1e-3, abs=1e-300)
    # the layout of the reference's writer
    lines = out.splitlines()
    assert lines[0] == "THE SOLUTIONS :" and lines[1] == "32 6" and set(lines[2]) == {"="}
    assert lines[3] == "solution 1 :"
    assert lines[4].startswith("t :  1.00000000000000E+00   0.00000000000000E+00")
    assert lines[5] == "m : 1" and lines[6] == "the solution for t :"
    assert lines[7].startswith(" x : -7.34679413715196E-02  -1.05421979432305E-81")
    assert lines[13].startswith("== err :  4.235E-16 = rco :  6.845E-02 = res :  5.551E-17 ==")


def test_system_round_trip():
    system, symbols = parse_system(read("utbikker"))
    again, sym2 = parse_system(write_system(system, symbols), symbols)
    assert sym2 == symbols
    for e1, e2 in zip(system.eqs, again.eqs):
        d1, d2 = dict((e, c) for c, e in e1), dict((e, c) for c, e in e2)
        assert d1.keys() == d2.keys()
        for k in d1:
            assert d2[k] == d1[k]
