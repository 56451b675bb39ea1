"""End to end on the reference's own example systems (tests/golden/demo,
from src/Demo): parse the system, track the total-degree homotopy on the
GPU (FP64) with double-double Newton at the endpoints, and find every
solution the reference stored (to ~15 digits, all the file holds)."""
import os

import numpy as np
import pytest

from phcpack_amd.formats import parse_solutions, parse_system, solve_total_degree, write_solutions

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(__file__), "golden", "demo")


def load(name):
    with open(os.path.join(DEMO, name)) as f:
        text = f.read()
    system, symbols = parse_system(text)
    return system, symbols, parse_solutions(text, symbols)


def scaled_residual(system, x):
    s = max(1.0, float(np.max(np.abs(x))))
    out = 0.0
    for eq in system.eqs:
        d = max(sum(e) for _, e in eq)
        v = abs(sum(c * np.prod(np.power(x, e)) for c, e in eq))
        out = max(out, v / (sum(abs(c) for c, _ in eq) * s ** d))
    return out


def matches(X, x, tol=1e-9):
    """indices of the rows of X within tol (relative) of x"""
    if len(X) == 0:
        return np.zeros(0, int)
    d = np.max(np.abs(X - x), axis=1)
    return np.nonzero(d <= tol * max(1.0, float(np.max(np.abs(x)))))[0]


@pytest.mark.parametrize("name", ["mickey", "redeco5", "katsura5", "utbikker"])
def test_every_stored_solution_found(name):
    """systems whose solution count equals the total degree: every path
    ends at a regular solution and the solutions are the stored ones, one to
    one"""
    system, symbols, stored = load(name)
    print(f"{name}: n = {system.n}, {len(stored)} stored solutions", flush=True)
    sols, out = solve_total_degree(system, seed=1)
    assert len(sols) == len(stored)
    assert all(s["track_status"] == 1 and s["refine_status"] == 1 for s in sols)
    X = np.array([s["x"] for s in sols])
    for s in sols:
        assert scaled_residual(system, s["x"]) < 1e-14
    used = set()
    for s in stored:
        m = matches(X, s["x"])
        assert len(m) == 1, (name, s["x"])
        used.add(int(m[0]))
    assert len(used) == len(stored)


def test_cyclic5_finite_solutions():
    """cyclic 5-roots: 120 paths, 70 isolated solutions (the mixed volume),
    the other 50 paths diverge; the 7 generating solutions the file stores
    are among the 70"""
    system, symbols, stored = load("cyclic5")
    sols, out = solve_total_degree(system, seed=1)
    assert len(sols) == 120
    good = [s for s in sols if s["track_status"] == 1 and s["refine_status"] == 1]
    assert len(good) == 70
    X = np.array([s["x"] for s in good])
    for i in range(len(X)):
        assert len(matches(X, X[i])) == 1  # distinct
        assert scaled_residual(system, X[i]) < 1e-14
    for s in stored:
        assert len(matches(X, s["x"])) == 1
    # the solution list in the reference's format reads back
    text = write_solutions(good, symbols)
    back = parse_solutions(text, symbols)
    assert len(back) == 70
    assert np.max(np.abs(np.array([b["x"] for b in back]) - X)) < 1e-13
