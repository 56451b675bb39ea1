"""The column Newton method on the GPU (phc_colsys_create, phc_column_newton)
against the oracle's restatement of the reference's host code
(onw_column_newton, itself bit-exact with the compiled reference):
every step's series, the tail indices, the skip counters and the last Q
and R, bit for bit, at K = 1, 2, 4, batched with shared and per-system
coefficients and right-hand sides.

One documented exception: the device's two_prod takes a fused multiply-add,
the reference's Dekker's split; the two differ only where a product's error
term is subnormal (|a b| < 2^-916).  Components that one Newton step solves
exactly leave limbs near 1e-140, whose quad-double products reach that range,
so at K = 4 the trailing limbs may differ there, and a later rounding may
then flip a last bit: such a run must still agree in every step count, tail
index and skip counter, and in value to 2^-200 relative (4096 units of the
quad-double precision 2^-212)."""
import numpy as np
import pytest

from oraclelib import ColumnSystem, oracle_column_newton, planted_column_system
from test_oracle_columns import golden_case

pytestmark = pytest.mark.gpu

KEYS = ("wdeg", "tail", "up", "bs")


def device_system(csys):
    from phcpack_amd.newton import ColumnSystem as DevColumns
    return DevColumns(csys.dim, csys.nbrcol, csys.nvr, csys.idx, csys.exps)


def check_against_oracle(K, csys, starts, nbsteps, cff=None, mb=None):
    """starts [K][nsys][dim][D]; cff/mb per system ([K][nsys][...]) or None
    for the system's own (shared)"""
    dev = device_system(csys)
    nsys = starts.shape[1]
    mbs = csys.mb[:, None] if mb is None else mb
    cffs = None if csys.cff is None else (csys.cff[:, None] if cff is None else cff)
    got = dev.column_newton(K, csys.deg, nbsteps, mbs, starts, cff=cffs, history=True)
    for s in range(nsys):
        c = ColumnSystem(csys.dim, csys.deg, csys.nbrcol, csys.nvr, csys.idx,
                         mbs[:, s if mbs.shape[1] > 1 else 0],
                         cff=None if cffs is None else cffs[:, s if cffs.shape[1] > 1 else 0],
                         exps=csys.exps)
        want = oracle_column_newton(K, c, np.ascontiguousarray(starts[:, s]), nbsteps)
        n = want["steps"]
        assert got["steps"][s] == n, s
        for k in KEYS:
            assert np.array_equal(got[k][s, :n], want[k]), (s, k)
        pairs = ((got["history"][:n, :, s], want["history"]), (got["x"][:, s], want["x"]),
                 (got["Q"][:, s], want["Q"]), (got["R"][:, s], want["R"]))
        for a, b in pairs:
            if K == 4 and not np.array_equal(a, b):
                # subnormal product errors (fused vs split two_prod), see above
                diff = np.abs((a - b).sum(axis=-3 if a.ndim == 4 else 0))
                assert diff.max() <= 2.0 ** -200 * max(1.0, np.abs(b[..., 0, :, :] if a.ndim == 4
                                                                  else b[0]).max()), s
            else:
                assert np.array_equal(a, b), s
    return got


@pytest.mark.parametrize("name", ["mono", "two", "cyclic"])
def test_reference_fixture_on_the_device(name):
    csys, g = golden_case(name)
    print("column fixture", name, flush=True)
    dev = device_system(csys)
    got = dev.column_newton(2, csys.deg, 12, csys.mb[:, None], g("start")[:, None],
                            cff=None if csys.cff is None else csys.cff[:, None], history=True)
    n = len(g("tail"))
    assert got["steps"][0] == n
    for k in KEYS:
        assert np.array_equal(got[k][0, :n], g(k)), k
    assert np.array_equal(got["history"][:n, :, 0], g("history"))
    assert np.array_equal(got["Q"][:, 0], g("Q")) and np.array_equal(got["R"][:, 0], g("R"))


@pytest.mark.parametrize("K", [1, 2, 4])
@pytest.mark.parametrize("dim,nbrcol", [(5, 1), (9, 1), (6, 3), (8, 8)])
def test_planted_batches_match_oracle(K, dim, nbrcol):
    rng = np.random.default_rng(31 * dim + nbrcol + K)
    print("column batch", K, dim, nbrcol, flush=True)
    csys, root, start = planted_column_system(rng, K, dim, 13, nbrcol)
    nsys = 5
    starts = np.repeat(start[:, None], nsys, axis=1)
    starts[0, :, :, 0] += rng.uniform(-1e-6, 1e-6, (nsys, dim))
    check_against_oracle(K, csys, starts, 10)


@pytest.mark.parametrize("K", [2, 4])
def test_per_system_coefficients_and_right_hand_sides(K):
    rng = np.random.default_rng(77 + K)
    nsys, dim, nbrcol, deg = 4, 6, 4, 9
    systems = [planted_column_system(rng, K, dim, deg, nbrcol) for _ in range(nsys)]
    csys = systems[0][0]
    cff = np.stack([s[0].cff for s in systems], axis=1)
    mb = np.stack([s[0].mb for s in systems], axis=1)
    starts = np.stack([s[2] for s in systems], axis=1)
    check_against_oracle(K, csys, starts, 8, cff=cff, mb=mb)


def test_monomial_system_ignores_coefficients():
    """a monomial system's coefficients are the unit series whatever the
    caller passes (the reference resets them each step)"""
    rng = np.random.default_rng(5)
    csys, root, start = planted_column_system(rng, 2, 4, 7, 1)
    dev = device_system(csys)
    a = dev.column_newton(2, 7, 6, csys.mb[:, None], start[:, None])
    b = dev.column_newton(2, 7, 6, csys.mb[:, None], start[:, None],
                          cff=rng.uniform(size=(2, 1, 4, 8)))
    assert np.array_equal(a["x"], b["x"])


def test_column_system_is_refused_by_the_row_entry_point():
    from phcpack_amd import _lib
    rng = np.random.default_rng(6)
    csys, root, start = planted_column_system(rng, 2, 4, 7, 3)
    dev = device_system(csys)
    lib = _lib.load()
    rc = lib.phc_row_newton(dev.handle, 2, 7, 4, 1, None, 0, None, 0, None, None, None, None,
                            None, None, None, None, None, None)
    assert rc != 0 and b"column" in lib.phc_last_error()
