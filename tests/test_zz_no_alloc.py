"""The allocation-free tiled back substitution
(phc_cmplx_upper_tiled_solver_noalloc: the tile inverses live in the
library's static device scratch) against the allocating entry point
phc_cmplx_upper_tiled_solver on the same inputs: x, the updated b and U with
its diagonal tiles inverted must be bit-identical (the same kernels run on
both).  Two calls in a row, the second smaller, so the second call runs on
scratch that still holds the first call's inverses."""
import ctypes

import numpy as np
import pytest

from oraclelib import rand_cmd

pytestmark = pytest.mark.gpu


def upper_system(K, dim, seed):
    rng = np.random.default_rng(seed)
    U = rand_cmd(rng, K, (dim, dim))
    for p in range(2 * K):
        U[p] = np.triu(U[p])
    for i in range(dim):
        U[0, i, i] += 4.0
    return U, rand_cmd(rng, K, (dim,))


def solve(entry, K, szt, nbt, U, b):
    from phcpack_amd._lib import DeviceArray, check, load
    lib = load()
    dim = szt * nbt
    dU, db = DeviceArray.from_host(U), DeviceArray.from_host(b)
    dx = DeviceArray.zeros((2 * K, dim))
    check(getattr(lib, entry)(K, dim, szt, nbt, dU.ptr, db.ptr, dx.ptr, None))
    check(lib.phc_synchronize())
    return dx.to_host(), dU.to_host(), db.to_host()


@pytest.mark.parametrize("K", (2, 4))
def test_noalloc_tiled_solver_matches_allocating_one(K):
    # the first call fills the whole scratch at K = 4 (dim 1024, tiles of 128)
    for seed, (szt, nbt) in enumerate(((128, 8), (16, 3)) if K == 4 else ((64, 4), (8, 5))):
        dim = szt * nbt
        print(f"noalloc tiled solver K={K} dim={dim} szt={szt} call {seed + 1}", flush=True)
        U, b = upper_system(K, dim, 40 + 10 * K + seed)
        got = solve("phc_cmplx_upper_tiled_solver_noalloc", K, szt, nbt, U, b)
        want = solve("phc_cmplx_upper_tiled_solver", K, szt, nbt, U, b)
        for g, w, name in zip(got, want, ("x", "U", "b")):
            assert np.array_equal(g, w), f"{name} differs (call {seed + 1})"
        # the solution solves the system: U x = b up to the rounding of the
        # leading limbs (relative to the size of the terms)
        x = want[0][0] + 1j * want[0][K]
        Uc = U[0] + 1j * U[K]
        bc = b[0] + 1j * b[K]
        assert np.abs(Uc @ x - bc).max() <= 1e-13 * dim * np.abs(Uc).max() * np.abs(x).max()


def test_noalloc_refuses_oversize():
    from phcpack_amd._lib import load
    lib = load()
    rc = lib.phc_cmplx_upper_tiled_solver_noalloc(4, 2048, 128, 16, ctypes.c_void_p(8),
                                                  ctypes.c_void_p(8), ctypes.c_void_p(8), None)
    assert rc == -2 and b"scratch" in lib.phc_last_error()
