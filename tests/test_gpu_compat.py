"""The reference-signature layer (include/phcmd_compat.h: GPU_*_product,
GPU_*_poly_evaldiff with ConvolutionJobs/AdditionJobs schedules,
GPU_cmplx{,2,4}vectorized_poly_evaldiff with the complex schedules,
GPU_cmplx{2,4,8}_blocked_houseqr, GPU_dbl{2,4,8}_blocked_houseqr, GPU_cmplx{2,4,8}_upper_tiled_solver,
GPU_dbl{2,4}_upper_tiled_solver, GPU_cmplx{2,4}_bals_qhb,
GPU_dbl{,2,4}_bals_solve, GPU_dbl{,2,4}_linear_residue) against the CPU
oracle, by the C++ program tests/compat/test_compat.cpp (built by the
package build; it links libphcmd.so and oracle/liboracle.so only).
Products, evaluation, Toeplitz solve and residue bit for bit; QR, solvers
and Q^H b within 64 n eps^K; a planted least-squares solution recovered."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "compat", "test_compat")


def test_reference_signature_layer_matches_oracle():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} was not built (python -m phcpack_amd.build)")
    print("running tests/compat/test_compat", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout
    for name in ("GPU_dbl8_poly_evaldiff", "GPU_cmplx4_poly_evaldiff", "GPU_cmplx8_product",
                 "GPU_cmplxvectorized_poly_evaldiff", "GPU_cmplx2vectorized_poly_evaldiff",
                 "GPU_cmplx4vectorized_poly_evaldiff",
                 "GPU_dbl4_bals_solve", "GPU_cmplx4_blocked_houseqr",
                 "GPU_cmplx8_blocked_houseqr", "GPU_cmplx8_upper_tiled_solver",
                 "GPU_dbl2_blocked_houseqr", "GPU_dbl8_blocked_houseqr"):
        assert name in r.stdout, f"{name} was not checked"
