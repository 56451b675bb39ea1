"""The reference-signature layer (include/phcmd_compat.h: GPU_*_product,
GPU_cmplx{2,4}_blocked_houseqr, GPU_cmplx{2,4}_upper_tiled_solver,
GPU_cmplx{2,4}_bals_qhb) against the reference's own host functions, by the
C++ program tests/compat/test_compat.cpp (built by the package build next to
the compiled reference, oracle/_ref).  Products bit for bit; QR, solver and
Q^H b within 64 n eps^K; a planted least-squares solution recovered."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "compat", "test_compat")


def test_reference_signature_layer_matches_reference_host_code():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} was not built (python -m phcpack_amd.build, with the reference "
                    "sources present for oracle/_ref)")
    print("running tests/compat/test_compat", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout
