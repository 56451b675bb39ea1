"""ConvolutionJobs / AdditionJobs (phcpack_amd/csrc/jobs.cpp, the
reference's job planners with the reference's public API) against the
reference's own classes: tests/compat/jobs_dump prints the schedules of 45
seeded supports (monomials of 1..20 variables, up to 64 monomials) through
the public API, job by job and with operator<<; the output must equal the
reference's, kept in tests/golden/jobs_dump.txt.gz (written by
tests/golden/make_jobs_golden.py from the reference classes compiled into
oracle/_ref) and, where the reference build is present, produced afresh by
the same program linked against it.  Host code only (no GPU)."""
import gzip
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MINE = os.path.join(HERE, "compat", "jobs_dump")
REF = os.path.join(HERE, "compat", "jobs_dump_ref")
GOLDEN = os.path.join(HERE, "golden", "jobs_dump.txt.gz")


def run(binary):
    if not os.path.exists(binary):
        pytest.fail(f"{binary} was not built (python -m phcpack_amd.build)")
    return subprocess.run([binary], capture_output=True, text=True, check=True,
                          timeout=60).stdout


def test_schedules_match_reference_fixture():
    with gzip.open(GOLDEN, "rt") as fh:
        want = fh.read()
    got = run(MINE)
    assert got.count("support dim") == 45
    if got != want:
        a, b = got.splitlines(), want.splitlines()
        k = next(i for i in range(min(len(a), len(b))) if a[i] != b[i]) \
            if any(x != y for x, y in zip(a, b)) else min(len(a), len(b))
        pytest.fail(f"first difference at line {k + 1}: {a[k:k + 1]} vs {b[k:k + 1]}")


def test_schedules_match_compiled_reference():
    if not os.path.exists(REF):
        pytest.skip("the reference's job classes are not built here (oracle/_ref)")
    assert run(MINE) == run(REF)


def test_schedule_shapes():
    """counts that the schedule construction fixes: a monomial of n >= 3
    variables takes 3n - 3 convolutions (n forward, n - 1 backward, n - 2
    cross), of 2 variables 3, of 1 variable 1; the value tree of nbr
    monomials takes nbr additions (one of them adds the constant)"""
    out = run(MINE)
    blocks = out.split("support dim ")[1:]
    for blk in blocks:
        head = blk.splitlines()[0]
        mons = head.split(":", 1)[1].split()
        sizes = [len(m.strip("()").split(",")) for m in mons]
        conv = sum(1 if n == 1 else 3 if n == 2 else 3 * n - 3 for n in sizes)
        line = next(l for l in blk.splitlines() if l.startswith("conv "))
        assert int(line.split()[4]) == conv
        adds = [l for l in blk.splitlines() if "+=" in l and "|" in l]
        assert sum(1 for l in adds if l.split("|")[1].strip().startswith("f[")
                   and l.split()[0] == "1" and l.split()[1] == "1") >= len(sizes)
