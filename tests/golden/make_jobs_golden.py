"""Writes tests/golden/jobs_dump.txt.gz: the convolution and addition job
schedules that the reference's own ConvolutionJobs and AdditionJobs classes
(Polynomials/convolution_jobs.cpp, addition_jobs.cpp, compiled into
oracle/_ref/libphcref.so by oracle/Makefile) produce for the seeded supports
of tests/compat/jobs_dump.cpp.  Needs the reference checkout; run after
`make -C oracle` and `make -C tests/compat`."""
import gzip
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "tests", "compat", "jobs_dump_ref")


def main():
    if not os.path.exists(BIN):
        sys.exit(f"{BIN} is missing: build it with make -C tests/compat (reference needed)")
    out = subprocess.run([BIN], capture_output=True, text=True, check=True).stdout
    with gzip.open(os.path.join(HERE, "jobs_dump.txt.gz"), "wt", compresslevel=9) as fh:
        fh.write(out)
    print(f"wrote {len(out.splitlines())} lines")


if __name__ == "__main__":
    main()
