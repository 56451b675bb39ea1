"""Build recipe for the HIP library libphcmd.so (gfx950) and the CPU oracle.

``python -m phcpack_amd.build`` compiles every ``csrc/*.hip`` source with
hipcc for gfx950 into ``phcpack_amd/libphcmd.so`` (in-tree, so it travels with
the repository), links the PHCtoC entry (``phctoc/``) into
``phcpack_amd/libphcmd_phctoc.so`` on top of it, and runs ``make -C oracle``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libphcmd.so")
PHCTOC_LIB = os.path.join(HERE, "libphcmd_phctoc.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PHC_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off keeps the two roundings of every plain product/sum of the
# multiple-double algorithms (fused multiply-adds are written explicitly).
FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def sources():
    """HIP kernels and C-ABI (*.hip) and host-only C++ (*.cpp: the layer with
    the reference's C++ signatures)"""
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build_hip(jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), 8)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_phctoc():
    """libphcmd_phctoc.so: gpumanypaths_dd (include/phcmd_phctoc.h) over
    libphcmd.so; PHCpack's container functions stay undefined, to be
    resolved from the program that loads it"""
    src = os.path.join(HERE, "phctoc", "gpumanypaths_dd.cpp")
    tmp = PHCTOC_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", src, "-o", tmp,
           "-L" + HERE, "-lphcmd", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath-link," + os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"PHCtoC build failed:\n{r.stderr}")
    os.replace(tmp, PHCTOC_LIB)
    return PHCTOC_LIB


def build_oracle():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")


def build_compat_test():
    """the test programs of the reference-signature layer (tests/compat):
    test_compat checks the GPU_* entry points against the CPU oracle,
    jobs_dump prints the job schedules; jobs_dump_ref (the same program
    against the reference's job classes) is built where the reference
    sources and oracle/_ref are present"""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "compat")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compat test build failed:\n{r.stdout}\n{r.stderr}")


def build_all():
    build_hip()
    build_phctoc()
    build_oracle()
    build_compat_test()


if __name__ == "__main__":
    build_all()
    print("built", LIB, file=sys.stderr)
