"""The HIP library loads on a CPU-only host and exports every symbol that
include/phcmd.h declares (no compute calls without a GPU)."""
import os
import re

from phcpack_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms |= set(re.findall(r"\b(phc_\w+)\s*\(", text))
    return syms


def test_library_exports_declared_symbols():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) > 10
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert declared_symbols() <= set(_lib.SIGNATURES)


def test_version_and_error_string():
    lib = _lib.load()
    assert lib.phc_version() >= 100
    assert isinstance(lib.phc_last_error(), bytes)


def test_programs_validate_on_cpu():
    """Compiled homotopies pass the library's index validation; a corrupted
    one is rejected with an error (no device work involved)."""
    import ctypes

    import numpy as np

    from phcpack_amd.homotopy import compile_homotopy, dense_system, total_degree_homotopy
    from phcpack_amd.tracker import program_struct
    lib = _lib.load()
    rng = np.random.default_rng(0)
    for n, d, k in [(3, 2, 1), (4, 3, 1), (3, 2, 2), (12, 2, 1)]:
        hom, _ = total_degree_homotopy(dense_system(n, d, rng), rng, k=k)
        for deg in (0, 1, 8):
            ps = program_struct(compile_homotopy(hom, deg))
            assert lib.phc_program_validate(ctypes.byref(ps)) == 0, lib.phc_last_error()
    prog = compile_homotopy(hom, 8)
    prog.term_pair = prog.term_pair.copy()
    prog.term_pair[3] = 10 ** 6
    ps = program_struct(prog)
    assert lib.phc_program_validate(ctypes.byref(ps)) == -2
    assert b"bad term" in lib.phc_last_error()


def test_reference_signature_layer_exported():
    """every GPU_* entry point of include/phcmd_compat.h is exported with
    C++ linkage (demangled name and parameter list as declared)"""
    import subprocess
    text = open(os.path.join(ROOT, "include", "phcmd_compat.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = re.findall(r"void\s+(GPU_\w+)\s*\(([^;]*)\);", text)
    assert len(decls) >= 12
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {}
    for line in out.splitlines():
        m = re.search(r"\b(GPU_\w+)\((.*)\)$", line)
        if m:
            exported[m.group(1)] = m.group(2)
    for name, params in decls:
        assert name in exported, name
        # parameter types in order (names and default values dropped)
        types = []
        for p in params.split(","):
            p = p.split("=")[0].strip()
            p = re.sub(r"\s*\b\w+$", "", p) if re.search(r"[\w\*]\s+\**\w+$", p) else p
            types.append(re.sub(r"\s+", "", p.replace("long long int", "long long")))
        got = [re.sub(r"\s+", "", t) for t in exported[name].split(",")]
        assert got == types, (name, got, types)


def test_invalid_arguments_are_refused_before_device_work():
    """entry points check their arguments first and return -2 with a
    message (no GPU is touched here)"""
    import ctypes
    lib = _lib.load()
    null = None

    def refused(rc, words):
        msg = lib.phc_last_error().decode()
        return rc == -2 and all(w in msg for w in words)

    assert refused(lib.phc_series_product(3, 0, 4, 1, null, null, null, null), ["K must be"])
    assert refused(lib.phc_series_product(2, 0, -1, 1, null, null, null, null), ["negative"])
    assert refused(lib.phc_md_batch(9, 2, 4, null, null, null, null), ["op"])
    assert refused(lib.phc_md_batch(0, 3, 4, null, null, null, null), ["K must be"])
    h = ctypes.c_void_p()
    assert refused(lib.phc_baqr_create(ctypes.byref(h), 3, 64, 64, 16), ["K must be"])
    assert refused(lib.phc_baqr_create(ctypes.byref(h), 4, 64, 64, 0), ["szt"])
    assert refused(lib.phc_baqr_create(ctypes.byref(h), 4, 512, 512, 256), ["szt"])
    assert refused(lib.phc_baqr_create(ctypes.byref(h), 4, 64, 60, 16), ["szt"])
    assert refused(lib.phc_baqr_create(ctypes.byref(h), 4, 32, 64, 16), ["nrows"])
    assert refused(lib.phc_cmplx_upper_tiled_solver(4, 65, 16, 4, null, null, null, null),
                   ["dim must equal"])
    assert refused(lib.phc_cmplx_gemm(3, 0, 0, 4, 4, 4, null, 0, 1, 1, null, 0, 1, 1, null, 0,
                                      1, 1, 0, null), ["K must be"])
    assert refused(lib.phc_refine(null, 2, 1, 1.0, null, 2, null, 1, 1e-20, null, null, null,
                                  null, 8, null), ["null tracker"])
    nvr = (ctypes.c_int * 1)(1)
    idx = (ctypes.c_int * 1)(0)
    assert refused(lib.phc_poly_create(ctypes.byref(h), 2, 0, 1, 1, 300, nvr, idx, None, None),
                   ["degree"])
    mono = (ctypes.c_int * 2)(0, 1)
    assert refused(lib.phc_rowsys_create(ctypes.byref(h), 65, mono, nvr, mono, idx), ["dim"])
    two = (ctypes.c_int * 2)(1, 1)
    assert refused(lib.phc_colsys_create(ctypes.byref(h), 1, 2, two, mono, two), ["exponents"])
    zero = (ctypes.c_int * 1)(0)
    assert refused(lib.phc_colsys_create(ctypes.byref(h), 1, 1, zero, idx, None), ["count"])
    rep = (ctypes.c_int * 3)(0, 0, 1)
    two1 = (ctypes.c_int * 2)(2, 1)
    assert refused(lib.phc_colsys_create(ctypes.byref(h), 2, 1, two1, rep, None), ["repeated"])
    assert refused(lib.phc_column_newton(null, 2, 4, 1, 1, null, 0, null, 0, null, null, null,
                                         null, null, null, null, null, null, null),
                   ["not a column system"])
