"""Test helpers for the PHCtoC entry gpumanypaths_dd (libphcmd_phctoc.so):
the in-memory containers of tests/compat/libcontainers.so (PHCpack's C
interface for systems and solutions) loaded into the process first, so that
the entry finds its container functions, then the entry itself."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONTAINERS = os.path.join(ROOT, "tests", "compat", "libcontainers.so")
PHCTOC = os.path.join(ROOT, "phcpack_amd", "libphcmd_phctoc.so")

_libs = {}


def libs():
    """(containers, phctoc) ctypes libraries"""
    if not _libs:
        for path in (CONTAINERS, PHCTOC):
            if not os.path.exists(path):
                raise RuntimeError(f"{path} is not built (python -m phcpack_amd.build; "
                                   "make -C tests/compat)")
        # libphcmd.so first, so that libphcmd_phctoc.so resolves it by soname
        from phcpack_amd import _lib
        _lib.load()
        ctr = ctypes.CDLL(CONTAINERS, mode=ctypes.RTLD_GLOBAL)
        ctr.ctr_number_of_solutions.restype = ctypes.c_int
        phc = ctypes.CDLL(PHCTOC)
        phc.gpumanypaths_dd.restype = ctypes.c_int
        phc.gpumanypaths_dd.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double]
        _libs["ctr"], _libs["phc"] = ctr, phc
    return _libs["ctr"], _libs["phc"]


def _dd_coefficients(cf):
    """[m] complex or [4][m] limb planes -> [m][4] (re hi, re lo, im hi, im lo)"""
    cf = np.asarray(cf)
    if cf.dtype == np.complex128:
        out = np.zeros((len(cf), 4))
        out[:, 0], out[:, 2] = cf.real, cf.imag
        return out
    return np.ascontiguousarray(np.asarray(cf, np.float64).T)


def set_system(which, n, eqs, exps, cf):
    """which 0: target, 1: start; eqs [m], exps [m][n], cf as compile_terms takes"""
    ctr, _ = libs()
    eqs = np.ascontiguousarray(eqs, np.int32)
    exps = np.ascontiguousarray(exps, np.int32).reshape(len(eqs), n)
    c4 = np.ascontiguousarray(_dd_coefficients(cf))
    ctr.ctr_set_system(which, n, n, len(eqs), eqs.ctypes.data_as(ctypes.c_void_p),
                       exps.ctypes.data_as(ctypes.c_void_p), c4.ctypes.data_as(ctypes.c_void_p))


def set_solutions(planes):
    """start solutions from [4][npaths][n] double-double limb planes"""
    ctr, _ = libs()
    _, npaths, n = planes.shape
    sols = np.zeros((npaths, 4 * n + 10))
    for q in range(4):
        sols[:, 4 + q:4 + 4 * n:4] = planes[q]
    sols = np.ascontiguousarray(sols)
    ctr.ctr_set_solutions(n, npaths, sols.ctypes.data_as(ctypes.c_void_p))


def get_solutions(n):
    """(planes [4][len][n], t [len], err [len], mult [len]) in the container"""
    ctr, _ = libs()
    m = ctr.ctr_number_of_solutions()
    sols = np.zeros((m, 4 * n + 10))
    mult = np.zeros(m, np.int32)
    ctr.ctr_get_solutions(n, sols.ctypes.data_as(ctypes.c_void_p), mult.ctypes.data_as(ctypes.c_void_p))
    planes = np.stack([sols[:, 4 + q:4 + 4 * n:4] for q in range(4)])
    return planes, sols[:, 0], sols[:, 4 * n + 4], mult


def gpumanypaths_dd(mode, verbose, gamma):
    _, phc = libs()
    return phc.gpumanypaths_dd(mode, verbose, gamma.real, gamma.imag)
