"""The PHCtoC entry gpumanypaths_dd (libphcmd_phctoc.so) on a CPU-only
host: it loads against PHCpack's container interface (here the in-memory
containers of tests/compat/libcontainers.so), exports the reference's
symbol, and refuses what it cannot do before any device work: mode 1 (CPU
tracking) and a system that is not square."""
import numpy as np

from phctoclib import gpumanypaths_dd, libs, set_solutions, set_system


def test_library_loads_and_exports_the_entry():
    _, phc = libs()
    assert hasattr(phc, "gpumanypaths_dd")


def test_mode_1_is_refused():
    assert gpumanypaths_dd(1, 0, complex(0.6, 0.8)) == 1


def test_non_square_system_is_refused():
    n = 3
    # a target of 2 equations in 3 unknowns
    eqs = np.array([0, 1])
    exps = np.array([[1, 0, 0], [0, 1, 1]])
    ctr, _ = libs()
    import ctypes
    c4 = np.ascontiguousarray(np.ones((2, 4)))
    ctr.ctr_set_system(0, n, 2, 2, eqs.astype(np.int32).ctypes.data_as(ctypes.c_void_p),
                       np.ascontiguousarray(exps, np.int32).ctypes.data_as(ctypes.c_void_p),
                       c4.ctypes.data_as(ctypes.c_void_p))
    set_system(1, n, np.arange(n), np.eye(n, dtype=np.int32), np.ones(n, complex))
    set_solutions(np.zeros((4, 1, n)))
    assert gpumanypaths_dd(2, 0, complex(0.6, 0.8)) == 2
