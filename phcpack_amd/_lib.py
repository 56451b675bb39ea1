"""ctypes binding of libphcmd.so (the HIP library) with declared signatures.

The library must have been built (``python -m phcpack_amd.build``); there is no
CPU fallback: every compute entry point runs on the GPU or raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphcmd.so")

c_int = ctypes.c_int
c_double = ctypes.c_double
c_size_t = ctypes.c_size_t
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
P_int = ctypes.POINTER(ctypes.c_int)
P_double = ctypes.POINTER(ctypes.c_double)
P_void_p = ctypes.POINTER(ctypes.c_void_p)
P_float = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes); every symbol of include/phcmd.h
SIGNATURES = {
    "phc_last_error": (c_char_p, []),
    "phc_version": (c_int, []),
    "phc_device_count": (c_int, [P_int]),
    "phc_malloc": (c_int, [P_void_p, c_size_t]),
    "phc_free": (c_int, [c_void_p]),
    "phc_memcpy_htod": (c_int, [c_void_p, c_void_p, c_size_t]),
    "phc_memcpy_dtoh": (c_int, [c_void_p, c_void_p, c_size_t]),
    "phc_memset": (c_int, [c_void_p, c_int, c_size_t]),
    "phc_synchronize": (c_int, []),
    "phc_stream_synchronize": (c_int, [c_void_p]),
    "phc_event_create": (c_int, [P_void_p]),
    "phc_event_destroy": (c_int, [c_void_p]),
    "phc_event_record": (c_int, [c_void_p, c_void_p]),
    "phc_event_elapsed_ms": (c_int, [c_void_p, c_void_p, P_float]),
    "phc_device_copy": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "phc_md_batch": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "phc_series_product": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                   c_void_p]),
    "phc_poly_create": (c_int, [P_void_p, c_int, c_int, c_int, c_int, c_int, P_int, P_int,
                                P_double, P_double]),
    "phc_poly_destroy": (c_int, [c_void_p]),
    "phc_poly_evaldiff": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
}

class PhcProgram(ctypes.Structure):
    _fields_ = [
        ("n", c_int), ("deg", c_int), ("ncomp", c_int), ("wdeg", c_int), ("G", c_int),
        ("nprod", c_int), ("nlev", c_int), ("nslots", c_int), ("nfactor", c_int),
        ("nodes", P_int), ("lev_ptr", P_int),
        ("npair", c_int), ("pairs", P_int),
        ("nterm", c_int), ("seg_ptr", P_int), ("term_pair", P_int), ("term_coef", P_double),
        ("njpair", c_int), ("jpairs", P_int),
        ("njterm", c_int), ("jent_ptr", P_int), ("jterms", P_int),
        ("weights", P_double),
        ("hsig", P_double),
        ("nhterm", c_int), ("hent_ptr", P_int), ("hterms", P_int),
        ("nhpair", c_int), ("hpairs", P_int),
    ]


class PhcTrackParams(ctypes.Structure):
    _fields_ = [
        ("maxsteps", c_int), ("corsteps", c_int), ("endsteps", c_int), ("fabry_offset", c_int),
        ("maxstep", c_double), ("minstep", c_double), ("pbeta", c_double), ("alpha", c_double),
        ("tolcor", c_double), ("tolend", c_double), ("divergence", c_double),
        ("pade_numdeg", c_int), ("pade_dendeg", c_int),
        ("cbeta", c_double),
    ]



CTL_MAXCAND = 16


class PhcController(ctypes.Structure):
    _fields_ = [
        ("mode", c_int), ("ncand", c_int), ("mult", c_double * CTL_MAXCAND),
        ("weights", c_void_p), ("mu", c_void_p), ("sd", c_void_p),
        ("logit_threshold", ctypes.c_float), ("seed", ctypes.c_uint), ("log_capacity", c_int),
        ("log_x", c_void_p), ("log_c", c_void_p), ("log_y", c_void_p), ("log_n", c_void_p),
        ("pole_cap", c_double),
        ("deg0", c_int), ("degree", c_int * CTL_MAXCAND), ("retry_below", c_double),
    ]


SIGNATURES.update({
    "phc_tracker_create": (c_int, [ctypes.POINTER(c_void_p), ctypes.POINTER(PhcProgram)]),
    "phc_tracker_destroy": (c_int, [c_void_p]),
    "phc_program_validate": (c_int, [ctypes.POINTER(PhcProgram)]),
    "phc_program_compile": (c_int, [P_void_p, c_int, c_int, c_int, c_int, P_double, P_int, P_int,
                                    P_int, c_int, P_double]),
    "phc_compiled_program": (c_int, [c_void_p, ctypes.POINTER(PhcProgram)]),
    "phc_compiled_destroy": (c_int, [c_void_p]),
    "phc_tracker_lds_bytes": (c_int, [c_void_p, c_int, ctypes.POINTER(ctypes.c_size_t),
                                      ctypes.POINTER(ctypes.c_size_t)]),
    "phc_track": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                          c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(PhcTrackParams),
                          ctypes.POINTER(PhcController), c_int, c_void_p]),
    "phc_track_from": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               ctypes.POINTER(PhcTrackParams), ctypes.POINTER(PhcController),
                               c_int, c_void_p]),
    "phc_mlp_size": (c_int, [c_int, c_int, c_int, P_int]),
    "phc_mlp_forward": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "phc_mlp_grad": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p]),
    "phc_mlp_train": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                              c_int, c_int, ctypes.c_float, ctypes.c_uint,
                              ctypes.POINTER(ctypes.c_float), c_void_p]),
    "phc_endgame": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p,
                            c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "phc_refine": (c_int, [c_void_p, c_int, c_int, c_double, c_void_p, c_int, c_void_p, c_int,
                           c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "phc_houseqr_batch": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "phc_qrbs_solve_batch": (c_int, [c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "phc_linear_residue_batch": (c_int, [c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_void_p]),
    "phc_rowsys_create": (c_int, [ctypes.POINTER(c_void_p), c_int, P_int, P_int, P_int, P_int]),
    "phc_rowsys_create_exp": (c_int, [ctypes.POINTER(c_void_p), c_int, P_int, P_int, P_int, P_int,
                                      P_int]),
    "phc_rowsys_destroy": (c_int, [c_void_p]),
    "phc_colsys_create": (c_int, [ctypes.POINTER(c_void_p), c_int, c_int, P_int, P_int, P_int]),
    "phc_column_newton": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int,
                                  c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "phc_rowsys_lds_bytes": (c_int, [c_void_p, c_int, c_int, ctypes.POINTER(ctypes.c_size_t)]),
    "phc_row_newton": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p,
                               c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_void_p]),
    "phc_cmplx_row_newton": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int,
                                     c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "phc_cmplx_houseqr_batch": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    "phc_cmplx_qrbs_solve_batch": (c_int, [c_int, c_int, c_int, c_int, c_int, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    "phc_cmplx_linear_residue_batch": (c_int, [c_int, c_int, c_int, c_int, c_int, c_void_p,
                                               c_void_p, c_void_p, c_void_p, c_void_p,
                                               c_void_p]),
    "phc_baqr_create": (c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_int]),
    "phc_baqr_destroy": (c_int, [c_void_p]),
    "phc_baqr_set_tuning": (c_int, [c_void_p, c_int, c_int]),
    "phc_baqr_get_tuning": (c_int, [c_void_p, P_int, P_int]),
    "phc_cmplx_blocked_houseqr": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "phc_real_blocked_houseqr": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "phc_real_qtb": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "phc_cmplx_qhb": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "phc_cmplx_upper_tiled_solver": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                             c_void_p, c_void_p]),
    "phc_cmplx_upper_tiled_solver_noalloc": (c_int, [c_int, c_int, c_int, c_int, c_void_p,
                                                     c_void_p, c_void_p, c_void_p]),
    "phc_cmplx_gemm": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int,
                               c_void_p, c_size_t, ctypes.c_long, ctypes.c_long,
                               c_void_p, c_size_t, ctypes.c_long, ctypes.c_long,
                               c_void_p, c_size_t, ctypes.c_long, ctypes.c_long, c_int, c_void_p]),
})



_lib = None


class PhcError(RuntimeError):
    pass


def load():
    """Load libphcmd.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PhcError(f"{LIB_PATH} is missing: build it with `python -m phcpack_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != 0:
        raise PhcError(load().phc_last_error().decode())
    return status


def ptr_i(a):
    return a.ctypes.data_as(P_int)


def ptr_d(a):
    return a.ctypes.data_as(P_double)


class DeviceArray:
    """A float64/int32 device buffer owned by the HIP library."""

    def __init__(self, shape, dtype=np.float64):
        lib = load()
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = c_void_p()
        check(lib.phc_malloc(ctypes.byref(p), max(self.nbytes, 8)))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(load().phc_memcpy_htod(d.ptr, a.ctypes.data_as(c_void_p), d.nbytes))
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float64):
        d = cls(shape, dtype)
        check(load().phc_memset(d.ptr, 0, d.nbytes))
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(load().phc_memcpy_htod(self.ptr, a.ctypes.data_as(c_void_p), self.nbytes))

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        check(load().phc_memcpy_dtoh(out.ctypes.data_as(c_void_p), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            load().phc_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
