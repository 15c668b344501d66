"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.  The shipped library never uses it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class Options(ctypes.Structure):
    _fields_ = [
        ("solver", ctypes.c_int),
        ("max_iters", ctypes.c_int),
        ("rel_tol", ctypes.c_double),
        ("abs_tol", ctypes.c_double),
        ("fact_method", ctypes.c_int),
        ("precond_type", ctypes.c_int),
        ("safe_fact", ctypes.c_int),
        ("prec_double", ctypes.c_int),
        ("refine", ctypes.c_int),
        ("refine_tol", ctypes.c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("num_iterations", ctypes.c_float),
        ("rhs_norm", ctypes.c_double),
        ("residual_norm", ctypes.c_double),
        ("rel_residual_norm", ctypes.c_double),
        ("code", ctypes.c_int),
        ("num_partitions", ctypes.c_int),
        ("bandwidth", ctypes.c_int),
        ("time_setup", ctypes.c_double),
        ("time_solve", ctypes.c_double),
        ("time_bandLU", ctypes.c_double),
        ("time_bandUL", ctypes.c_double),
        ("time_assembly", ctypes.c_double),
        ("time_fullLU", ctypes.c_double),
        ("num_applies", ctypes.c_int),
        ("num_spmv", ctypes.c_int),
        ("num_refine", ctypes.c_int),
    ]


_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_gen_banded.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _dp]
        L.orc_gen_vector.argtypes = [ctypes.c_int64, ctypes.c_uint64, _dp]
        L.orc_band_spmv.argtypes = [ctypes.c_int64, ctypes.c_int, _dp, _dp, _dp]
        L.orc_band_solve_pp.argtypes = [ctypes.c_int64, ctypes.c_int, _dp, _dp, _dp]
        L.orc_band_solve_pp.restype = ctypes.c_int
        L.orc_create.argtypes = [ctypes.c_int, ctypes.POINTER(Options)]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_setup_banded.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, _dp]
        L.orc_setup_banded.restype = ctypes.c_int
        L.orc_solve.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_solve.restype = ctypes.c_int
        L.orc_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
        L.orc_monitor_message.argtypes = [ctypes.c_void_p]
        L.orc_monitor_message.restype = ctypes.c_char_p
        L.orc_precond_apply.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_precond_apply.restype = ctypes.c_int
        L.orc_num_partitions.argtypes = [ctypes.c_void_p]
        L.orc_num_partitions.restype = ctypes.c_int
        L.orc_get_factors.argtypes = [ctypes.c_void_p, _dp]
        L.orc_get_reduced.argtypes = [ctypes.c_void_p, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def gen_banded(n, k, dominance, seed):
    band = np.empty((n, 2 * k + 1), dtype=np.float64)
    lib().orc_gen_banded(n, k, dominance, seed, _p(band))
    return band


def gen_vector(n, seed):
    x = np.empty(n, dtype=np.float64)
    lib().orc_gen_vector(n, seed, _p(x))
    return x


def band_spmv(band, x):
    n, w = band.shape
    y = np.empty(n, dtype=np.float64)
    lib().orc_band_spmv(n, (w - 1) // 2, _p(band), _p(np.ascontiguousarray(x, dtype=np.float64)), _p(y))
    return y


def band_solve_pp(band, b):
    n, w = band.shape
    x = np.empty(n, dtype=np.float64)
    rc = lib().orc_band_solve_pp(n, (w - 1) // 2, _p(band), _p(np.ascontiguousarray(b, dtype=np.float64)), _p(x))
    if rc != 0:
        raise ZeroDivisionError("singular band matrix")
    return x


SOLVERS = {"bicgstab2": 0, "bicgstab1": 1, "minres": 2}
FACT = {"LU_UL": 0, "LU_only": 1}
PREC = {"Spike": 0, "Block": 1, "None": 2}


class Solver:
    """CPU restatement of sap::Solver for a banded FP64 matrix."""

    def __init__(self, P, solver="bicgstab2", max_iters=100, rel_tol=1e-6, abs_tol=0.0,
                 fact="LU_UL", precond="Spike", safe=False, prec_double=False,
                 refine=0, refine_tol=1e-6):
        self.opts = Options(SOLVERS[solver], max_iters, rel_tol, abs_tol, FACT[fact],
                            PREC[precond], int(safe), int(prec_double), int(refine), float(refine_tol))
        self.h = lib().orc_create(P, ctypes.byref(self.opts))
        self.n = 0
        self.k = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def setup(self, band):
        band = np.ascontiguousarray(band, dtype=np.float64)
        self.n, w = band.shape
        self.k = (w - 1) // 2
        return lib().orc_setup_banded(self.h, self.n, self.k, _p(band))

    def solve(self, b, x0=None):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        ok = lib().orc_solve(self.h, _p(np.ascontiguousarray(b, dtype=np.float64)), _p(x))
        return ok, x

    def stats(self):
        st = Stats()
        lib().orc_get_stats(self.h, ctypes.byref(st))
        return {f: getattr(st, f) for f, _ in Stats._fields_}

    def message(self):
        return lib().orc_monitor_message(self.h).decode()

    def apply(self, v):
        z = np.empty(self.n)
        rc = lib().orc_precond_apply(self.h, _p(np.ascontiguousarray(v, dtype=np.float64)), _p(z))
        assert rc == 0
        return z

    @property
    def num_partitions(self):
        return lib().orc_num_partitions(self.h)

    def factors(self):
        out = np.empty((self.n, 2 * self.k + 1))
        lib().orc_get_factors(self.h, _p(out))
        return out

    def reduced(self):
        P = self.num_partitions
        k = self.k
        R = np.empty((max(P - 1, 0), 2 * k, 2 * k))
        WV = np.empty((max(P - 1, 0), 2, k, k))
        lib().orc_get_reduced(self.h, _p(R), _p(WV))
        return R, WV
