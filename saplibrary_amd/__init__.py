"""MI355X-native SaP (SPIKE) solver -- Python binding of libsap_amd.so.

The compute path is the HIP library built from saplibrary_amd/csrc (C ABI
in include/sap_c.h).  This module mirrors the reference's C++ interface
(spikegpu/SaPLibrary sap/solver.h: ``Options``, ``Stats``,
``Solver(numPartitions, opts)`` with ``setup``/``solve``/``getStats``/
``getMonitorCode``/``getMonitorMessage``; sap/banded_matrix.h:
``BandedMatrix``; cusp::csr_matrix: ``CsrMatrix``; sap/spmv.h: ``MVBanded``,
``SpmvCusp``) over device tensors.  PyTorch is
only used for device memory; it is imported first so the library binds to
the HIP runtime already loaded in the process.

There is no CPU fallback: every entry point raises if the library is
missing or the call fails.
"""
import ctypes
import os

import torch  # noqa: F401  (device memory; must be loaded before the HIP library)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsap_amd.so")

# sap::KrylovSolverType (sap/common.h:71-82)
BiCGStab_C, GMRES_C, CG_C, CR_C, BiCGStab1, BiCGStab2, BiCGStab, MINRES = range(8)
# sap::FactorizationMethod (sap/common.h:84-87)
LU_UL, LU_only = 0, 1
# sap::PreconditionerType (sap/common.h:92-96)
Spike, Block, NoPrecond = 0, 1, 2
PREC_FLOAT, PREC_DOUBLE = 0, 1
# largest n of sap_solve_noalloc (SAP_NOALLOC_MAX_N of include/sap_c.h)
NOALLOC_MAX_N = 131072

# sap::system_error::Reason (sap/exception.h:16-23) plus the ABI's own codes
ERRORS = {
    -1: "Zero_pivoting",
    -2: "Negative_DB_weight",
    -3: "Illegal_update",
    -4: "Illegal_solve",
    -5: "Matrix_singular",
    -6: "Out_of_memory",
    -7: "HIP_error",
    -8: "Bad_argument",
    -9: "Unsupported",
}


class SapError(RuntimeError):
    """Counterpart of sap::system_error; ``reason`` is the negative code."""

    def __init__(self, code, message=""):
        self.reason = code
        super().__init__(f"{ERRORS.get(code, code)}: {message}")


class Options(ctypes.Structure):
    """sap::Options (sap/solver.h:52-81) followed by this library's additions."""

    _fields_ = [
        ("solverType", ctypes.c_int),
        ("maxNumIterations", ctypes.c_int),
        ("relTol", ctypes.c_double),
        ("absTol", ctypes.c_double),
        ("testDB", ctypes.c_int),
        ("isSPD", ctypes.c_int),
        ("saveMem", ctypes.c_int),
        ("performReorder", ctypes.c_int),
        ("performDB", ctypes.c_int),
        ("dbFirstStageOnly", ctypes.c_int),
        ("applyScaling", ctypes.c_int),
        ("maxBandwidth", ctypes.c_int),
        ("gpuCount", ctypes.c_int),
        ("dropOffFraction", ctypes.c_double),
        ("factMethod", ctypes.c_int),
        ("precondType", ctypes.c_int),
        ("safeFactorization", ctypes.c_int),
        ("variableBandwidth", ctypes.c_int),
        ("trackReordering", ctypes.c_int),
        ("useBCR", ctypes.c_int),
        ("ilu_level", ctypes.c_int),
        ("precValueType", ctypes.c_int),
        ("refine", ctypes.c_int),
        ("refineTol", ctypes.c_double),
        ("profile", ctypes.c_int),
        ("batch", ctypes.c_int),
        ("graph", ctypes.c_int),
        ("fuse", ctypes.c_int),
        ("autotune", ctypes.c_int),
    ]

    def __init__(self, **kw):
        super().__init__()
        lib().sap_default_options(ctypes.byref(self))
        for key, val in kw.items():
            if not hasattr(self, key):
                raise AttributeError(key)
            setattr(self, key, val)


class Stats(ctypes.Structure):
    """sap::Stats (sap/solver.h:89-140), times in ms, plus additions."""

    _fields_ = [(f, ctypes.c_double) for f in (
        "timeSetup", "timeUpdate", "timeSolve", "time_DB", "time_DB_pre", "time_DB_first", "time_DB_second",
        "time_DB_post", "d_p1", "d_p1_ori", "diag_dom", "diag_dom_ori", "time_reorder", "time_dropOff",
        "time_cpu_assemble", "time_transfer", "time_toBanded", "time_offDiags", "time_bandLU", "time_bandUL",
        "time_fullLU", "time_assembly", "time_shuffle", "time_bcr_lu", "time_bcr_sweep_deflation",
        "time_bcr_mat_mul_deflation", "time_bcr_sweep_inflation", "time_bcr_mv_inflation")] + [
        ("bandwidthReorder", ctypes.c_int),
        ("bandwidthDB", ctypes.c_int),
        ("bandwidth", ctypes.c_int),
        ("nuKf", ctypes.c_double),
        ("flops_LU", ctypes.c_double),
        ("numPartitions", ctypes.c_int),
        ("actualDropOff", ctypes.c_double),
        ("numIterations", ctypes.c_float),
        ("rhsNorm", ctypes.c_double),
        ("residualNorm", ctypes.c_double),
        ("relResidualNorm", ctypes.c_double),
        ("actual_nnz", ctypes.c_int),
        ("numApplies", ctypes.c_int),
        ("numSpmv", ctypes.c_int),
        ("numRefine", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KernelProfile(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("launches", ctypes.c_int), ("total_ms", ctypes.c_double)]


SPMV_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int)

# every symbol include/sap_c.h declares
EXPORTS = (
    "sap_default_options", "sap_create", "sap_destroy", "sap_set_stream", "sap_setup_banded", "sap_solve",
    "sap_get_stats", "sap_monitor_code", "sap_monitor_message", "sap_last_error", "sap_precond_apply",
    "sap_band_spmv", "sap_get_profile", "sap_reset_profile", "sap_get_factors", "sap_get_spikes",
    "sap_num_partitions", "sap_gen_banded", "sap_gen_vector", "sap_setup_csr", "sap_csr_spmv",
    "sap_reorder_csr", "sap_reordering_info", "sap_reordering_get", "sap_reordering_free", "sap_update",
    "sap_reordering_assemble_var", "sap_reordering_var_get", "sap_mm_read", "sap_mm_info", "sap_mm_get_csr",
    "sap_mm_free", "sap_solve_noalloc", "sap_batch_create", "sap_batch_destroy", "sap_batch_set_stream",
    "sap_batch_setup_banded", "sap_batch_solve", "sap_batch_result", "sap_batch_times", "sap_batch_last_error",
    "sap_autotune_entries", "sap_stream_read_ms", "sap_reordering_dominance", "sap_mlp_create", "sap_mlp_destroy", "sap_mlp_num_params", "sap_mlp_get_params",
    "sap_mlp_set_params", "sap_mlp_forward", "sap_mlp_gradient", "sap_mlp_train", "sap_mlp_last_error",
)

class TuneEntry(ctypes.Structure):
    """sap_tune_entry: one choice of the autotuner."""

    _fields_ = [("family", ctypes.c_int), ("k", ctypes.c_int), ("elem_bytes", ctypes.c_int), ("value", ctypes.c_int),
                ("default_value", ctypes.c_int), ("candidates", ctypes.c_int), ("best_ms", ctypes.c_double),
                ("default_ms", ctypes.c_double)]


def stream_read_gbs(nbytes=1 << 31, reps=5):
    """The practical streaming-read bandwidth (GB/s) of the library's own
    read kernel over a device buffer of nbytes."""
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    ms = ctypes.c_double()
    _check(lib().sap_stream_read_ms(ctypes.c_void_p(x.data_ptr()), x.numel(), reps, ctypes.byref(ms)))
    return nbytes / (ms.value * 1e-3) / 1e9


def autotune_entries():
    """The autotuner's choices so far: [dict] (family 0 = conversion tile
    width, 1 = sweep staging depth)."""
    m = lib().sap_autotune_entries(None, 0)
    arr = (TuneEntry * max(m, 1))()
    lib().sap_autotune_entries(arr, m)
    return [{f: getattr(arr[j], f) for f, _ in TuneEntry._fields_} for j in range(m)]


_lib = None


def lib():
    """Loads libsap_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C saplibrary_amd`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, i64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    L.sap_default_options.argtypes = [ctypes.POINTER(Options)]
    L.sap_default_options.restype = None
    L.sap_create.argtypes = [i, ctypes.POINTER(Options)]
    L.sap_create.restype = vp
    L.sap_destroy.argtypes = [vp]
    L.sap_destroy.restype = None
    L.sap_set_stream.argtypes = [vp, vp]
    L.sap_setup_banded.argtypes = [vp, i, i, vp]
    L.sap_solve.argtypes = [vp, SPMV_FN, vp, vp, vp, i]
    L.sap_solve_noalloc.argtypes = [vp, SPMV_FN, vp, vp, vp, i]
    L.sap_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.sap_monitor_code.argtypes = [vp]
    L.sap_monitor_message.argtypes = [vp]
    L.sap_monitor_message.restype = ctypes.c_char_p
    L.sap_last_error.argtypes = [vp]
    L.sap_last_error.restype = ctypes.c_char_p
    L.sap_precond_apply.argtypes = [vp, vp, vp]
    L.sap_band_spmv.argtypes = [i, i, vp, vp, vp, vp]
    L.sap_get_profile.argtypes = [vp, ctypes.POINTER(KernelProfile), i]
    L.sap_reset_profile.argtypes = [vp]
    L.sap_reset_profile.restype = None
    L.sap_get_factors.argtypes = [vp, ctypes.POINTER(d), i64]
    L.sap_get_spikes.argtypes = [vp, ctypes.POINTER(d), ctypes.POINTER(d), i64]
    L.sap_num_partitions.argtypes = [vp]
    L.sap_gen_banded.argtypes = [i64, i, d, ctypes.c_uint64, vp, vp]
    L.sap_gen_vector.argtypes = [i64, ctypes.c_uint64, vp, vp]
    L.sap_setup_csr.argtypes = [vp, i, i, vp, vp, vp]
    L.sap_update.argtypes = [vp, vp, i64]
    L.sap_csr_spmv.argtypes = [i, i, vp, vp, vp, vp, vp, vp]
    ip, dp = ctypes.POINTER(i), ctypes.POINTER(d)
    L.sap_reorder_csr.argtypes = [i, ip, ip, dp, ctypes.POINTER(Options), ip]
    L.sap_reorder_csr.restype = vp
    L.sap_reordering_info.argtypes = [vp, ip, ip, ip, dp, ip]
    L.sap_reordering_get.argtypes = [vp, ip, ip, ip, dp, dp, ip, ip, dp]
    L.sap_reordering_dominance.argtypes = [vp, dp, dp, dp, dp]
    L.sap_reordering_free.argtypes = [vp]
    L.sap_reordering_free.restype = None
    L.sap_reordering_assemble_var.argtypes = [vp, i]
    L.sap_reordering_var_get.argtypes = [vp, ip, ctypes.POINTER(i64), ip, ip, ip, ip, ip, ip, ip, dp, dp, ip, ip,
                                         ip, ip, dp]
    L.sap_mm_read.argtypes = [ctypes.c_char_p, ip, ctypes.c_char_p, i]
    L.sap_mm_read.restype = vp
    L.sap_mm_info.argtypes = [vp, ip, ip, ctypes.POINTER(i64)]
    L.sap_mm_get_csr.argtypes = [vp, ip, ip, dp]
    L.sap_mm_free.argtypes = [vp]
    L.sap_mm_free.restype = None
    L.sap_batch_create.argtypes = [i, ip, ctypes.POINTER(Options)]
    L.sap_batch_create.restype = vp
    L.sap_batch_destroy.argtypes = [vp]
    L.sap_batch_destroy.restype = None
    L.sap_batch_set_stream.argtypes = [vp, vp]
    L.sap_batch_setup_banded.argtypes = [vp, ip, ip, ctypes.POINTER(vp)]
    L.sap_batch_solve.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ip]
    L.sap_batch_result.argtypes = [vp, i, ip, ctypes.POINTER(ctypes.c_float), dp]
    L.sap_batch_times.argtypes = [vp, dp, dp]
    L.sap_autotune_entries.argtypes = [ctypes.POINTER(TuneEntry), i]
    L.sap_stream_read_ms.argtypes = [vp, i64, i, dp]
    fp = ctypes.POINTER(ctypes.c_float)
    L.sap_mlp_create.argtypes = [i, i, i, ctypes.c_uint64]
    L.sap_mlp_create.restype = vp
    L.sap_mlp_destroy.argtypes = [vp]
    L.sap_mlp_destroy.restype = None
    L.sap_mlp_num_params.argtypes = [vp]
    L.sap_mlp_num_params.restype = i64
    L.sap_mlp_get_params.argtypes = [vp, fp]
    L.sap_mlp_set_params.argtypes = [vp, fp]
    L.sap_mlp_forward.argtypes = [vp, vp, i, vp]
    L.sap_mlp_gradient.argtypes = [vp, vp, vp, vp, i, fp, fp]
    L.sap_mlp_train.argtypes = [vp, vp, vp, vp, i, i, ctypes.c_float, fp]
    L.sap_mlp_last_error.argtypes = [vp]
    L.sap_mlp_last_error.restype = ctypes.c_char_p
    L.sap_batch_last_error.argtypes = [vp]
    L.sap_batch_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _dev_f64(t, name):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous float64 device tensor")
    return ctypes.c_void_p(t.data_ptr())


def _dev_i32(t, name):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.int32 or not t.is_cuda or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous int32 device tensor")
    return ctypes.c_void_p(t.data_ptr())


def _check(code, handle=None):
    if code < 0:
        msg = lib().sap_last_error(handle).decode() if handle else ""
        raise SapError(code, msg)
    return code


class BandedMatrix:
    """Row-major FP64 band on the device (sap/banded_matrix.h:25-95):
    A(i, j) at band[i, k + j - i]."""

    def __init__(self, band):
        if band.dim() != 2 or (band.shape[1] - 1) % 2:
            raise ValueError("band must be n x (2k+1)")
        _dev_f64(band, "band")
        self.band = band
        self.num_rows = band.shape[0]
        self.m_k = (band.shape[1] - 1) // 2

    @staticmethod
    def generate(n, k, dominance, seed, device="cuda"):
        band = torch.empty((n, 2 * k + 1), dtype=torch.float64, device=device)
        _check(lib().sap_gen_banded(n, k, dominance, seed, ctypes.c_void_p(band.data_ptr()), None))
        return BandedMatrix(band)


def gen_vector(n, seed, device="cuda"):
    x = torch.empty(n, dtype=torch.float64, device=device)
    _check(lib().sap_gen_vector(n, seed, ctypes.c_void_p(x.data_ptr()), None))
    return x


def band_spmv(A, x, y):
    """y = A x for a BandedMatrix (the MVBanded functor, sap/spmv.h:27-74)."""
    _check(lib().sap_band_spmv(A.num_rows, A.m_k, _dev_f64(A.band, "A"), _dev_f64(x, "x"), _dev_f64(y, "y"), None))
    return y


class CsrMatrix:
    """FP64 CSR matrix on the device with zero-based int32 indices (the
    cusp::csr_matrix<int, double, device_memory> the reference's Solver::setup
    and SpmvCusp take)."""

    def __init__(self, row_offsets, column_indices, values, num_cols=None):
        _dev_i32(row_offsets, "row_offsets")
        _dev_i32(column_indices, "column_indices")
        _dev_f64(values, "values")
        if row_offsets.dim() != 1 or column_indices.shape != values.shape:
            raise ValueError("inconsistent CSR arrays")
        self.row_offsets = row_offsets
        self.column_indices = column_indices
        self.values = values
        self.num_rows = row_offsets.shape[0] - 1
        self.num_cols = self.num_rows if num_cols is None else num_cols
        self.num_entries = values.shape[0]
        if self.num_cols != self.num_rows:
            raise ValueError("the matrix must be square")

    @staticmethod
    def from_host(row_offsets, column_indices, values, device="cuda"):
        """From numpy / host arrays (scipy.sparse.csr_matrix fields)."""
        return CsrMatrix(torch.as_tensor(row_offsets, dtype=torch.int32).to(device).contiguous(),
                         torch.as_tensor(column_indices, dtype=torch.int32).to(device).contiguous(),
                         torch.as_tensor(values, dtype=torch.float64).to(device).contiguous())

    @staticmethod
    def from_matrix_market(path, device="cuda"):
        """Reads a Matrix Market file (read_matrix_market) onto the device."""
        rows, cols, rp, ci, val = read_matrix_market(path)
        if rows != cols:
            raise ValueError(f"{path}: the matrix must be square ({rows} x {cols})")
        return CsrMatrix.from_host(rp, ci, val, device=device)

    def _ptrs(self):
        return (_dev_i32(self.row_offsets, "row_offsets"), _dev_i32(self.column_indices, "column_indices"),
                _dev_f64(self.values, "values"))


def csr_spmv(A, x, y):
    """y = A x for a CsrMatrix (the SpmvCusp functor, sap/spmv.h:83-129)."""
    rp, ci, val = A._ptrs()
    _check(lib().sap_csr_spmv(A.num_rows, A.num_entries, rp, ci, val, _dev_f64(x, "x"), _dev_f64(y, "y"), None))
    return y


def read_matrix_market(path):
    """A Matrix Market file as host CSR arrays (row_offsets, column_indices,
    values; int32, int32, float64) through the library's reader (the
    cusp::io::read_matrix_market_file call of the reference's drivers,
    examples/matrix_market/driver_mm.cu:135-140).  Returns
    (rows, cols, row_offsets, column_indices, values); raises SapError."""
    import numpy as np
    L = lib()
    err = ctypes.c_int(0)
    msg = ctypes.create_string_buffer(512)
    h = L.sap_mm_read(os.fsencode(path), ctypes.byref(err), msg, len(msg))
    if not h:
        raise SapError(err.value, msg.value.decode(errors="replace"))
    h = ctypes.c_void_p(h)
    try:
        rows, cols, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        _check(L.sap_mm_info(h, ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz)))
        rp = np.empty(rows.value + 1, dtype=np.int32)
        ci = np.empty(max(nnz.value, 1), dtype=np.int32)
        val = np.empty(max(nnz.value, 1))
        ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        _check(L.sap_mm_get_csr(h, rp.ctypes.data_as(ip), ci.ctypes.data_as(ip), val.ctypes.data_as(dp)))
        return rows.value, cols.value, rp, ci[:nnz.value], val[:nnz.value]
    finally:
        L.sap_mm_free(h)


class SpmvCusp:
    """SPMV functor over a CsrMatrix (sap/spmv.h:83-129)."""

    def __init__(self, A):
        self.A = A
        self.num_rows = A.num_rows

    def __call__(self, v, Av):
        csr_spmv(self.A, v, Av)


def reorder_csr(row_offsets, column_indices, values, opts=None, numPartitions=None):
    """The host reordering of a CSR setup (DB, RCM, drop-off; sap/graph.h:446-747)
    on host arrays, without a GPU.  Returns a dict with the permutations
    (db_perm, row_perm, col_reord), scalings, bandwidths (k, k_db, k_reorder),
    the actual drop-off fraction and the reordered CSR (rp, ci, val).
    Raises SapError (Matrix_singular, Negative_DB_weight)."""
    import numpy as np
    rp = np.ascontiguousarray(row_offsets, dtype=np.int32)
    ci = np.ascontiguousarray(column_indices, dtype=np.int32)
    val = np.ascontiguousarray(values, dtype=np.float64)
    n = rp.shape[0] - 1
    o = opts if opts is not None else Options()
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    err = ctypes.c_int(0)
    L = lib()
    r = L.sap_reorder_csr(n, rp.ctypes.data_as(ip), ci.ctypes.data_as(ip), val.ctypes.data_as(dp),
                          ctypes.byref(o), ctypes.byref(err))
    if not r:
        raise SapError(err.value, "reorder_csr failed")
    r = ctypes.c_void_p(r)
    try:
        k, kdb, kre, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        drop = ctypes.c_double()
        _check(L.sap_reordering_info(r, ctypes.byref(k), ctypes.byref(kdb), ctypes.byref(kre), ctypes.byref(drop),
                                     ctypes.byref(nnz)))
        out = {name: np.empty(n, dtype=np.int32) for name in ("db_perm", "row_perm", "col_reord")}
        out["row_scale"] = np.empty(n)
        out["col_scale"] = np.empty(n)
        out["rp"] = np.empty(n + 1, dtype=np.int32)
        out["ci"] = np.empty(max(nnz.value, 1), dtype=np.int32)
        out["val"] = np.empty(max(nnz.value, 1))
        _check(L.sap_reordering_get(r, *(out[f].ctypes.data_as(ip) for f in ("db_perm", "row_perm", "col_reord")),
                                    out["row_scale"].ctypes.data_as(dp), out["col_scale"].ctypes.data_as(dp),
                                    out["rp"].ctypes.data_as(ip), out["ci"].ctypes.data_as(ip),
                                    out["val"].ctypes.data_as(dp)))
        out["ci"] = out["ci"][:nnz.value]
        out["val"] = out["val"][:nnz.value]
        dd = [ctypes.c_double() for _ in range(4)]
        _check(L.sap_reordering_dominance(r, *(ctypes.byref(x) for x in dd)))
        out["diag_dom_ori"], out["d_p1_ori"], out["diag_dom"], out["d_p1"] = (x.value for x in dd)
        out.update(k=k.value, k_db=kdb.value, k_reorder=kre.value, dropoff=drop.value)
        if numPartitions is not None:
            P = L.sap_reordering_assemble_var(r, numPartitions)
            _check(P)
            out["P"] = P
            if P > 1:
                kk = k.value
                var = dict(ks=np.empty(P, dtype=np.int32), boffsets=np.empty(P + 1, dtype=np.int64),
                           first_rows=np.empty(P - 1, dtype=np.int32), second_reorder=np.empty(n, dtype=np.int32),
                           second_perm=np.empty(n, dtype=np.int32), widths_left=np.empty(P - 1, dtype=np.int32),
                           widths_right=np.empty(P - 1, dtype=np.int32),
                           perms_left=np.empty((P - 1) * kk, dtype=np.int32),
                           perms_right=np.empty((P - 1) * kk, dtype=np.int32),
                           wv=np.empty(2 * kk * kk * (P - 1)), offdiags=np.empty(2 * kk * kk * (P - 1)),
                           ks_col=np.empty(n, dtype=np.int32), ks_row=np.empty(n, dtype=np.int32),
                           drp=np.empty(n + 1, dtype=np.int32), dci=np.empty(nnz.value, dtype=np.int32),
                           dval=np.empty(max(nnz.value, 1)))
                names = ("ks", "boffsets", "first_rows", "second_reorder", "second_perm", "widths_left",
                         "widths_right", "perms_left", "perms_right", "wv", "offdiags", "ks_col", "ks_row", "drp",
                         "dci", "dval")
                ptrs = [var[f].ctypes.data_as(ctypes.POINTER(ctypes.c_int64) if f == "boffsets" else
                                              dp if var[f].dtype == np.float64 else ip) for f in names]
                _check(L.sap_reordering_var_get(r, *ptrs))
                nd = int(var["drp"][-1])
                var["dci"] = var["dci"][:nd]
                var["dval"] = var["dval"][:nd]
                out["var"] = var
        return out
    finally:
        L.sap_reordering_free(r)


class MVBanded:
    """SPMV functor over a BandedMatrix (sap/spmv.h:27-74)."""

    def __init__(self, A):
        self.A = A
        self.num_rows = A.num_rows

    def __call__(self, v, Av):
        band_spmv(self.A, v, Av)


class Solver:
    """sap::Solver<Array, PrecValueType> (sap/solver.h:153-225) over the C ABI."""

    def __init__(self, numPartitions, opts=None):
        self.opts = opts if opts is not None else Options()
        h = lib().sap_create(numPartitions, ctypes.byref(self.opts))
        if not h:
            raise SapError(-9, "sap_create rejected the options")
        self.h = ctypes.c_void_p(h)
        self._A = None
        self._keep = None

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.sap_destroy(h)
            self.h = None

    def setup(self, A):
        """Builds the preconditioner (sap/solver.h:339-404) from a BandedMatrix
        or a CsrMatrix.  Raises SapError."""
        if isinstance(A, BandedMatrix):
            _check(lib().sap_setup_banded(self.h, A.num_rows, A.m_k, _dev_f64(A.band, "A")), self.h)
        elif isinstance(A, CsrMatrix):
            _check(lib().sap_setup_csr(self.h, A.num_rows, A.num_entries, *A._ptrs()), self.h)
        else:
            raise TypeError("setup() takes a BandedMatrix or a CsrMatrix")
        self._A = A
        return True

    def update(self, entries):
        """Refactors with new values of the setup matrix (sap/solver.h:421-462):
        a float64 device tensor of the CSR values (CsrMatrix setup) or of the
        n x (2k+1) band (BandedMatrix setup).  Returns False when the number
        of entries does not match; raises SapError (Illegal_update without
        trackReordering or before setup).  The built-in SpMV then uses
        ``entries``, which must stay alive."""
        rc = lib().sap_update(self.h, _dev_f64(entries, "entries"), entries.numel())
        _check(rc, self.h)
        if rc == 1:
            if isinstance(self._A, CsrMatrix):
                self._A = CsrMatrix(self._A.row_offsets, self._A.column_indices, entries)
            elif isinstance(self._A, BandedMatrix):
                self._A = BandedMatrix(entries.view(self._A.band.shape))
        return rc == 1

    def _n(self):
        if self._A is None:
            raise SapError(-4, "Illegal call to solve() before setup().")
        return self._A.num_rows

    def solve(self, spmv, b, x):
        """Solves A x = b (sap/solver.h:477-562); x holds the initial guess.
        spmv: None (built-in SpMV of the setup matrix), an MVBanded / SpmvCusp
        of that matrix, or any callable f(v, Av) on device tensors.
        b and x are vectors of the setup's n rows, or n x nrhs matrices whose
        columns are solved one after another (the C ABI takes them
        column-major with leading dimension n; they are transposed here and
        x is written back)."""
        return self._solve(lib().sap_solve, spmv, b, x)

    def solve_noalloc(self, spmv, b, x):
        """solve() through sap_solve_noalloc: the Krylov vectors live in the
        library's static device array instead of a pool allocated on the
        first solve (n <= NOALLOC_MAX_N, profile off).  Same results."""
        return self._solve(lib().sap_solve_noalloc, spmv, b, x)

    def _solve(self, entry, spmv, b, x):
        if self._A is None:
            # the library reports Illegal_solve itself
            _check(entry(self.h, SPMV_FN(), None, _dev_f64(b, "b"), _dev_f64(x, "x"), 1), self.h)
        n = self._n()
        if b.dim() not in (1, 2) or b.shape != x.shape or b.shape[0] != n:
            raise ValueError(f"b and x must both be ({n},) or ({n}, nrhs); got {tuple(b.shape)} and {tuple(x.shape)}")
        nrhs = 1 if b.dim() == 1 else b.shape[1]
        if spmv is None or (isinstance(spmv, (MVBanded, SpmvCusp)) and spmv.A is self._A):
            fn = SPMV_FN()
        else:
            def _cb(ctx, pv, pav, nn):
                v = _wrap(pv, nn, b.device)
                av = _wrap(pav, nn, b.device)
                spmv(v, av)
            fn = SPMV_FN(_cb)
        self._keep = fn
        if nrhs > 1 or b.dim() == 2:
            bc = b.t().contiguous()
            xc = x.t().contiguous()
        else:
            bc, xc = b, x
        rc = entry(self.h, fn, None, _dev_f64(bc, "b"), _dev_f64(xc, "x"), nrhs)
        _check(rc, self.h)
        if xc is not x:
            x.copy_(xc.t())
        return rc == 1

    def precond_apply(self, v, z):
        """z = M^{-1} v (Precond::operator(), sap/precond.h:1370-1386)."""
        n = self._n()
        if v.numel() != n or z.numel() != n:
            raise ValueError(f"v and z must hold the setup's {n} rows")
        _check(lib().sap_precond_apply(self.h, _dev_f64(v, "v"), _dev_f64(z, "z")), self.h)
        return z

    def getStats(self):
        st = Stats()
        _check(lib().sap_get_stats(self.h, ctypes.byref(st)), self.h)
        return st

    def getMonitorCode(self):
        return lib().sap_monitor_code(self.h)

    def getMonitorMessage(self):
        return lib().sap_monitor_message(self.h).decode()

    @property
    def numPartitions(self):
        return lib().sap_num_partitions(self.h)

    def profile(self):
        m = lib().sap_get_profile(self.h, None, 0)
        _check(m, self.h)
        arr = (KernelProfile * max(m, 1))()
        lib().sap_get_profile(self.h, arr, m)
        return {arr[j].name.decode(): (arr[j].launches, arr[j].total_ms) for j in range(m)}

    def reset_profile(self):
        lib().sap_reset_profile(self.h)

    def factors(self):
        import numpy as np
        n, k = self._A.num_rows, self.getStats().bandwidth
        out = np.empty((n, 2 * k + 1))
        _check(lib().sap_get_factors(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size), self.h)
        return out

    def factors_flat(self, count):
        """The factored storage as a flat array of count values: on the
        variable-bandwidth path the partitions' own bands one after the other
        (partition p in a (2 K_p + 1) x n_p column-major band, second-level
        order)."""
        import numpy as np
        out = np.empty(count)
        _check(lib().sap_get_factors(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size), self.h)
        return out

    def spikes(self):
        """(WV, S): spike tips (P-1, 2, k, k) stored RHS-major and the factored
        Schur blocks (P-1, k, k) stored column-major."""
        import numpy as np
        k = self.getStats().bandwidth
        P = self.numPartitions
        wv = np.empty((max(P - 1, 0), 2, k, k))
        s = np.empty((max(P - 1, 0), k, k))
        _check(lib().sap_get_spikes(self.h, wv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), wv.size), self.h)
        return wv, s


class Mlp:
    """The option selector's network (sap_mlp_* of the C ABI): in -> hidden
    -> hidden -> hidden -> out, ReLU, FP32, trained by the library's own
    HIP kernels.  Arrays are float32 device tensors."""

    def __init__(self, n_in, n_out, n_hidden=64, seed=0):
        h = lib().sap_mlp_create(n_in, n_hidden, n_out, seed)
        if not h:
            raise SapError(-8, "sap_mlp_create: bad sizes")
        self.h = ctypes.c_void_p(h)
        self.n_in, self.n_out, self.n_hidden = n_in, n_out, n_hidden

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.sap_mlp_destroy(h)
            self.h = None

    def _check(self, rc):
        if rc < 0:
            raise SapError(rc, lib().sap_mlp_last_error(self.h).decode())
        return rc

    @staticmethod
    def _f32(t, name):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise TypeError(f"{name} must be a contiguous float32 device tensor")
        return ctypes.c_void_p(t.data_ptr())

    def num_params(self):
        return lib().sap_mlp_num_params(self.h)

    def params(self):
        import numpy as np
        out = np.empty(self.num_params(), dtype=np.float32)
        self._check(lib().sap_mlp_get_params(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params()
        self._check(lib().sap_mlp_set_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def forward(self, X):
        Y = torch.empty((X.shape[0], self.n_out), dtype=torch.float32, device=X.device)
        self._check(lib().sap_mlp_forward(self.h, self._f32(X, "X"), X.shape[0], self._f32(Y, "Y")))
        return Y

    def gradient(self, X, T, M):
        import numpy as np
        g = np.empty(self.num_params(), dtype=np.float32)
        loss = ctypes.c_float()
        self._check(lib().sap_mlp_gradient(self.h, self._f32(X, "X"), self._f32(T, "T"), self._f32(M, "M"),
                                           X.shape[0], g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                           ctypes.byref(loss)))
        return g, loss.value

    def train(self, X, T, M, epochs, lr=1e-3):
        import numpy as np
        losses = np.empty(max(epochs, 1), dtype=np.float32)
        self._check(lib().sap_mlp_train(self.h, self._f32(X, "X"), self._f32(T, "T"), self._f32(M, "M"), X.shape[0],
                                        epochs, lr, losses.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return losses[:epochs]


class SystemBatch:
    """Independent banded systems whose right-hand sides are solved together
    (sap_batch_* of the C ABI): one BiCGStab(2) batch over the columns of all
    systems, each system with its own preconditioner."""

    def __init__(self, numPartitions, opts=None):
        self.opts = opts if opts is not None else Options()
        P = (ctypes.c_int * len(numPartitions))(*numPartitions)
        h = lib().sap_batch_create(len(numPartitions), P, ctypes.byref(self.opts))
        if not h:
            raise SapError(-9, "sap_batch_create rejected the options")
        self.h = ctypes.c_void_p(h)
        self.nsys = len(numPartitions)
        self._A = None
        self._ncols = 0

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.sap_batch_destroy(h)
            self.h = None

    def _check(self, rc):
        if rc < 0:
            raise SapError(rc, lib().sap_batch_last_error(self.h).decode())
        return rc

    def setup(self, mats):
        """mats: one BandedMatrix per system."""
        if len(mats) != self.nsys:
            raise ValueError(f"{self.nsys} matrices expected")
        n = (ctypes.c_int * self.nsys)(*[A.num_rows for A in mats])
        k = (ctypes.c_int * self.nsys)(*[A.m_k for A in mats])
        p = (ctypes.c_void_p * self.nsys)(*[_dev_f64(A.band, "A").value for A in mats])
        self._check(lib().sap_batch_setup_banded(self.h, n, k, p))
        self._A = list(mats)

    def solve(self, B, X):
        """B[s], X[s]: (n_s, nrhs_s) float64 device tensors (X: initial
        guesses, overwritten with the solutions).  Returns True when every
        column converged."""
        if self._A is None:
            raise SapError(-4, "Illegal call to solve() before setup().")
        bc, xc, nr = [], [], []
        for A, b, x in zip(self._A, B, X):
            if b.dim() != 2 or b.shape != x.shape or b.shape[0] != A.num_rows:
                raise ValueError("B[s] and X[s] must be (n_s, nrhs_s)")
            bc.append(b.t().contiguous())
            xc.append(x.t().contiguous())
            nr.append(b.shape[1])
        pb = (ctypes.c_void_p * self.nsys)(*[_dev_f64(t, "B").value for t in bc])
        px = (ctypes.c_void_p * self.nsys)(*[_dev_f64(t, "X").value for t in xc])
        rc = self._check(lib().sap_batch_solve(self.h, pb, px, (ctypes.c_int * self.nsys)(*nr)))
        for x, t in zip(X, xc):
            x.copy_(t.t())
        self._ncols = sum(nr)
        return rc == 1

    def results(self):
        """[(code, iterations, relative residual)] per column, systems in order."""
        out = []
        for c in range(self._ncols):
            code, its, rel = ctypes.c_int(), ctypes.c_float(), ctypes.c_double()
            self._check(lib().sap_batch_result(self.h, c, ctypes.byref(code), ctypes.byref(its), ctypes.byref(rel)))
            out.append((code.value, its.value, rel.value))
        return out

    def times(self):
        """(setup ms, solve ms) of the last setup and solve."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(lib().sap_batch_times(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def _wrap(ptr, n, device):
    """A float64 device tensor view of n doubles at ptr (for SPMV callbacks)."""
    class _Arr:
        def __init__(self, p, nn):
            self.__cuda_array_interface__ = {"shape": (nn,), "typestr": "<f8", "data": (p, False), "version": 2}
    return torch.as_tensor(_Arr(ptr, n), device=device)
