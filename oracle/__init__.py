"""Oracle: CPU restatement of Spira's locked-mode lock() and SpMV.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or the timed CPU baseline -- never as the thing measured or shipped.
The product library (``spira_amd``) neither imports nor links it.

Every function wraps ``libspira_oracle.so`` (plain C, ``oracle/spira_oracle.c``)
and cites the reference code it restates:

===================  ===========================================================
``dedup``            aos_array_buffer.hpp:76-101 (reverse, stable_sort, keep first)
``build_csr``        csr_build.hpp:32-98
``merge_csr``        csr_build.hpp:126-488
``spmv_f64``         serial/spmv.hpp:48-54 (and dot_scalar.cpp:4-11)
``spmv_f32``         dot_scalar.cpp:13 (FP32 loop, FP32 accumulator)
``spmv_dd``          double-double error reference (not in the reference)
``Pool``             parallel_matrix.hpp:289-405, thread_pool.hpp:102-134,
                     parallel/algorithms/spmv.hpp:127-161
``partition_bounds`` partition.hpp:78-126
``transpose``        serial/transpose.hpp:35-157; with drop_zeros,
                     parallel/algorithms/transpose.hpp:37-69 (insert + lock)
``add``              serial/matrix_addition.hpp:34-112 (subtract:
                     matrix_operators.hpp:32-41 with scalars.hpp:23-35)
``spgemm``           serial/spgemm.hpp:12-50
===================  ===========================================================
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libspira_oracle.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "spira_or_is_zero": (ctypes.c_int, [ctypes.c_double]),
            "spira_or_dedup": (ctypes.c_uint64, [ctypes.c_uint64, _u64p, _u32p, _f64p, _u64p, _u32p, _f64p]),
            "spira_or_build_csr": (ctypes.c_uint64, [ctypes.c_uint64, _u64p, _u32p, _f64p, _u64p, _u32p, _f64p]),
            "spira_or_merge_csr": (ctypes.c_uint64, [ctypes.c_uint64, _u64p, _u32p, _f64p, _u64p, _u32p, _f64p,
                                                     _u8p, _u64p, _u32p, _f64p]),
            "spira_or_spmv_f64": (None, [ctypes.c_uint64, _u64p, _u32p, _f64p, _f64p, _f64p]),
            "spira_or_spmv_f32": (None, [ctypes.c_uint64, _u64p, _u32p, _f32p, _f32p, _f32p]),
            "spira_or_spmv_dd": (None, [ctypes.c_uint64, _u64p, _u32p, _f64p, _f64p, _f64p, _f64p]),
            "spira_or_spmm_f64": (None, [ctypes.c_uint64, _u64p, _u32p, _f64p, ctypes.c_uint32, _f64p, _f64p]),
            "spira_or_pool_create": (ctypes.c_void_p, [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64]),
            "spira_or_pool_destroy": (None, [ctypes.c_void_p]),
            "spira_or_pool_threads": (ctypes.c_uint32, [ctypes.c_void_p]),
            "spira_or_pool_set_kernel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
            "spira_or_pool_force_kernel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
            "spira_or_pool_pin": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
            "spira_or_pool_flush": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
            "spira_or_pool_partition": (None, [ctypes.c_void_p, ctypes.c_uint32, _u64p, _u64p]),
            "spira_or_pool_load_csr": (ctypes.c_int, [ctypes.c_void_p, _u64p, _u32p, _f64p]),
            "spira_or_pool_spmv": (ctypes.c_int, [ctypes.c_void_p, _f64p, _f64p]),
            "spira_or_pool_lock": (ctypes.c_int, [ctypes.c_void_p, _u64p, _u32p, _f64p]),
            "spira_or_pool_nnz": (ctypes.c_uint64, [ctypes.c_void_p]),
            "spira_or_pool_export_csr": (ctypes.c_int, [ctypes.c_void_p, _u64p, _u32p, _f64p]),
            "spira_or_partition_boundaries": (None, [ctypes.c_uint64, _u64p, ctypes.c_uint64, _u64p]),
            "spira_or_add": (ctypes.c_uint64, [ctypes.c_uint64, _u64p, _u32p, _f64p, _u64p, _u32p, _f64p,
                                               ctypes.c_int, _u64p, _u32p, _f64p]),
            "spira_or_spgemm": (ctypes.c_int64, [ctypes.c_uint64, ctypes.c_uint64, _u64p, _u32p, _f64p, _u64p,
                                                 _u32p, _f64p, _u64p, _u32p, _f64p, ctypes.c_uint64]),
            "spira_or_transpose": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64, _u64p, _u32p, _f64p,
                                                     ctypes.c_int, _u64p, _u32p, _f64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def is_zero(v: float) -> bool:
    return bool(lib().spira_or_is_zero(float(v)))


def dedup(stage_ptr, stage_col, stage_val):
    """Per-row last-write-wins dedup, sorted by column, zeros kept."""
    sp, sc, sv = _u64(stage_ptr), _u32(stage_col), _f64(stage_val)
    n = sp.size - 1
    m = int(sp[-1]) if n >= 0 else 0
    dptr = np.zeros(n + 1, np.uint64)
    dcol = np.zeros(max(m, 1), np.uint32)
    dval = np.zeros(max(m, 1), np.float64)
    w = lib().spira_or_dedup(n, _p(sp, _u64p), _p(sc, _u32p), _p(sv, _f64p),
                             _p(dptr, _u64p), _p(dcol, _u32p), _p(dval, _f64p))
    return dptr, dcol[:w].copy(), dval[:w].copy()


def build_csr(dptr, dcol, dval):
    dp, dc, dv = _u64(dptr), _u32(dcol), _f64(dval)
    n = dp.size - 1
    m = int(dp[-1])
    off = np.zeros(n + 1, np.uint64)
    col = np.zeros(max(m, 1), np.uint32)
    val = np.zeros(max(m, 1), np.float64)
    nnz = lib().spira_or_build_csr(n, _p(dp, _u64p), _p(dc, _u32p), _p(dv, _f64p),
                                   _p(off, _u64p), _p(col, _u32p), _p(val, _f64p))
    return off, col[:nnz].copy(), val[:nnz].copy()


def merge_csr(old_off, old_col, old_val, dptr, dcol, dval, dirty=None):
    oo, oc, ov = _u64(old_off), _u32(old_col), _f64(old_val)
    dp, dc, dv = _u64(dptr), _u32(dcol), _f64(dval)
    n = oo.size - 1
    assert dp.size == n + 1
    cap = int(oo[-1]) + int(dp[-1])
    off = np.zeros(n + 1, np.uint64)
    col = np.zeros(max(cap, 1), np.uint32)
    val = np.zeros(max(cap, 1), np.float64)
    dirty_p = None
    if dirty is not None:
        d = np.ascontiguousarray(dirty, dtype=np.uint8)
        assert d.size == n
        dirty_p = _p(d, _u8p)
    nnz = lib().spira_or_merge_csr(n, _p(oo, _u64p), _p(oc, _u32p), _p(ov, _f64p),
                                   _p(dp, _u64p), _p(dc, _u32p), _p(dv, _f64p), dirty_p,
                                   _p(off, _u64p), _p(col, _u32p), _p(val, _f64p))
    return off, col[:nnz].copy(), val[:nnz].copy()


def lock(stage_ptr, stage_col, stage_val, old=None, dirty=None):
    """matrix::lock() (matrix.hpp:277-342): dedup, then build or merge."""
    dptr, dcol, dval = dedup(stage_ptr, stage_col, stage_val)
    if old is None:
        return build_csr(dptr, dcol, dval)
    return merge_csr(*old, dptr, dcol, dval, dirty)


def spmv_f64(off, col, val, x):
    o, c, v, xx = _u64(off), _u32(col), _f64(val), _f64(x)
    n = o.size - 1
    y = np.empty(n, np.float64)
    lib().spira_or_spmv_f64(n, _p(o, _u64p), _p(c, _u32p), _p(v, _f64p), _p(xx, _f64p), _p(y, _f64p))
    return y


def spmv_f32(off, col, val, x):
    o, c = _u64(off), _u32(col)
    v = np.ascontiguousarray(val, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    n = o.size - 1
    y = np.empty(n, np.float32)
    lib().spira_or_spmv_f32(n, _p(o, _u64p), _p(c, _u32p), _p(v, _f32p), _p(xx, _f32p), _p(y, _f32p))
    return y


def spmv_dd(off, col, val, x):
    """Double-double SpMV; returns (hi, lo) with hi+lo accurate to ~2^-104."""
    o, c, v, xx = _u64(off), _u32(col), _f64(val), _f64(x)
    n = o.size - 1
    hi = np.empty(n, np.float64)
    lo = np.empty(n, np.float64)
    lib().spira_or_spmv_dd(n, _p(o, _u64p), _p(c, _u32p), _p(v, _f64p), _p(xx, _f64p),
                           _p(hi, _f64p), _p(lo, _f64p))
    return hi, lo


def spmm_f64(off, col, val, X):
    o, c, v = _u64(off), _u32(col), _f64(val)
    Xc = np.ascontiguousarray(X, dtype=np.float64)
    k = Xc.shape[1]
    n = o.size - 1
    Y = np.empty((n, k), np.float64)
    lib().spira_or_spmm_f64(n, _p(o, _u64p), _p(c, _u32p), _p(v, _f64p), k, _p(Xc, _f64p), _p(Y, _f64p))
    return Y


def partition_bounds(row_nnz, n_threads):
    rn = _u64(row_nnz)
    b = np.zeros(n_threads + 1, np.uint64)
    lib().spira_or_partition_boundaries(rn.size, _p(rn, _u64p), n_threads, _p(b, _u64p))
    return b


def transpose(off, col, val, n_cols, drop_zeros=False):
    """CSR of A^T (n_cols rows): columns ascending in every row, values as stored
    (serial); drop_zeros=True also removes explicit zeros (parallel)."""
    o, c, v = _u64(off), _u32(col), _f64(val)
    n = o.size - 1
    t_off = np.zeros(int(n_cols) + 1, np.uint64)
    t_col = np.zeros(max(c.size, 1), np.uint32)
    t_val = np.zeros(max(c.size, 1), np.float64)
    nz = lib().spira_or_transpose(n, int(n_cols), _p(o, _u64p), _p(c, _u32p), _p(v, _f64p), 1 if drop_zeros else 0,
                                  _p(t_off, _u64p), _p(t_col, _u32p), _p(t_val, _f64p))
    return t_off, t_col[:nz].copy(), t_val[:nz].copy()


def add(a, b, subtract=False):
    """C = A + B (A - B): a and b are (off, col, val) of the same row count."""
    ao, ac, av = _u64(a[0]), _u32(a[1]), _f64(a[2])
    bo, bc, bv = _u64(b[0]), _u32(b[1]), _f64(b[2])
    n = ao.size - 1
    cap = max(ac.size + bc.size, 1)
    co = np.zeros(n + 1, np.uint64)
    cc = np.zeros(cap, np.uint32)
    cv = np.zeros(cap, np.float64)
    nz = lib().spira_or_add(n, _p(ao, _u64p), _p(ac, _u32p), _p(av, _f64p), _p(bo, _u64p), _p(bc, _u32p),
                            _p(bv, _f64p), 1 if subtract else 0, _p(co, _u64p), _p(cc, _u32p), _p(cv, _f64p))
    return co, cc[:nz].copy(), cv[:nz].copy()


def spgemm_products(a, b):
    """Number of products a * b of C = A * B (the work of the row-wise form)."""
    ao, ac = _u64(a[0]), _u32(a[1])
    blen = np.diff(_u64(b[0]).astype(np.int64))
    return int(blen[ac].sum()) if ac.size else 0


def spgemm(a, b, n_cols):
    """C = A * B: a = (off, col, val) of A, b of B, n_cols = B's columns."""
    ao, ac, av = _u64(a[0]), _u32(a[1]), _f64(a[2])
    bo, bc, bv = _u64(b[0]), _u32(b[1]), _f64(b[2])
    n = ao.size - 1
    cap = max(min(spgemm_products(a, b), n * max(int(n_cols), 1)), 1)
    co = np.zeros(n + 1, np.uint64)
    cc = np.zeros(cap, np.uint32)
    cv = np.zeros(cap, np.float64)
    nz = lib().spira_or_spgemm(n, int(n_cols), _p(ao, _u64p), _p(ac, _u32p), _p(av, _f64p), _p(bo, _u64p),
                               _p(bc, _u32p), _p(bv, _f64p), _p(co, _u64p), _p(cc, _u32p), _p(cv, _f64p), cap)
    assert nz >= 0
    return co, cc[:nz].copy(), cv[:nz].copy()


def normwise_rel_err(y, y_ref_hi, y_ref_lo=None):
    """||y - y_ref||_2 / ||y_ref||_2, with y_ref optionally double-double."""
    y = np.asarray(y, dtype=np.float64)
    d = (y - y_ref_hi)
    if y_ref_lo is not None:
        d = d - y_ref_lo
    den = np.linalg.norm(y_ref_hi)
    num = np.linalg.norm(d)
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return float(num / den)


class Pool:
    """Threaded parallel_matrix restatement (the timed CPU baseline)."""

    def __init__(self, n_threads: int, n_rows: int, n_cols: int):
        self._h = lib().spira_or_pool_create(n_threads, n_rows, n_cols)
        if not self._h:
            raise ValueError("parallel_matrix requires at least one thread")
        self.n_rows, self.n_cols, self.n_threads = n_rows, n_cols, n_threads
        self._keep = []

    def close(self):
        if self._h:
            lib().spira_or_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    KERNELS = {0: "scalar loop", 1: "scalar dot", 2: "avx2", 3: "avx512"}

    def set_simd(self, on=True):
        """Row dot of spmv(): the SIMD gather dot the reference would dispatch
        to on this CPU (AVX-512F, else AVX2+FMA), or the scalar loop.  Returns
        the name of the dot in use."""
        return self.KERNELS[int(lib().spira_or_pool_set_kernel(self._h, 1 if on else 0))]

    def force_kernel(self, name):
        which = {v: k for k, v in self.KERNELS.items()}[name]
        if lib().spira_or_pool_force_kernel(self._h, which) != 0:
            raise RuntimeError(f"{name} is not supported on this CPU")

    def pin(self, cpus):
        """Pin worker t to CPU cpus[t]."""
        arr = (ctypes.c_int * self.n_threads)(*[int(c) for c in cpus[: self.n_threads]])
        if len(cpus) < self.n_threads or lib().spira_or_pool_pin(self._h, arr) != 0:
            raise RuntimeError("pinning the pool's workers failed")

    def flush_llc(self, nbytes):
        """Every worker reads an arena of nbytes (the cold-LLC mode)."""
        if lib().spira_or_pool_flush(self._h, int(nbytes)) != 0:
            raise MemoryError("flush arena allocation failed")

    def partition(self, t):
        a = ctypes.c_uint64()
        b = ctypes.c_uint64()
        lib().spira_or_pool_partition(self._h, t, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def load_csr(self, off, col, val):
        o, c, v = _u64(off), _u32(col), _f64(val)
        rc = lib().spira_or_pool_load_csr(self._h, _p(o, _u64p), _p(c, _u32p), _p(v, _f64p))
        if rc:
            raise RuntimeError("pool load failed")

    def spmv(self, x, y=None):
        xx = _f64(x)
        if y is None:
            y = np.empty(self.n_rows, np.float64)
        rc = lib().spira_or_pool_spmv(self._h, _p(xx, _f64p), _p(y, _f64p))
        if rc == -2:
            raise RuntimeError("spmv: matrix must be locked")
        return y

    def lock(self, stage_ptr, stage_col, stage_val):
        sp, sc, sv = _u64(stage_ptr), _u32(stage_col), _f64(stage_val)
        rc = lib().spira_or_pool_lock(self._h, _p(sp, _u64p), _p(sc, _u32p), _p(sv, _f64p))
        if rc:
            raise MemoryError("pool lock failed")

    def nnz(self):
        return int(lib().spira_or_pool_nnz(self._h))

    def export_csr(self):
        nnz = self.nnz()
        off = np.zeros(self.n_rows + 1, np.uint64)
        col = np.zeros(max(nnz, 1), np.uint32)
        val = np.zeros(max(nnz, 1), np.float64)
        rc = lib().spira_or_pool_export_csr(self._h, _p(off, _u64p), _p(col, _u32p), _p(val, _f64p))
        if rc:
            raise RuntimeError("export on unlocked pool")
        return off, col[:nnz].copy(), val[:nnz].copy()
