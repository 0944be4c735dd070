"""spira_amd -- Spira's locked-mode hot path (``lock()`` and SpMV) on MI355X.

Python mirror of the reference's operator surface for this path, over the
C-ABI of ``libspira_amd.so`` (``include/spira_amd.h``):

* :class:`Matrix` mirrors ``spira::matrix<soa_tag, uint32_t, double>``
  (``include/spira/matrix/matrix.hpp``): open-mode ``insert`` into per-row
  buffers kept in insertion order on the host, ``lock()`` on the device
  (``build_csr`` / ``merge_csr``), ``open()``, ``clear()``, ``load_csr``,
  ``csr()``, ``get``/``contains``/``row_nnz``/``nnz``.
* :class:`ParallelMatrix` mirrors ``spira::parallel::parallel_matrix``
  (``include/spira/parallel/parallel_matrix.hpp``): same lifecycle, uniform row
  partitions; the device holds one concatenated CSR and ``partition_at(t)``
  returns the partition's rebased slice.
* :func:`spmv` mirrors ``serial::algorithms::spmv`` / ``parallel::algorithms::spmv``
  (``serial/spmv.hpp:142-186``, ``parallel/algorithms/spmv.hpp:127-161``).

Exceptions mirror the reference's: ``InvalidArgument`` (std::invalid_argument),
``LogicError`` (std::logic_error), ``OutOfRange`` (std::out_of_range).

There is no CPU fallback: importing this package fails if the HIP library has
not been built, and every compute call runs the HIP kernels.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = [
    "Device", "DeviceCSR", "Matrix", "ParallelMatrix", "spmv", "lib", "lib_path",
    "InvalidArgument", "LogicError", "OutOfRange", "UnsupportedInput", "DeviceError",
    "VARIANTS", "SPMV_AUTO", "transpose", "transpose_itself", "MatrixAddition", "spgemm",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libspira_amd.so")

# ---- constants mirrored from include/spira_amd.h ---------------------------
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_LOGIC = 2
ERR_OUT_OF_RANGE = 3
ERR_NOMEM = 4
ERR_DEVICE = 5
ERR_UNSUPPORTED = 6

X_DEVICE = 1
Y_DEVICE = 2
IN_DEVICE = 4
OUT_DEVICE = 8
NO_SYNC = 16
REPRODUCIBLE = 32
COUNT_GIVEN = 64

SPMV_AUTO = -1
VARIANTS = {
    "vec1": 0, "vec2": 1, "vec4": 2, "vec8": 3, "vec16": 4, "vec32": 5, "vec64": 6,
    "stream1k": 7, "stream2k": 8, "xslice1": 9, "xslice2": 10, "xslice4": 11, "xstream": 12, "xphase": 13,
    "xphase_paced": 14,
}
VARIANT_NAMES = {v: k for k, v in VARIANTS.items()}
STORAGE = {"f32": 1, "bf16": 2, "split_bf16": 3}
MIXED_VARIANTS = [VARIANTS[k] for k in ("vec4", "vec8", "vec16", "vec32", "stream1k", "stream2k", "xphase",
                                         "xphase_paced")]
ROW_MAJOR_MIXED_VARIANTS = MIXED_VARIANTS[:6]  # the ones that can run the FP64 recompute
N_VARIANTS = 15
BATCH_XLDS_FLOATS = 36864
BATCH_XLDS_TWO_PASS_COLS = 65536
BATCH_MAX_PLOG2 = 5
BATCH_STREAM = 1
NOALLOC_SCRATCH_BYTES = 32 << 20
BATCH_SCHED_DTYPE = np.dtype([("plog2", "u1"), ("xlds", "u1"), ("flags", "<u2"), ("item_nnz", "<u4")])
TRANSPOSE_DROP_ZEROS = 1
ADD_SUBTRACT = 1


class SpiraError(Exception):
    pass


class InvalidArgument(SpiraError, ValueError):
    """std::invalid_argument"""


class LogicError(SpiraError, RuntimeError):
    """std::logic_error"""


class OutOfRange(SpiraError, IndexError):
    """std::out_of_range"""


class UnsupportedInput(SpiraError, ValueError):
    """Input outside this build's limits (SPIRA_ERR_UNSUPPORTED)."""


class DeviceError(SpiraError, RuntimeError):
    """HIP runtime failure or device allocation failure."""


_ERR = {
    ERR_INVALID_ARGUMENT: InvalidArgument,
    ERR_LOGIC: LogicError,
    ERR_OUT_OF_RANGE: OutOfRange,
    ERR_NOMEM: MemoryError,
    ERR_DEVICE: DeviceError,
    ERR_UNSUPPORTED: UnsupportedInput,
}

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

_lib = None

# Live handles, released in order (matrices, then devices) before the
# interpreter tears down the HIP runtime; afterwards finalizers do nothing.
import atexit
import weakref

_live_csr = weakref.WeakSet()
_live_dev = weakref.WeakSet()
_shutdown = False


@atexit.register
def _release_all():
    global _shutdown
    for B in (list(globals().get("_live_batch", ())) + list(globals().get("_live_mlp", ()))
              + list(globals().get("_live_stage", ()))):
        B.close()
    for A in list(_live_csr):
        A.close()
    for d in list(_live_dev):
        d.close()
    _shutdown = True


def lib_path() -> str:
    return _LIB_PATH


def lib():
    """Load libspira_amd.so; fail loudly when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ImportError(
                f"{_LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        # torch ships its own HIP runtime: load it first so that the library
        # binds to the same one (one runtime per process, whichever of the two
        # touches the GPU first)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "spira_dev_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _vp]),
            "spira_dev_destroy": (None, [_vp]),
            "spira_dev_last_error": (ctypes.c_char_p, [_vp]),
            "spira_dev_stream": (_vp, [_vp]),
            "spira_dev_synchronize": (ctypes.c_int, [_vp]),
            "spira_dcsr_create": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(_vp)]),
            "spira_dcsr_destroy": (None, [_vp]),
            "spira_dcsr_info": (ctypes.c_int, [_vp, _u64p, _u64p, _u64p, ctypes.POINTER(ctypes.c_int)]),
            "spira_dcsr_upload_f64": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp, _vp, _vp, ctypes.c_uint32]),
            "spira_dcsr_download_f64": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_uint32]),
            "spira_lock_f64": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_uint32, _u64p]),
            "spira_spmv_f64": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64, ctypes.c_int,
                                              ctypes.c_uint32]),
            "spira_spmv_f32": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64, ctypes.c_int,
                                              ctypes.c_uint32]),
            "spira_spmv_f64_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64,
                                                      ctypes.c_int, ctypes.c_uint32]),
            "spira_spmv_choose": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
            "spira_spmv_set_variant": (ctypes.c_int, [_vp, ctypes.c_int]),
            "spira_partition_boundaries": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp]),
            "spira_transpose_f64": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32]),
            "spira_add_f64": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint32]),
            "spira_spgemm_f64": (ctypes.c_int, [_vp, _vp, _vp]),
            "spira_stage_create": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(_vp)]),
            "spira_stage_destroy": (None, [_vp]),
            "spira_stage_append": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_uint64, ctypes.c_uint32]),
            "spira_stage_size": (ctypes.c_int, [_vp, _u64p]),
            "spira_stage_clear": (ctypes.c_int, [_vp]),
            "spira_lock_staged_f64": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint32, _u64p]),
            "spira_scan_u64": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint64]),
            "spira_probe_ceilings": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double),
                                                    ctypes.POINTER(ctypes.c_double)]),
            "spira_spmv_autotune_f64": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                                       ctypes.POINTER(ctypes.c_int)]),
            "spira_spmv_mixed_f64": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64,
                                                    ctypes.c_int, ctypes.c_float, ctypes.c_uint32, _u64p]),
            "spira_batch_create": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(_vp), ctypes.c_int,
                                                  ctypes.POINTER(_vp)]),
            "spira_batch_destroy": (None, [_vp]),
            "spira_batch_info": (ctypes.c_int, [_vp, _u64p, _u64p, _u64p, _u32p]),
            "spira_batch_bases": (ctypes.c_int, [_vp, _vp, _vp]),
            "spira_batch_set_schedule": (ctypes.c_int, [_vp, ctypes.c_uint32, _vp]),
            "spira_batch_profile": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
            "spira_batch_spmm": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint32]),
            "spira_batch_graph_capture": (ctypes.c_int, [_vp, _vp, _vp]),
            "spira_batch_graph_launch": (ctypes.c_int, [_vp, ctypes.c_uint32]),
            "spira_matrix_features": (ctypes.c_int, [_vp, _vp]),
            "spira_mlp_create": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                ctypes.c_uint64, ctypes.POINTER(_vp)]),
            "spira_mlp_destroy": (None, [_vp]),
            "spira_mlp_params": (ctypes.c_int, [_vp, _vp, _vp, _u64p]),
            "spira_mlp_forward": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp]),
            "spira_mlp_loss_grad": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_uint32, ctypes.c_float, _vp, _vp]),
            "spira_mlp_adam_step": (ctypes.c_int, [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                                   ctypes.c_float]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a) -> int:
    """Address of a numpy array or a torch tensor (or an int address)."""
    if a is None:
        return 0
    if isinstance(a, int):
        return a
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())


def _producers_done(dev, *arrays):
    """Device inputs are read on the library's stream: unless torch's current
    stream (where the given tensors were produced) is that stream, wait for
    it first."""
    if any(hasattr(a, "is_cuda") and a.is_cuda for a in arrays):
        import torch
        cur = torch.cuda.current_stream()
        if cur.cuda_stream != dev.stream:
            cur.synchronize()


def check_device_vector(t, n: int, name: str, ordinal: int, dtype: str = "float64", what: str = "spmv"):
    """The C entry points take bare device pointers, so a device vector is
    checked here before any HIP call: a contiguous CUDA tensor of `dtype` on
    GPU `ordinal` with exactly n elements (n = -1: at least one check of the
    element count is left to the caller).  Raises InvalidArgument otherwise."""
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise InvalidArgument(f"{what}: {name} must be a {dtype} device tensor (got {type(t).__name__}"
                              + ("" if not isinstance(t, torch.Tensor) else f" on {t.device}") + ")")
    if t.dtype != getattr(torch, dtype):
        raise InvalidArgument(f"{what}: {name} must be {dtype}, got {str(t.dtype).replace('torch.', '')}")
    if not t.is_contiguous():
        raise InvalidArgument(f"{what}: {name} must be contiguous")
    if t.device.index != ordinal:
        raise InvalidArgument(f"{what}: {name} is on cuda:{t.device.index}, the handle on cuda:{ordinal}")
    if n >= 0 and t.numel() != n:
        raise InvalidArgument(f"{what}: {name} holds {t.numel()} elements, needs {n}")


def check_host_vector(a, n: int, name: str, dtype=np.float64, what: str = "spmv"):
    """A host output vector the library writes through a bare pointer: a
    contiguous numpy array of `dtype` with exactly n elements."""
    if not (isinstance(a, np.ndarray) and a.dtype == dtype and a.flags.c_contiguous):
        raise InvalidArgument(f"{what}: {name} must be a contiguous {np.dtype(dtype).name} array")
    if a.size != n:
        raise InvalidArgument(f"{what}: {name} holds {a.size} elements, needs {n}")


class Device:
    """spira_dev: one GPU and one HIP stream (``stream`` = raw hipStream_t or None)."""

    def __init__(self, ordinal: int = 0, stream: int | None = None):
        self.ordinal = int(ordinal)
        h = _vp()
        rc = lib().spira_dev_create(ctypes.byref(h), ordinal, stream or None)
        if rc != OK:
            raise _ERR.get(rc, DeviceError)(lib().spira_dev_last_error(None).decode())
        self._h = h
        _live_dev.add(self)

    @property
    def handle(self):
        return self._h

    def check(self, rc: int):
        if rc != OK:
            msg = lib().spira_dev_last_error(self._h).decode()
            raise _ERR.get(rc, DeviceError)(msg)

    def synchronize(self):
        self.check(lib().spira_dev_synchronize(self._h))

    @property
    def stream(self) -> int:
        return lib().spira_dev_stream(self._h) or 0

    def scan_u64(self, src, n=None, dst=None):
        """Exclusive prefix sum of the first n words of a device int64 tensor
        (spira_scan_u64): dst[:n] = prefixes, dst[n] = total.  n defaults to
        src.numel(); dst (n + 1 words, may be src) defaults to a new tensor.
        Returns dst."""
        import torch
        n = src.numel() if n is None else int(n)
        if dst is None:
            dst = torch.empty(n + 1, dtype=torch.int64, device=src.device)
        check_device_vector(src, -1, "src", self.ordinal, dtype="int64", what="scan")
        check_device_vector(dst, -1, "dst", self.ordinal, dtype="int64", what="scan")
        if src.numel() < n or dst.numel() < n + 1:
            raise InvalidArgument(f"scan: src needs {n} words and dst {n + 1}")
        _producers_done(self, src, dst)
        self.check(lib().spira_scan_u64(self._h, _ptr(src), _ptr(dst), n))
        return dst

    def probe_ceilings(self):
        """Measured ceilings of this GPU (spira_probe_ceilings): L2-resident
        random 8-byte gathers in requests/s and a 1 GiB streaming read in
        bytes/s."""
        g, b = ctypes.c_double(), ctypes.c_double()
        self.check(lib().spira_probe_ceilings(self._h, ctypes.byref(g), ctypes.byref(b)))
        return g.value, b.value

    def close(self):
        if getattr(self, "_h", None) and not _shutdown:
            for B in list(_live_batch) + list(_live_mlp) + list(_live_stage):  # these, then matrices, go first
                if B.dev is self:
                    B.close()
            for A in list(_live_csr):
                if A.dev is self:
                    A.close()
            lib().spira_dev_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_device: Device | None = None


def default_device() -> Device:
    global _default_device
    if _default_device is None:
        _default_device = Device(0)
    return _default_device


class DeviceCSR:
    """spira_dcsr: csr_storage<soa_tag, uint32_t, double> resident in HBM."""

    def __init__(self, dev: Device, n_rows: int, n_cols: int):
        self.dev = dev
        h = _vp()
        dev.check(lib().spira_dcsr_create(dev.handle, n_rows, n_cols, ctypes.byref(h)))
        self._h = h
        _live_csr.add(self)
        self.n_rows, self.n_cols = int(n_rows), int(n_cols)

    @property
    def handle(self):
        return self._h

    def info(self):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        built = ctypes.c_int()
        self.dev.check(lib().spira_dcsr_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                             ctypes.byref(built)))
        return {"n_rows": a.value, "n_cols": b.value, "nnz": c.value, "built": bool(built.value)}

    @property
    def nnz(self) -> int:
        return self.info()["nnz"]

    @property
    def built(self) -> bool:
        return self.info()["built"]

    def upload(self, offsets, cols, vals, device: bool = False):
        """load_csr (matrix.hpp:520-605): validate and install."""
        if not device:
            offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
            cols = np.ascontiguousarray(cols, dtype=np.uint32)
            vals = np.ascontiguousarray(vals, dtype=np.float64)
            if offsets.size != self.n_rows + 1:
                raise InvalidArgument("spira::matrix::load_csr: CSR row count does not match matrix dimensions")
            nnz = int(cols.size)
            if vals.size != nnz:
                raise InvalidArgument("load_csr: cols and vals differ in length")
        else:
            nnz = int(cols.numel())
            _producers_done(self.dev, offsets, cols, vals)
        self.dev.check(lib().spira_dcsr_upload_f64(self._h, nnz, _ptr(offsets), _ptr(cols), _ptr(vals),
                                                   IN_DEVICE if device else 0))

    def download(self):
        """csr() read-out: (offsets u64[n+1], cols u32[nnz], vals f64[nnz])."""
        nnz = self.nnz
        off = np.empty(self.n_rows + 1, np.uint64)
        col = np.empty(max(nnz, 1), np.uint32)
        val = np.empty(max(nnz, 1), np.float64)
        self.dev.check(lib().spira_dcsr_download_f64(self._h, _ptr(off), _ptr(col), _ptr(val), 0))
        return off, col[:nnz], val[:nnz]

    def lock(self, stage_ptr, stage_col, stage_val, dirty=None, device: bool = False) -> int:
        """Device lock(): build (first call) or merge the staged inserts."""
        out = ctypes.c_uint64()
        if not device:
            stage_ptr = np.ascontiguousarray(stage_ptr, dtype=np.uint64)
            stage_col = np.ascontiguousarray(stage_col, dtype=np.uint32)
            stage_val = np.ascontiguousarray(stage_val, dtype=np.float64)
            if stage_ptr.size != self.n_rows + 1:
                raise InvalidArgument("lock: stage_ptr must have n_rows+1 entries")
            if dirty is not None:
                dirty = np.ascontiguousarray(dirty, dtype=np.uint8)
                if dirty.size != self.n_rows:
                    raise InvalidArgument("lock: dirty must have n_rows entries")
            flags = 0
        else:
            _producers_done(self.dev, stage_ptr, stage_col, stage_val, dirty)
            flags = IN_DEVICE
            if hasattr(stage_col, "numel") and hasattr(stage_val, "numel"):
                # the staged count is the arrays' length: no device read of stage_ptr[n]
                if stage_col.numel() != stage_val.numel():
                    raise InvalidArgument("lock: stage_col and stage_val differ in length")
                out.value = int(stage_col.numel())
                flags |= COUNT_GIVEN
        self.dev.check(lib().spira_lock_f64(self._h, _ptr(stage_ptr), _ptr(stage_col), _ptr(stage_val),
                                            _ptr(dirty), flags, ctypes.byref(out)))
        return int(out.value)

    def spmv_f32(self, x, y=None, variant: int = SPMV_AUTO, device: bool = False, sync: bool = True):
        """The float instantiation's y = A*x: FP32 x and y (host numpy float32
        arrays, or float32 device tensors with device=True), A's values in FP32."""
        if not device:
            x = np.ascontiguousarray(x, dtype=np.float32)
            if y is None:
                y = np.empty(self.n_rows, np.float32)
            if x.size != self.n_cols:
                raise InvalidArgument("The size of the input vector x does not match the number of columns of the matrix")
            if not (isinstance(y, np.ndarray) and y.size == self.n_rows):
                raise InvalidArgument("The size of the output vector y does not match the number of rows of the matrix")
            check_host_vector(y, self.n_rows, "y", dtype=np.float32, what="spmv (float)")
            flags = 0
        else:
            import torch
            if isinstance(x, torch.Tensor) and x.numel() != self.n_cols:
                raise InvalidArgument("The size of the input vector x does not match the number of columns of the matrix")
            if isinstance(y, torch.Tensor) and y.numel() != self.n_rows:
                raise InvalidArgument("The size of the output vector y does not match the number of rows of the matrix")
            check_device_vector(x, self.n_cols, "x", self.dev.ordinal, dtype="float32", what="spmv (float)")
            check_device_vector(y, self.n_rows, "y", self.dev.ordinal, dtype="float32", what="spmv (float)")
            _producers_done(self.dev, x, y)
            flags = X_DEVICE | Y_DEVICE | (0 if sync else NO_SYNC)
        self.dev.check(lib().spira_spmv_f32(self._h, _ptr(x), self.n_cols, _ptr(y), self.n_rows, int(variant), flags))
        return y

    def _vectors(self, x, y, device, what="spmv"):
        """Checked (x, y, x_len, y_len, flags) for an FP64 SpMV call.  Sizes
        must match the matrix (the reference's checks, in its order); device
        vectors must be float64 CUDA tensors on this handle's GPU, and the
        library's stream waits for torch's current stream first."""
        if not device:
            x = np.ascontiguousarray(x, dtype=np.float64)
            if y is None:
                y = np.empty(self.n_rows, np.float64)
            if x.size != self.n_cols:
                raise InvalidArgument("The size of the input vector x does not match the number of columns of the matrix")
            if not (isinstance(y, np.ndarray) and y.size == self.n_rows):
                raise InvalidArgument("The size of the output vector y does not match the number of rows of the matrix")
            check_host_vector(y, self.n_rows, "y", what=what)
            return x, y, x.size, y.size, 0
        import torch
        if isinstance(x, torch.Tensor) and x.numel() != self.n_cols:
            raise InvalidArgument("The size of the input vector x does not match the number of columns of the matrix")
        if isinstance(y, torch.Tensor) and y.numel() != self.n_rows:
            raise InvalidArgument("The size of the output vector y does not match the number of rows of the matrix")
        check_device_vector(x, self.n_cols, "x", self.dev.ordinal, what=what)
        check_device_vector(y, self.n_rows, "y", self.dev.ordinal, what=what)
        _producers_done(self.dev, x, y)
        return x, y, x.numel(), y.numel(), X_DEVICE | Y_DEVICE

    def spmv(self, x, y=None, variant: int = SPMV_AUTO, device: bool = False, sync: bool = True,
             reproducible: bool = False):
        """y = A*x.  Host numpy arrays, or device tensors with device=True.
        reproducible: bit-identical y on every call (AUTO keeps to variants
        whose summation order is fixed; an XPHASE variant is refused)."""
        x, y, xl, yl, flags = self._vectors(x, y, device)
        if device and not sync:
            flags |= NO_SYNC
        if reproducible:
            flags |= REPRODUCIBLE
        self.dev.check(lib().spira_spmv_f64(self._h, _ptr(x), xl, _ptr(y), yl, int(variant), flags))
        return y

    def spmv_noalloc(self, x, y, variant: int = SPMV_AUTO, sync: bool = True, reproducible: bool = False):
        """y = A*x through spira_spmv_f64_noalloc: device tensors only, no
        allocation (the plan for the variant must exist: one spmv() call with
        the variant after the last lock builds it)."""
        x, y, xl, yl, flags = self._vectors(x, y, True)
        if not sync:
            flags |= NO_SYNC
        if reproducible:
            flags |= REPRODUCIBLE
        self.dev.check(lib().spira_spmv_f64_noalloc(self._h, _ptr(x), xl, _ptr(y), yl, int(variant), flags))
        return y

    def spmv_mixed(self, x, y=None, storage: str = "f32", variant: int = SPMV_AUTO, recompute_ratio: float = 0.0,
                   device: bool = False, sync: bool = True, count: bool = True):
        """Mixed-precision y = A*x (values stored as f32 / bf16 / split_bf16,
        compensated FP32 accumulation, FP64 recompute of rows whose
        cancellation ratio exceeds recompute_ratio).  Returns (y, n_recomputed)."""
        st = STORAGE[storage]
        x, y, xl, yl, flags = self._vectors(x, y, device, what="mixed spmv")
        if device and not sync:
            flags |= NO_SYNC
        nrec = ctypes.c_uint64()
        self.dev.check(lib().spira_spmv_mixed_f64(self._h, st, _ptr(x), xl, _ptr(y), yl, int(variant),
                                                  float(recompute_ratio), flags,
                                                  ctypes.byref(nrec) if count else None))
        return y, (int(nrec.value) if count else None)

    def lock_staged(self, stage: "Stage", dirty=None, device: bool = False) -> int:
        """lock() on a device Stage's inserts (sorted by row on the device);
        the stage is emptied.  dirty: optional u8[n_rows] (host, or device
        with device=True)."""
        out = ctypes.c_uint64()
        if dirty is not None and not device:
            dirty = np.ascontiguousarray(dirty, dtype=np.uint8)
            if dirty.size != self.n_rows:
                raise InvalidArgument("lock: dirty must have n_rows entries")
        self.dev.check(lib().spira_lock_staged_f64(self._h, stage.handle, _ptr(dirty), IN_DEVICE if device else 0,
                                                   ctypes.byref(out)))
        return int(out.value)

    def partition_boundaries(self, n_threads: int) -> np.ndarray:
        """compute_partition_boundaries (partition.hpp:78-126) on the device."""
        b = np.zeros(int(n_threads) + 1, np.uint64)
        self.dev.check(lib().spira_partition_boundaries(self._h, int(n_threads), _ptr(b)))
        return b

    def transpose(self, out: "DeviceCSR" = None, drop_zeros: bool = False) -> "DeviceCSR":
        """A^T on the device (serial/transpose.hpp:35-157; drop_zeros: the
        parallel form, parallel/algorithms/transpose.hpp:37-69).  out: a handle
        of the swapped shape (default: a new one); self for square in place."""
        if out is None:
            out = DeviceCSR(self.dev, self.n_cols, self.n_rows)
        self.dev.check(lib().spira_transpose_f64(self._h, out._h, TRANSPOSE_DROP_ZEROS if drop_zeros else 0))
        return out

    def add(self, B: "DeviceCSR", out: "DeviceCSR" = None, subtract: bool = False) -> "DeviceCSR":
        """C = self + B (subtract: self - B) on the device
        (serial/matrix_addition.hpp:88-112).  out: a handle of the same shape
        (default: a new one); may be self or B."""
        if out is None:
            out = DeviceCSR(self.dev, self.n_rows, self.n_cols)
        self.dev.check(lib().spira_add_f64(self._h, B._h, out._h, ADD_SUBTRACT if subtract else 0))
        return out

    def spgemm(self, B: "DeviceCSR", out: "DeviceCSR" = None) -> "DeviceCSR":
        """C = self * B on the device (serial/spgemm.hpp:12-50).  out: a handle
        with self's rows and B's columns (default: a new one)."""
        if out is None:
            out = DeviceCSR(self.dev, self.n_rows, B.n_cols)
        self.dev.check(lib().spira_spgemm_f64(self._h, B._h, out._h))
        return out

    def choose_variant(self) -> int:
        v = ctypes.c_int()
        self.dev.check(lib().spira_spmv_choose(self._h, ctypes.byref(v)))
        return v.value

    def set_variant(self, variant: int):
        self.dev.check(lib().spira_spmv_set_variant(self._h, int(variant)))

    def autotune(self, x_dev, y_dev, reps: int = 5):
        """Exhaustive per-matrix autotune on device vectors; pins the fastest.
        Returns (best_variant, {variant: ms per call})."""
        ms = (ctypes.c_float * N_VARIANTS)()
        best = ctypes.c_int()
        self.dev.check(lib().spira_spmv_autotune_f64(self._h, _ptr(x_dev), _ptr(y_dev), int(reps), ms,
                                                     ctypes.byref(best)))
        return best.value, {v: float(ms[v]) for v in range(N_VARIANTS)}

    def close(self):
        if getattr(self, "_h", None) and not _shutdown:
            lib().spira_dcsr_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_live_stage = weakref.WeakSet()


class Stage:
    """spira_stage: open-mode inserts as a device COO buffer in insertion order
    (matrix::insert / parallel_matrix::insert staging), consumed by
    DeviceCSR.lock_staged()."""

    def __init__(self, dev: Device, n_rows: int, n_cols: int):
        self.dev = dev
        h = _vp()
        dev.check(lib().spira_stage_create(dev.handle, int(n_rows), int(n_cols), ctypes.byref(h)))
        self._h = h
        self.n_rows, self.n_cols = int(n_rows), int(n_cols)
        _live_stage.add(self)

    @property
    def handle(self):
        return self._h

    def append(self, rows, cols, vals, device: bool = False):
        """Append triplets in order (host arrays, or device tensors with
        device=True: rows and cols int32/uint32, vals float64)."""
        if device:
            n = int(cols.numel())
            if int(rows.numel()) != n or int(vals.numel()) != n:
                raise InvalidArgument("stage: rows, cols and vals differ in length")
            _producers_done(self.dev, rows, cols, vals)
        else:
            rows = np.asarray(rows)
            cols = np.asarray(cols)
            if rows.size and (rows.min() < 0 or rows.max() >= self.n_rows):
                raise OutOfRange("spira::matrix: row_index out of range")
            if cols.size and (cols.min() < 0 or cols.max() >= self.n_cols):
                raise OutOfRange("spira::matrix: col_index out of range")
            rows = np.ascontiguousarray(rows, dtype=np.uint32)
            cols = np.ascontiguousarray(cols, dtype=np.uint32)
            vals = np.ascontiguousarray(vals, dtype=np.float64)
            n = int(cols.size)
            if rows.size != n or vals.size != n:
                raise InvalidArgument("stage: rows, cols and vals differ in length")
        self.dev.check(lib().spira_stage_append(self._h, _ptr(rows), _ptr(cols), _ptr(vals), n,
                                                IN_DEVICE if device else 0))

    def __len__(self):
        n = ctypes.c_uint64()
        self.dev.check(lib().spira_stage_size(self._h, ctypes.byref(n)))
        return int(n.value)

    def clear(self):
        self.dev.check(lib().spira_stage_clear(self._h))

    def close(self):
        if getattr(self, "_h", None) and not _shutdown:
            lib().spira_stage_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_live_batch = weakref.WeakSet()


class Batch:
    """spira_batch: locked matrices as one block-diagonal batch (reduced-
    precision values), k right-hand sides, one persistent launch per call."""

    def __init__(self, dev: Device, mats, storage: str = "f32"):
        self.dev = dev
        self.mats = list(mats)
        arr = (_vp * max(1, len(self.mats)))(*[m.handle for m in self.mats])
        h = _vp()
        dev.check(lib().spira_batch_create(dev.handle, len(self.mats), arr, STORAGE[storage], ctypes.byref(h)))
        self._h = h
        _live_batch.add(self)
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        t = ctypes.c_uint32()
        lib().spira_batch_info(h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(t))
        self.n_rows, self.n_cols, self.nnz = a.value, b.value, c.value
        rb = np.zeros(len(self.mats) + 1, np.uint64)
        cb = np.zeros(len(self.mats) + 1, np.uint64)
        lib().spira_batch_bases(h, _ptr(rb), _ptr(cb))
        self.row_base, self.col_base = rb, cb
        self.k = 0

    def n_tiles(self) -> int:
        t = ctypes.c_uint32()
        lib().spira_batch_info(self._h, None, None, None, ctypes.byref(t))
        return t.value

    def set_schedule(self, k: int, plog2=None, xlds=None, item_nnz=None, stream=0):
        """Per-matrix schedule (spira_batch_sched): plog2 (target nonzeros
        per lane 2^plog2), xlds (stage the X block in LDS), item_nnz (work-item
        size), stream (1: entries read from global memory, not staged in LDS).
        plog2, xlds and item_nnz all None: the library's built-in choices."""
        given = [a is not None for a in (plog2, xlds, item_nnz)]
        if any(given) and not all(given):
            raise InvalidArgument("set_schedule: give plog2, xlds and item_nnz together (or none)")
        s = None
        if all(given):
            n = len(self.mats)
            s = np.zeros(n, dtype=BATCH_SCHED_DTYPE)
            s["plog2"] = np.broadcast_to(np.asarray(plog2, np.int64), n)
            s["xlds"] = np.broadcast_to(np.asarray(xlds, np.int64), n)
            s["item_nnz"] = np.broadcast_to(np.asarray(item_nnz, np.int64), n)
            s["flags"] = np.broadcast_to(np.asarray(stream, np.int64), n) * BATCH_STREAM
        self.dev.check(lib().spira_batch_set_schedule(self._h, int(k), _ptr(s)))
        self.k = int(k)

    def _check_xy(self, X, Y):
        """The C entry points take bare device pointers to X (n_cols x k) and
        Y (n_rows x k) FP32 values, row-major, so their sizes are checked here.
        Before a schedule is set (k == 0) the library reports the mode error."""
        if self.k == 0:
            return
        import torch
        for name, T, rows in (("X", X, self.n_cols), ("Y", Y, self.n_rows)):
            check_device_vector(T, -1, name, self.dev.ordinal, dtype="float32", what="batch")
            if T.numel() < rows * self.k:
                raise InvalidArgument(f"batch: {name} holds {T.numel()} values, needs {rows} x k = {rows * self.k}")

    def spmm(self, X, Y, sync: bool = True):
        self._check_xy(X, Y)
        _producers_done(self.dev, X, Y)
        self.dev.check(lib().spira_batch_spmm(self._h, _ptr(X), _ptr(Y), 0 if sync else NO_SYNC))
        return Y

    def profile(self, X, Y) -> np.ndarray:
        """One profiled spmm: per-matrix cost in CU-microseconds."""
        self._check_xy(X, Y)
        _producers_done(self.dev, X, Y)
        out = np.zeros(max(1, len(self.mats)), np.float64)
        self.dev.check(lib().spira_batch_profile(self._h, _ptr(X), _ptr(Y), _ptr(out)))
        return out[:len(self.mats)]

    def capture(self, X, Y):
        self._check_xy(X, Y)
        _producers_done(self.dev, X, Y)
        self.dev.check(lib().spira_batch_graph_capture(self._h, _ptr(X), _ptr(Y)))

    def replay(self, sync: bool = True):
        self.dev.check(lib().spira_batch_graph_launch(self._h, 0 if sync else NO_SYNC))

    def close(self):
        if getattr(self, "_h", None) and not _shutdown:
            lib().spira_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


N_FEATURES = 18
_live_mlp = weakref.WeakSet()


def matrix_features(A: DeviceCSR) -> np.ndarray:
    """SPIRA_FEATURES floats describing a locked matrix (computed on the device)."""
    f = np.zeros(N_FEATURES, np.float32)
    A.dev.check(lib().spira_matrix_features(A.handle, _ptr(f)))
    return f


class MLP:
    """spira_mlp: n_in -> hidden -> hidden -> n_out ReLU network on the device
    (hand-written HIP forward, backward and Adam)."""

    def __init__(self, dev: Device, n_in: int, hidden: int, n_out: int, seed: int = 0):
        self.dev = dev
        h = _vp()
        dev.check(lib().spira_mlp_create(dev.handle, n_in, hidden, n_out, seed, ctypes.byref(h)))
        self._h = h
        _live_mlp.add(self)
        self.n_in, self.hidden, self.n_out = n_in, hidden, n_out
        n = ctypes.c_uint64()
        lib().spira_mlp_params(h, None, None, ctypes.byref(n))
        self.n_params = int(n.value)

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        self.dev.check(lib().spira_mlp_params(self._h, _ptr(p), None, None))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self.dev.check(lib().spira_mlp_params(self._h, None, _ptr(p), None))

    def forward(self, X_dev, Y_dev):
        """X_dev, Y_dev: float32 device tensors (batch x n_in, batch x n_out)."""
        b = int(X_dev.shape[0])
        self.dev.check(lib().spira_mlp_forward(self._h, _ptr(X_dev), b, _ptr(Y_dev)))
        return Y_dev

    def loss_grad(self, X_dev, T_dev, M_dev, mask_count: float, want_grad: bool = False):
        b = int(X_dev.shape[0])
        loss = np.zeros(1, np.float32)
        g = np.empty(self.n_params, np.float32) if want_grad else None
        self.dev.check(lib().spira_mlp_loss_grad(self._h, _ptr(X_dev), _ptr(T_dev), _ptr(M_dev), b,
                                                 float(mask_count), _ptr(loss), _ptr(g)))
        return float(loss[0]), g

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.dev.check(lib().spira_mlp_adam_step(self._h, lr, beta1, beta2, eps))

    def close(self):
        if getattr(self, "_h", None) and not _shutdown:
            lib().spira_mlp_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stage_from_coo(n_rows: int, rows, cols, vals):
    """Per-row segments in insertion order from a COO insert stream (stable)."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.uint32)
    vals = np.asarray(vals, dtype=np.float64)
    order = np.argsort(rows, kind="stable")
    counts = np.bincount(rows, minlength=n_rows) if rows.size else np.zeros(n_rows, np.int64)
    ptr = np.zeros(n_rows + 1, np.uint64)
    np.cumsum(counts, out=ptr[1:])
    return ptr, cols[order], vals[order]


class _RowBuffers:
    """Open-mode staging: the insert stream, kept in insertion order."""

    def __init__(self):
        self.clear()

    def clear(self):
        self._r, self._c, self._v = [], [], []
        self._bulk = []

    def append(self, r, c, v):
        self._r.append(r)
        self._c.append(c)
        self._v.append(v)

    def extend(self, rows, cols, vals):
        self._flush_small()
        self._bulk.append((np.asarray(rows, np.int64), np.asarray(cols, np.uint32), np.asarray(vals, np.float64)))

    def _flush_small(self):
        if self._r:
            self._bulk.append((np.asarray(self._r, np.int64), np.asarray(self._c, np.uint32),
                               np.asarray(self._v, np.float64)))
            self._r, self._c, self._v = [], [], []

    def coo(self):
        self._flush_small()
        if not self._bulk:
            return np.zeros(0, np.int64), np.zeros(0, np.uint32), np.zeros(0, np.float64)
        if len(self._bulk) > 1:
            self._bulk = [tuple(np.concatenate([b[i] for b in self._bulk]) for i in range(3))]
        return self._bulk[0]

    def __len__(self):
        return len(self._r) + sum(b[0].size for b in self._bulk)

    def lookup(self, r, c):
        """Last write to (r, c) in the buffer, or None (buffer-first read)."""
        for i in range(len(self._r) - 1, -1, -1):
            if self._r[i] == r and self._c[i] == c:
                return self._v[i]
        for rows, cols, vals in reversed(self._bulk):
            hit = np.nonzero((rows == r) & (cols == c))[0]
            if hit.size:
                return float(vals[hit[-1]])
        return None

    def row_count(self, r):
        n = sum(1 for x in self._r if x == r)
        return n + sum(int(np.count_nonzero(b[0] == r)) for b in self._bulk)

    def rows_touched(self, n_rows):
        d = np.zeros(n_rows, np.uint8)
        rows, _, _ = self.coo()
        d[rows] = 1
        return d


class _RowRef:
    """One staged row handed out by ParallelMatrix.parallel_fill."""

    __slots__ = ("_p", "_r")

    def __init__(self, part, r):
        self._p, self._r = part, r

    def insert(self, col: int, val: float):
        """row::insert (row.hpp:162-168)."""
        p = self._p
        if col < 0 or col >= p.n_cols:
            raise OutOfRange("Column index out of range")
        p.r.append(self._r)
        p.c.append(int(col))
        p.v.append(float(val))


class _PartitionRows:
    """A partition's rows for parallel_fill: rows[i] is global row row_start + i."""

    def __init__(self, row_start, row_end, n_cols):
        self.a, self.z, self.n_cols = int(row_start), int(row_end), int(n_cols)
        self.r, self.c, self.v = [], [], []

    def __len__(self):
        return self.z - self.a

    def __getitem__(self, i):
        i = int(i)
        if i < 0 or i >= self.z - self.a:
            raise IndexError("partition row index out of range")
        return _RowRef(self, self.a + i)


class Matrix:
    """spira::matrix<soa_tag, uint32_t, double> with the device backend."""

    def __init__(self, n_rows: int, n_cols: int, device: Device | None = None):
        self.dev = device or default_device()
        self._n_rows, self._n_cols = int(n_rows), int(n_cols)
        self._locked = False
        self._buf = _RowBuffers()
        self._dirty = np.zeros(self._n_rows, np.uint8)
        self._csr = DeviceCSR(self.dev, n_rows, n_cols)
        self._host = None  # cached host copy of the CSR for point queries

    # shape
    def shape(self):
        return (self._n_rows, self._n_cols)

    def n_rows(self):
        return self._n_rows

    def n_cols(self):
        return self._n_cols

    # modes
    def is_locked(self):
        return self._locked

    def is_open(self):
        return not self._locked

    def lock(self):
        """matrix.hpp:277-342."""
        if self._locked:
            return
        rows, cols, vals = self._buf.coo()
        ptr, c, v = stage_from_coo(self._n_rows, rows, cols, vals)
        dirty = self._dirty if self._csr.built else None
        self._csr.lock(ptr, c, v, dirty=dirty)
        self._buf.clear()
        self._dirty[:] = 0
        self._locked = True
        self._host = None

    def open(self):
        """matrix.hpp:348-357: the committed CSR is kept."""
        self._locked = False

    def _check_row(self, r):
        if r < 0 or r >= self._n_rows:
            raise OutOfRange("spira::matrix: row_index out of range")

    def _check_col(self, c):
        if c < 0 or c >= self._n_cols:
            raise OutOfRange("spira::matrix: col_index out of range")

    def insert(self, r: int, c: int, v: float):
        """matrix.hpp:453-462."""
        if self._locked:
            raise LogicError("matrix::insert() requires open mode")
        self._check_row(r)
        self._check_col(c)
        self._buf.append(int(r), int(c), float(v))
        self._dirty[r] = 1

    def insert_many(self, rows, cols, vals):
        """Bulk form of insert() (same checks, same insertion order)."""
        if self._locked:
            raise LogicError("matrix::insert() requires open mode")
        rows = np.asarray(rows, np.int64)
        cols = np.asarray(cols, np.int64)
        if rows.size and (rows.min() < 0 or rows.max() >= self._n_rows):
            raise OutOfRange("spira::matrix: row_index out of range")
        if cols.size and (cols.min() < 0 or cols.max() >= self._n_cols):
            raise OutOfRange("spira::matrix: col_index out of range")
        self._buf.extend(rows, cols, vals)
        self._dirty[rows] = 1

    def clear(self):
        """matrix.hpp:468-477: pending inserts dropped, CSR kept."""
        if self._locked:
            raise LogicError("matrix::clear() requires open mode")
        self._buf.clear()
        self._dirty[:] = 0

    def mark_dirty(self, r: int):
        """Effect of row_at_mut() without an insert (matrix.hpp:483-491)."""
        if self._locked:
            raise LogicError("matrix::row_at_mut() requires open mode")
        self._check_row(r)
        self._dirty[r] = 1

    # queries
    def csr(self):
        """matrix.hpp:88-91: None before the first lock, else (offsets, cols, vals)."""
        if not self._csr.built:
            return None
        if self._host is None:
            self._host = self._csr.download()
        return self._host

    def device_csr(self) -> DeviceCSR:
        return self._csr

    def _committed(self, r, c):
        h = self.csr()
        if h is None:
            return None
        off, col, val = h
        a, b = int(off[r]), int(off[r + 1])
        i = a + int(np.searchsorted(col[a:b], c))
        if i < b and int(col[i]) == c:
            return float(val[i])
        return None

    def get(self, r, c):
        self._check_row(r)
        self._check_col(c)
        if not self._locked:
            v = self._buf.lookup(r, c)
            if v is not None:
                return v
        v = self._committed(r, c)
        return 0.0 if v is None else v

    def contains(self, r, c):
        self._check_row(r)
        self._check_col(c)
        if not self._locked and self._buf.lookup(r, c) is not None:
            return True
        return self._committed(r, c) is not None

    def row_nnz(self, r):
        self._check_row(r)
        h = self.csr()
        committed = 0 if h is None else int(h[0][r + 1] - h[0][r])
        return committed + (0 if self._locked else self._buf.row_count(r))

    def nnz(self):
        committed = self._csr.nnz if self._csr.built else 0
        return committed + (0 if self._locked else len(self._buf))

    def load_csr(self, offsets, cols, vals):
        """matrix.hpp:520-605 (validation on the device)."""
        self._csr.upload(offsets, cols, vals)
        self._buf.clear()
        self._dirty[:] = 0
        self._locked = True
        self._host = None

    def __matmul__(self, x):
        """A @ B: spgemm for a matrix (matrix_operators.hpp:67-74), else y = A x."""
        if isinstance(x, Matrix):
            return spgemm(self, x)
        y = np.empty(self._n_rows, np.float64)
        spmv(self, x, y)
        return y

    def __mul__(self, other):
        """operator* of two matrices (matrix_operators.hpp:67-74)."""
        if not isinstance(other, Matrix):
            return NotImplemented
        return spgemm(self, other)

    def __imul__(self, other):
        """operator*= (matrix_operators.hpp:76-83)."""
        if not isinstance(other, Matrix):
            return NotImplemented
        return spgemm(self, other)

    def __add__(self, other):
        """operator+ (matrix_operators.hpp:21-27)."""
        return MatrixAddition(self, other)

    def __iadd__(self, other):
        """operator+= (matrix_operators.hpp:41-47)."""
        return MatrixAddition(self, other)

    def __sub__(self, other):
        """operator- (matrix_operators.hpp:29-39): the shape check first, then
        A + B * -1."""
        if self.shape() != other.shape():
            raise InvalidArgument("operator-: matrix shapes must match")
        if not other.is_locked():
            raise LogicError("multiplication_scaler: input matrix must be locked")
        return _add(self, other, subtract=True)

    def __isub__(self, other):
        """operator-= (matrix_operators.hpp:50-62)."""
        if self.shape() != other.shape():
            raise InvalidArgument("operator-=: matrix shapes must match")
        if not other.is_locked():
            raise LogicError("multiplication_scaler: input matrix must be locked")
        return _add(self, other, subtract=True)


class ParallelMatrix(Matrix):
    """parallel_matrix<soa_tag, uint32_t, double> with the device backend.

    Rows are split uniformly into n_threads partitions (parallel_matrix.hpp:296-306);
    the device keeps one concatenated CSR and partition_at(t) returns the
    partition's rebased (offsets, cols, vals).  dirty follows the parallel rule
    "row buffer non-empty" (parallel_matrix.hpp:351-355)."""

    def __init__(self, n_rows, n_cols, n_threads, device: Device | None = None):
        if n_threads < 1:
            raise InvalidArgument("parallel_matrix requires at least one thread")
        super().__init__(n_rows, n_cols, device)
        self._n_threads = int(n_threads)
        self._bounds = [t * n_rows // n_threads for t in range(n_threads + 1)]

    def n_threads(self):
        return self._n_threads

    def lock(self):
        if self._locked:
            return
        rows, cols, vals = self._buf.coo()
        ptr, c, v = stage_from_coo(self._n_rows, rows, cols, vals)
        self._csr.lock(ptr, c, v, dirty=None)
        self._buf.clear()
        self._dirty[:] = 0
        self._locked = True
        self._host = None

    def partition_bounds(self):
        return list(self._bounds)

    def parallel_fill(self, f):
        """parallel_matrix::parallel_fill (parallel_matrix.hpp:166-189):
        f(rows, row_start, row_end, t) runs once per partition on a worker
        thread; rows[global_row - row_start].insert(col, val) stages an insert
        into that partition's rows (row::insert, row.hpp:162-168, checks the
        column).  The matrix stays open; lock() then builds or merges as after
        insert().  Each partition's inserts are appended in partition order,
        so every row keeps its own insertion order.  The first exception
        raised by f is re-raised after every worker finished, and inserts
        staged before it stay staged (as in the C++ mirror)."""
        if self._locked:
            raise LogicError("parallel_fill: matrix must be open")
        from concurrent.futures import ThreadPoolExecutor

        parts = [_PartitionRows(self._bounds[t], self._bounds[t + 1], self._n_cols)
                 for t in range(self._n_threads)]

        def work(t):
            f(parts[t], self._bounds[t], self._bounds[t + 1], t)

        with ThreadPoolExecutor(max_workers=self._n_threads) as ex:
            futs = [ex.submit(work, t) for t in range(self._n_threads)]
            errs = [fu.exception() for fu in futs]
        for p in parts:
            if p.r:
                rows = np.asarray(p.r, np.int64)
                self._buf.extend(rows, np.asarray(p.c, np.uint32), np.asarray(p.v, np.float64))
                self._dirty[rows] = 1
        for e in errs:
            if e is not None:
                raise e

    def rebalance(self):
        """parallel_matrix::rebalance() (parallel_matrix.hpp:434-497): move the
        partition boundaries to the nnz-balanced ones (computed on the device);
        the device CSR is one concatenated array, so no entry moves."""
        if not self._locked:
            raise LogicError("rebalance: matrix must be locked")
        self._bounds = [int(v) for v in self._csr.partition_boundaries(self._n_threads)]

    def partition_at(self, t):
        if t < 0 or t >= self._n_threads:
            raise OutOfRange("parallel_matrix::partition_at: thread index out of range")
        h = self.csr()
        a, b = self._bounds[t], self._bounds[t + 1]
        if h is None:
            return None
        off, col, val = h
        base = int(off[a])
        return (off[a:b + 1] - np.uint64(base), col[base:int(off[b])], val[base:int(off[b])])


def transpose(mat: Matrix) -> Matrix:
    """serial::algorithms::transpose (serial/transpose.hpp:35-157) for a Matrix,
    parallel::algorithms::transpose (parallel/algorithms/transpose.hpp:37-69:
    explicit zeros dropped by its lock()) for a ParallelMatrix: a new locked
    matrix of the swapped shape, built on the device."""
    if not mat.is_locked():
        raise LogicError("transpose: matrix must be locked")
    if isinstance(mat, ParallelMatrix):
        out = ParallelMatrix(mat.n_cols(), mat.n_rows(), mat.n_threads(), mat.dev)
    else:
        out = Matrix(mat.n_cols(), mat.n_rows(), mat.dev)
    mat.device_csr().transpose(out.device_csr(), drop_zeros=isinstance(mat, ParallelMatrix))
    out._locked = True
    return out


def transpose_itself(mat: Matrix):
    """serial/transpose.hpp:158-173: an open square matrix is locked,
    transposed in place on the device and opened again."""
    if not mat.is_open():
        raise LogicError("transpose_itself: matrix must be open")
    r, c = mat.shape()
    if r != c:
        raise LogicError("in-place transpose requires square matrix")
    mat.lock()
    mat.device_csr().transpose(mat.device_csr())
    mat._host = None
    mat.open()


def _like(mat: Matrix, n_rows: int, n_cols: int) -> Matrix:
    if isinstance(mat, ParallelMatrix):
        return ParallelMatrix(n_rows, n_cols, mat.n_threads(), mat.dev)
    return Matrix(n_rows, n_cols, mat.dev)


def _add(A: Matrix, B: Matrix, subtract: bool) -> Matrix:
    if not A.is_locked():
        raise LogicError("MatrixAddition: A must be locked")
    if not B.is_locked():
        raise LogicError("MatrixAddition: B must be locked")
    C = _like(A, A.n_rows(), A.n_cols())
    A.device_csr().add(B.device_csr(), C.device_csr(), subtract=subtract)
    C._locked = True
    return C


def MatrixAddition(A: Matrix, B: Matrix) -> Matrix:
    """serial::algorithms::MatrixAddition (serial/matrix_addition.hpp:88-112)
    or, for ParallelMatrix inputs, parallel::algorithms::MatrixAddition
    (parallel/algorithms/matrix_addition.hpp:24-99; same thread count
    required): a new locked matrix, added on the device."""
    if isinstance(A, ParallelMatrix) and isinstance(B, ParallelMatrix):
        if A.shape() != B.shape():
            raise InvalidArgument("MatrixAddition: matrices must have the same shape")
        if A.n_threads() != B.n_threads():
            raise InvalidArgument("MatrixAddition: matrices must have the same thread count")
    elif A.shape() != B.shape():
        raise InvalidArgument("Matrices aren't the same size")
    return _add(A, B, subtract=False)


def spgemm(A: Matrix, B: Matrix) -> Matrix:
    """serial::algorithms::spgemm (serial/spgemm.hpp:12-50) or, for
    ParallelMatrix inputs, parallel::algorithms::spgemm
    (parallel/algorithms/spgemm.hpp:24-69; C takes A's thread count): a new
    locked matrix A * B, multiplied on the device."""
    par = isinstance(A, ParallelMatrix) and isinstance(B, ParallelMatrix)
    if A.n_cols() != B.n_rows():
        raise InvalidArgument("spgemm: A.n_cols() must equal B.n_rows()" if par else "A.cols must equal B.rows")
    if not A.is_locked():
        raise LogicError("spgemm: A must be locked")
    if not B.is_locked():
        raise LogicError("spgemm: B must be locked")
    C = _like(A, A.n_rows(), B.n_cols())
    A.device_csr().spgemm(B.device_csr(), C.device_csr())
    C._locked = True
    return C


def spmv(mat: Matrix, x, y, variant: int = SPMV_AUTO, reproducible: bool = False):
    """serial::algorithms::spmv / parallel::algorithms::spmv (FP64, device).

    x and y are numpy float64 arrays (host) or torch float64 CUDA tensors.
    reproducible: bit-identical y on every call with the same matrix and x."""
    is_dev = hasattr(x, "numel") or hasattr(y, "numel")  # torch tensors: checked as device vectors below
    if str(getattr(y, "dtype", "")) in ("float32", "torch.float32"):
        # the float instantiation (serial/spmv.hpp:90-135): FP32 x and y
        xl = int(x.numel()) if hasattr(x, "numel") else int(np.asarray(x).size)
        if xl != mat.n_cols():
            raise InvalidArgument("The size of the input vector x does not match the number of columns of the matrix")
        yl = int(y.numel()) if hasattr(y, "numel") else int(np.asarray(y).size)
        if yl != mat.n_rows():
            raise InvalidArgument("The size of the output vector y does not match the number of rows of the matrix")
        if not mat.is_locked():
            raise LogicError("spmv: matrix must be locked")
        mat.device_csr().spmv_f32(x, y, variant=variant, device=is_dev)
        return y
    xl = int(x.numel()) if hasattr(x, "numel") else int(np.asarray(x).size)
    yl = int(y.numel()) if hasattr(y, "numel") else int(np.asarray(y).size)
    if xl != mat.n_cols():
        raise InvalidArgument("The size of the input vector x does not match the number of columns of the matrix")
    if yl != mat.n_rows():
        raise InvalidArgument("The size of the output vector y does not match the number of rows of the matrix")
    if not mat.is_locked():
        raise LogicError("spmv: matrix must be locked")
    if is_dev:
        mat.device_csr().spmv(x, y, variant=variant, device=True, reproducible=reproducible)
    else:
        if not (isinstance(y, np.ndarray) and y.dtype == np.float64 and y.flags.c_contiguous):
            raise InvalidArgument("spmv: y must be a contiguous float64 array")
        mat.device_csr().spmv(np.ascontiguousarray(x, dtype=np.float64), y, variant=variant,
                              reproducible=reproducible)
    return y
