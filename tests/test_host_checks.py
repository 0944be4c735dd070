"""Host-side argument checks that run before any HIP call: the SpMV and batch
entry points take bare pointers, so wrong tensors must be refused in Python
(a float32 or CPU tensor would otherwise be read as FP64 device memory)."""
import numpy as np
import pytest
import torch

import spira_amd
from spira_amd import InvalidArgument, check_device_vector, check_host_vector


def test_cpu_tensor_is_refused_as_device_vector():
    with pytest.raises(InvalidArgument, match="device tensor"):
        check_device_vector(torch.zeros(8, dtype=torch.float64), 8, "x", 0)


def test_numpy_array_is_refused_as_device_vector():
    with pytest.raises(InvalidArgument, match="device tensor"):
        check_device_vector(np.zeros(8), 8, "x", 0)


def test_host_output_vector_checks():
    check_host_vector(np.zeros(5), 5, "y")
    with pytest.raises(InvalidArgument, match="float64"):
        check_host_vector(np.zeros(5, np.float32), 5, "y")
    with pytest.raises(InvalidArgument, match="float64"):
        check_host_vector(np.zeros(10)[::2], 5, "y")
    with pytest.raises(InvalidArgument, match="needs 5"):
        check_host_vector(np.zeros(6), 5, "y")


def test_top_level_spmv_refuses_cpu_tensors_before_any_call():
    class Locked:  # the size and mode checks come first, as in the reference
        def n_cols(self):
            return 4

        def n_rows(self):
            return 3

        def is_locked(self):
            return True

        def device_csr(self):
            return spira_amd.DeviceCSR.__new__(spira_amd.DeviceCSR)

    m = Locked()
    with pytest.raises(InvalidArgument, match="columns"):
        spira_amd.spmv(m, torch.zeros(5, dtype=torch.float64), torch.zeros(3, dtype=torch.float64))
    A = m.device_csr()
    A.n_rows, A.n_cols, A.dev = 3, 4, type("D", (), {"ordinal": 0})()
    with pytest.raises(InvalidArgument, match="device tensor"):
        A.spmv(torch.zeros(4, dtype=torch.float64), torch.zeros(3, dtype=torch.float64), device=True)
    with pytest.raises(InvalidArgument, match="rows"):
        A.spmv(np.zeros(4), np.zeros(2))
    with pytest.raises(InvalidArgument, match="float64"):
        A.spmv(np.zeros(4), np.zeros(3, np.float32))


def _hostonly_parallel_matrix(n_rows, n_cols, n_threads):
    """A ParallelMatrix's open-mode staging without a device handle (the
    staging path of parallel_fill makes no device call)."""
    pm = object.__new__(spira_amd.ParallelMatrix)
    pm._n_rows, pm._n_cols, pm._n_threads = n_rows, n_cols, n_threads
    pm._bounds = [t * n_rows // n_threads for t in range(n_threads + 1)]
    pm._locked = False
    pm._buf = spira_amd._RowBuffers()
    pm._dirty = np.zeros(n_rows, np.uint8)
    return pm


def test_parallel_fill_stages_each_row_in_insertion_order():
    """What lock() would stage after parallel_fill equals what it stages after
    the same triples went through insert() in order: per-row segments in
    per-row insertion order (threads run concurrently, but each owns its rows)."""
    rng = np.random.default_rng(3)
    n_rows, n_cols, T = 997, 50, 5
    pm = _hostonly_parallel_matrix(n_rows, n_cols, T)
    n = 30000
    rows = rng.integers(0, n_rows, n)
    cols = rng.integers(0, n_cols, n)
    vals = rng.standard_normal(n)
    owner = np.searchsorted(pm.partition_bounds()[1:-1], rows, side="right")

    def fill(part_rows, r0, r1, t):
        assert len(part_rows) == r1 - r0
        for i in np.nonzero(owner == t)[0]:
            part_rows[int(rows[i]) - r0].insert(int(cols[i]), float(vals[i]))

    pm.parallel_fill(fill)
    got = spira_amd.stage_from_coo(n_rows, *pm._buf.coo())
    exp = spira_amd.stage_from_coo(n_rows, rows, cols, vals)
    for a, b in zip(got, exp):
        assert np.array_equal(a, b)
    assert np.array_equal(pm._dirty, np.bincount(rows, minlength=n_rows).astype(bool).astype(np.uint8))


def test_parallel_fill_host_checks():
    pm = _hostonly_parallel_matrix(6, 4, 3)
    with pytest.raises(spira_amd.OutOfRange, match="Column index out of range"):
        pm.parallel_fill(lambda rows, a, b, t: rows[0].insert(4, 1.0))
    with pytest.raises(IndexError):
        pm.parallel_fill(lambda rows, a, b, t: rows[2].insert(0, 1.0))
    pm._locked = True
    with pytest.raises(spira_amd.LogicError, match="must be open"):
        pm.parallel_fill(lambda *a: None)
