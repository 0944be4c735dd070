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
