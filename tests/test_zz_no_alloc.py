"""The allocation-free SpMV entry point (spira_spmv_f64_noalloc) through the
C-ABI: the same results as spira_spmv_f64 on the same matrix and vectors for
every variant (bit for bit where the variant's summation order is fixed,
within the FP64 bar of 1e-12 normwise for the XPHASE variants, whose LDS
atomics add in a varying order), on two calls in a row with the second on a
smaller matrix, checked against the FP64 oracle too; and its refusals: an
unprepared plan, host pointers, a variant whose scratch exceeds the fixed
array."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
import spira_amd
from spira_amd import workloads as W

pytestmark = pytest.mark.gpu
ORDER_VARIES = {spira_amd.VARIANTS["xphase"], spira_amd.VARIANTS["xphase_paced"]}


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


def _case(dev, n, seed):
    off, col, val = W.random_uniform_rows(n, n, 4, 40, seed=seed)
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(off, col, val)
    xh = np.random.default_rng(seed).uniform(-1, 1, n)
    return A, (off, col, val), xh, torch.from_numpy(xh).cuda()


@pytest.mark.parametrize("variant", [spira_amd.SPMV_AUTO] + sorted(spira_amd.VARIANTS.values()))
def test_noalloc_equals_spmv_on_two_calls(dev, variant):
    for n, seed in ((300_000, 1), (40_000, 2)):  # the second call on a smaller matrix
        print(f"noalloc: variant {variant}, n {n}", flush=True)
        A, csr, xh, x = _case(dev, n, seed)
        y_ref = torch.empty(n, dtype=torch.float64, device="cuda")
        y = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        A.spmv(x, y_ref, variant=variant, device=True)  # builds the plan
        A.spmv_noalloc(x, y, variant=variant)
        yh, yrh = y.cpu().numpy(), y_ref.cpu().numpy()
        ran = variant if variant >= 0 else A.choose_variant()  # AUTO's resolved choice
        if ran in ORDER_VARIES:
            assert oracle.normwise_rel_err(yh, yrh) <= 1e-12
        else:
            assert np.array_equal(yh.view(np.uint64), yrh.view(np.uint64))
        assert oracle.normwise_rel_err(yh, oracle.spmv_f64(*csr, xh)) <= 1e-12
        A.close()


def test_noalloc_refusals(dev):
    A, csr, xh, x = _case(dev, 20_000, 3)
    y = torch.empty(20_000, dtype=torch.float64, device="cuda")
    with pytest.raises(spira_amd.LogicError, match="no SpMV plan"):
        A.spmv_noalloc(x, y)  # AUTO before any plan
    with pytest.raises(spira_amd.LogicError, match="no SpMV plan"):
        A.spmv_noalloc(x, y, variant=spira_amd.VARIANTS["stream1k"])
    A.spmv(x, y, variant=spira_amd.VARIANTS["vec8"], device=True)
    with pytest.raises(spira_amd.LogicError, match="no SpMV plan"):
        A.spmv_noalloc(x, y, variant=spira_amd.VARIANTS["xphase"])  # plan parts of another variant
    L = spira_amd.lib()
    yh = np.empty(20_000)
    rc = L.spira_spmv_f64_noalloc(A.handle, xh.ctypes.data, 20_000, yh.ctypes.data, 20_000,
                                  spira_amd.VARIANTS["vec8"], 0)  # host pointers
    assert rc == spira_amd.ERR_INVALID_ARGUMENT
    assert b"device pointers" in L.spira_dev_last_error(dev.handle)
    with pytest.raises(spira_amd.InvalidArgument, match="columns"):
        A.spmv_noalloc(x[:-1], y)
    # XSLICE's partial sums need 8 * n_rows doubles: above 524,288 rows they
    # do not fit the fixed array
    n = 600_000
    B, csr2, xh2, x2 = _case(dev, n, 4)
    y2 = torch.empty(n, dtype=torch.float64, device="cuda")
    B.spmv(x2, y2, variant=spira_amd.VARIANTS["xslice1"], device=True)
    with pytest.raises(spira_amd.UnsupportedInput, match="SPIRA_NOALLOC_SCRATCH_BYTES"):
        B.spmv_noalloc(x2, y2, variant=spira_amd.VARIANTS["xslice1"])
    # what does fit still runs on the same matrix
    B.spmv(x2, y2, variant=spira_amd.VARIANTS["xphase_paced"], device=True)
    y3 = torch.empty_like(y2)
    B.spmv_noalloc(x2, y3, variant=spira_amd.VARIANTS["xphase_paced"])
    assert oracle.normwise_rel_err(y3.cpu().numpy(), oracle.spmv_f64(*csr2, xh2)) <= 1e-12
