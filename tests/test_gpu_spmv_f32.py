"""The reference's float instantiation (serial/spmv.hpp:90-135,
parallel/algorithms/spmv.hpp:85-119) on the device: spira_spmv_f32 with FP32
x and y and the matrix's values used in FP32, through the C-ABI and the
Python mirror's spmv(); the C++ mirror's float overloads through its test
driver.

Tolerance: normwise relative error <= 1e-6 against the FP64 sum of the same
FP32 data (and no worse than twice the error of the oracle's FP32 loop, the
reference's scalar kernel dot_scalar.cpp); the known answer
(spmv_tests.cpp:121-139) exactly."""
import numpy as np
import pytest

import oracle
import spira_amd
from spira_amd import workloads as W
from oracle_model import load_kats, stage

pytestmark = pytest.mark.gpu
KATS = load_kats()
F32_VARIANTS = [spira_amd.SPMV_AUTO] + spira_amd.MIXED_VARIANTS


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("case", KATS["spmv_f32"], ids=lambda c: c["source"])
@pytest.mark.parametrize("variant", F32_VARIANTS)
def test_float_known_answer(dev, case, variant):
    m = spira_amd.Matrix(case["n_rows"], case["n_cols"], dev)
    for r, c, v in case["inserts"]:
        m.insert(r, c, float(np.float32(v)))
    m.lock()
    x = np.array(case["x"], np.float32)
    y = np.zeros(case["n_rows"], np.float32)
    m.device_csr().spmv_f32(x, y, variant=variant)
    assert y.tolist() == case["y"]
    y2 = np.zeros(case["n_rows"], np.float32)
    spira_amd.spmv(m, x, y2)  # the float overload of the mirror
    assert y2.tolist() == case["y"]


@pytest.mark.parametrize("shape", [(5000, 7000, 0, 40), (300000, 300000, 8, 32), (20000, 20000, 200, 600)])
@pytest.mark.parametrize("variant", F32_VARIANTS)
def test_float_spmv_against_the_oracle(dev, shape, variant):
    n_rows, n_cols, lo, hi = shape
    off, col, val = W.random_uniform_rows(n_rows, n_cols, lo, hi, seed=n_rows + lo)
    val32 = val.astype(np.float32)
    A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    A.upload(off, col, val32.astype(np.float64))  # a float matrix's values, exact in FP64
    x = np.random.default_rng(3).uniform(-1, 1, n_cols).astype(np.float32)
    y = A.spmv_f32(x, variant=variant)
    assert y.dtype == np.float32
    exact = oracle.spmv_f64(off, col, val32.astype(np.float64), x.astype(np.float64))
    ref32 = oracle.spmv_f32(off, col, val32, x)
    err = oracle.normwise_rel_err(y.astype(np.float64), exact)
    err_ref = oracle.normwise_rel_err(ref32.astype(np.float64), exact)
    assert err <= 1e-6, err
    assert err <= max(2 * err_ref, 1e-7), (err, err_ref)


def test_float_spmv_device_tensors_and_checks(dev):
    import torch
    n = 100000
    off, col, val = W.random_uniform_rows(n, n, 4, 24, seed=9)
    val32 = val.astype(np.float32)
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(off, col, val32.astype(np.float64))
    xh = np.random.default_rng(4).uniform(-1, 1, n).astype(np.float32)
    x = torch.from_numpy(xh).cuda()
    y = torch.empty(n, dtype=torch.float32, device="cuda")
    A.spmv_f32(x, y, device=True)
    exact = oracle.spmv_f64(off, col, val32.astype(np.float64), xh.astype(np.float64))
    assert oracle.normwise_rel_err(y.cpu().numpy().astype(np.float64), exact) <= 1e-6
    with pytest.raises(spira_amd.InvalidArgument, match="float32"):
        A.spmv_f32(x.double(), y, device=True)
    with pytest.raises(spira_amd.InvalidArgument, match="columns"):
        A.spmv_f32(xh[:-1])
    with pytest.raises(spira_amd.InvalidArgument, match="variant"):
        A.spmv_f32(xh, variant=spira_amd.VARIANTS["xslice1"])
    B = spira_amd.DeviceCSR(dev, 3, 3)
    with pytest.raises(spira_amd.LogicError):
        B.spmv_f32(np.ones(3, np.float32))
