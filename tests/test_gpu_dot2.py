"""Compensated (Dot2) accumulation is error-free up to its proven bound.

Every FP32 path that claims compensated accumulation (the batch kernel in
every X / entry mode, the single-matrix mixed-precision variants and the float
instantiation) is checked row by row on rows built to cancel: FP32 values and
FP32 x, and a last entry chosen so that the row's sum is about 2^-24 of its
sum of |a x|.  Products of two FP32 numbers are exact in FP64, so math.fsum
of the FP64 products is the exact row sum S.  Dot2 (Ogita, Rump and Oishi,
"Accurate sum and dot product", SIAM J. Sci. Comput. 26 (2005), Prop. 5.5)
returns within u |S| + gamma_n^2 sum |a x| of S (u = 2^-24, gamma_n = n u /
(1 - n u)), and the final rounding to FP32 adds at most u |S|: each row is
held to 2 u |S| + 2 gamma_n^2 sum |a x| with n = row length + 64 (the lane
merges).  Plain FP32 summation, or a Dot2 whose s + a x the compiler fused
into one FMA (which counts each product's rounding error twice), misses this
bound by orders of magnitude on these rows.
"""
import numpy as np
import pytest
import torch

import spira_amd
from dot2_cases import assert_dot2_bound, cancelling_matrix

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def batch_case(dev):
    """Three matrices: two narrow (X block in LDS in one load) and one of
    40000 columns (two column halves at k = 1)."""
    rng = np.random.default_rng(31)
    mats, parts = [], []
    for n in (3000, 9000, 40000):
        x = rng.uniform(-1, 1, n).astype(np.float32).astype(np.float64)
        (off, col, val), exact, absum, lens = cancelling_matrix(rng, n, n, x)
        A = spira_amd.DeviceCSR(dev, n, n)
        A.upload(off, col, val)
        mats.append(A)
        parts.append((x, exact, absum, lens))
    B = spira_amd.Batch(dev, mats, "f32")
    x = np.concatenate([p[0] for p in parts])
    exact = np.concatenate([p[1] for p in parts])
    absum = np.concatenate([p[2] for p in parts])
    lens = np.concatenate([p[3] for p in parts])
    n_cols = np.array([A.n_cols for A in mats])
    yield B, x, exact, absum, lens, n_cols
    B.close()
    for A in mats:
        A.close()


@pytest.mark.parametrize("k", [1, 4])
def test_batch_dot2_is_error_free(batch_case, k):
    B, x, exact, absum, lens, n_cols = batch_case
    X = torch.from_numpy(np.repeat(x.astype(np.float32)[:, None], k, axis=1).copy()).cuda()
    fits = (n_cols * k <= spira_amd.BATCH_XLDS_FLOATS) | ((k == 1) & (n_cols <= spira_amd.BATCH_XLDS_TWO_PASS_COLS))
    for xl in (0, 1):
        for stream in (0, 1):
            for plog2 in (0, 3):
                B.set_schedule(k, plog2=plog2, xlds=fits.astype(int) * xl, item_nnz=4096, stream=stream)
                Y = torch.full((B.n_rows, k), float("nan"), dtype=torch.float32, device="cuda")
                torch.cuda.synchronize()
                B.spmm(X, Y)
                y = Y.cpu().numpy().astype(np.float64)
                for j in range(k):
                    assert_dot2_bound(y[:, j], exact, absum, lens, (k, xl, stream, plog2, j))


def test_mixed_and_float_paths_dot2_is_error_free(dev):
    rng = np.random.default_rng(32)
    n = 20000
    x = rng.uniform(-1, 1, n).astype(np.float32).astype(np.float64)
    (off, col, val), exact, absum, lens = cancelling_matrix(rng, n, n, x)
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(off, col, val)
    try:
        for variant in spira_amd.ROW_MAJOR_MIXED_VARIANTS:
            y, _ = A.spmv_mixed(x, storage="f32", variant=variant, recompute_ratio=0.0)
            assert_dot2_bound(y, exact, absum, lens, ("mixed", variant))
        for variant in spira_amd.ROW_MAJOR_MIXED_VARIANTS:
            y = A.spmv_f32(x.astype(np.float32), variant=variant)
            assert_dot2_bound(np.asarray(y, np.float64), exact, absum, lens, ("f32", variant))
    finally:
        A.close()
