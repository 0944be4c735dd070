"""Mixed-precision SpMV (FP32 / BF16 / split-BF16 value storage, compensated
FP32 accumulation, FP64 recompute of cancelling rows) against the FP64 oracle.

Tolerances, normwise relative error ||y - y_ref||_2 / ||y_ref||_2 against the
oracle's FP64 loop on the FP64 values:
  FP32 storage with recompute ratio 64:  <= 1e-6   (the target)
  split-BF16 storage:                    <= 1e-4
  BF16 storage:                          <= 2e-2
Rows recomputed in FP64 match the FP64 oracle to 1e-12 relative.
"""
import numpy as np
import pytest

import oracle
import spira_amd
from spira_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def powerlaw(dev):
    off, col, val = W.powerlaw_rows(200000, seed=5)
    A = spira_amd.DeviceCSR(dev, 200000, 200000)
    A.upload(off, col, val)
    x = np.random.default_rng(1).uniform(-1, 1, 200000)
    return A, (off, col, val), x, oracle.spmv_f64(off, col, val, x)


TOL = {"f32": 1e-6, "split_bf16": 1e-4, "bf16": 2e-2}


@pytest.mark.parametrize("storage", ["f32", "split_bf16", "bf16"])
@pytest.mark.parametrize("variant", spira_amd.ROW_MAJOR_MIXED_VARIANTS)
def test_mixed_error_bounds(powerlaw, storage, variant):
    A, csr, x, y_ref = powerlaw
    y, nrec = A.spmv_mixed(x, storage=storage, variant=variant, recompute_ratio=64.0)
    err = oracle.normwise_rel_err(y, y_ref)
    assert err <= TOL[storage], (storage, variant, err, nrec)


@pytest.mark.parametrize("storage", ["f32", "split_bf16", "bf16"])
@pytest.mark.parametrize("name", ["xphase", "xphase_paced"])
def test_mixed_xphase_layout(powerlaw, storage, name):
    """Reduced values in the XPHASE plan's order, rows summed in FP64 in LDS:
    the error is the value rounding's alone (FP32: <= 2e-6 without any
    recompute), and the recompute is refused (it needs row-major sums)."""
    A, csr, x, y_ref = powerlaw
    v = spira_amd.VARIANTS[name]
    y, nrec = A.spmv_mixed(x, storage=storage, variant=v, recompute_ratio=0.0)
    err = oracle.normwise_rel_err(y, y_ref)
    assert err <= {"f32": 2e-6, "split_bf16": 1e-4, "bf16": 2e-2}[storage], (storage, name, err)
    assert nrec == 0
    y_row, _ = A.spmv_mixed(x, storage=storage, variant=spira_amd.VARIANTS["stream2k"], recompute_ratio=0.0)
    assert err <= 1.5 * oracle.normwise_rel_err(y_row, y_ref) + 1e-12  # no worse than compensated FP32
    with pytest.raises(spira_amd.InvalidArgument, match="recompute"):
        A.spmv_mixed(x, storage=storage, variant=v, recompute_ratio=64.0)


def test_mixed_xphase_fp32_exact_values(dev):
    """FP32-exact values and x: the XPHASE layout's FP64 products (exact) and
    sums leave only FP64 rounding (<= 1e-14 normwise against the FP64 oracle)."""
    n = 300000
    off, col, val = W.random_uniform_rows(n, n, 8, 32, seed=17)
    val = val.astype(np.float32).astype(np.float64)
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(off, col, val)
    x = np.random.default_rng(2).uniform(-1, 1, n).astype(np.float32).astype(np.float64)
    y_ref = oracle.spmv_f64(off, col, val, x)
    for name in ("xphase", "xphase_paced"):
        y, _ = A.spmv_mixed(x, storage="f32", variant=spira_amd.VARIANTS[name])
        assert oracle.normwise_rel_err(y, y_ref) <= 1e-14, name


def _one_group(n_rows, n_cols, nnz, seed, long_row=0):
    """n_rows x n_cols CSR with exactly nnz distinct entries at random places
    (plus, with long_row > 0, row 0 filled with long_row leading columns), FP32-
    exact values: with 16,384 rows or fewer and at most 2^18 columns it is one
    (row block, slice) group of the XPHASE plan."""
    rng = np.random.default_rng(seed)
    keys = rng.choice(n_rows * n_cols, size=nnz, replace=False).astype(np.int64)
    if long_row:
        keys = np.union1d(keys, np.arange(long_row, dtype=np.int64))
    keys.sort()
    rows, cols = keys // n_cols, (keys % n_cols).astype(np.uint32)
    off = np.zeros(n_rows + 1, np.uint64)
    np.cumsum(np.bincount(rows, minlength=n_rows), out=off[1:])
    val = rng.standard_normal(keys.size).astype(np.float32).astype(np.float64)
    return off, cols, val


# entries per thread per step: 6 (XPHASE) and 20 (XPHASE_PACED), 1024 threads
STEP_SIZES = [1, 6143, 6144, 6145, 12288, 12289, 20479, 20480, 20481, 40960, 40961]


@pytest.mark.parametrize("nnz", STEP_SIZES)
def test_xphase_reduced_steps_at_group_boundaries(dev, nnz):
    """The reduced-precision XPHASE kernels take a group's full steps without
    branches and its remainder in one guarded step: groups of exactly one or
    two steps, and one entry either side, are exact with FP32-exact data
    (FP64 products and sums: <= 1e-14 normwise against the FP64 oracle; the
    FP32 result within FP32 rounding of it)."""
    n_rows, n_cols = 16384, 50000
    off, col, val = _one_group(n_rows, n_cols, nnz, seed=nnz)
    A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    A.upload(off, col, val)
    x = np.random.default_rng(nnz + 1).uniform(-1, 1, n_cols).astype(np.float32)
    y_ref = oracle.spmv_f64(off, col, val, x.astype(np.float64))
    for name in ("xphase", "xphase_paced"):
        v = spira_amd.VARIANTS[name]
        y, _ = A.spmv_mixed(x.astype(np.float64), storage="f32", variant=v)
        assert oracle.normwise_rel_err(y, y_ref) <= 1e-14, (name, nnz)
        y32 = A.spmv_f32(x, variant=v)
        assert np.array_equal(y32, y_ref.astype(np.float32)) or \
            oracle.normwise_rel_err(y32.astype(np.float64), y_ref) <= 1e-7, (name, nnz)


@pytest.mark.parametrize("long_row", [0, 4000])
def test_xphase_reduced_long_row_and_second_block(dev, long_row):
    """A row holding more than 1/32 of an average row block's entries makes
    the unpaced reduced kernel guard every step; both forms, and a second row
    block whose only group is a remainder, give the exact FP64 sums."""
    n_rows, n_cols = 20000, 30000  # two row blocks: 16,384 rows and 3,616 rows
    off, col, val = _one_group(n_rows, n_cols, 50000, seed=7 + long_row, long_row=long_row)
    A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    A.upload(off, col, val)
    x = np.random.default_rng(8).uniform(-1, 1, n_cols).astype(np.float32).astype(np.float64)
    y_ref = oracle.spmv_f64(off, col, val, x)
    for name in ("xphase", "xphase_paced"):
        y, _ = A.spmv_mixed(x, storage="f32", variant=spira_amd.VARIANTS[name])
        assert oracle.normwise_rel_err(y, y_ref) <= 1e-14, name
        for st in ("bf16", "split_bf16"):  # the other stored forms take the same steps
            yb, _ = A.spmv_mixed(x, storage=st, variant=spira_amd.VARIANTS[name])
            yb_row, _ = A.spmv_mixed(x, storage=st, variant=spira_amd.VARIANTS["vec8"])
            assert oracle.normwise_rel_err(yb, y_ref) <= 1.5 * oracle.normwise_rel_err(yb_row, y_ref) + 1e-12


def test_mixed_without_recompute_still_meets_fp32_bound(powerlaw):
    A, csr, x, y_ref = powerlaw
    y, nrec = A.spmv_mixed(x, storage="f32", recompute_ratio=0.0)
    assert nrec == 0
    assert oracle.normwise_rel_err(y, y_ref) <= 2e-6


def test_recomputed_rows_are_fp64(powerlaw):
    A, (off, col, val), x, y_ref = powerlaw
    v = spira_amd.VARIANTS["stream2k"]  # the same row-major kernel with and without the recompute
    y_plain, _ = A.spmv_mixed(x, storage="bf16", variant=v, recompute_ratio=0.0)
    y_rec, nrec = A.spmv_mixed(x, storage="bf16", variant=v, recompute_ratio=4.0)
    assert nrec > 0
    changed = y_rec != y_plain
    assert changed.any()
    # FP64 with a different summation order: error bounded by eps * sum|a||x|
    absum = oracle.spmv_f64(off, col, np.abs(val), np.abs(x))
    assert np.all(np.abs(y_rec[changed] - y_ref[changed]) <= 1e-14 * absum[changed])


def test_dot2_recovers_cancellation_exactly(dev):
    # 1e8 + 1 - 1e8: plain FP32 accumulation loses the 1; Dot2 keeps it.
    off = np.array([0, 3, 5], np.uint64)
    col = np.array([0, 1, 2, 0, 2], np.uint32)
    val = np.array([1e8, 1.0, -1e8, 3.0, -3.0])
    A = spira_amd.DeviceCSR(dev, 2, 3)
    A.upload(off, col, val)
    for variant in spira_amd.MIXED_VARIANTS:
        y, _ = A.spmv_mixed(np.ones(3), storage="f32", variant=variant, recompute_ratio=0.0)
        assert y.tolist() == [1.0, 0.0], variant


def test_recompute_fixes_values_not_representable_in_fp32(dev):
    # a = 1 + 2^-30 rounds to 1.0f: FP32 gives 0, the FP64 recompute gives 2^-30.
    off = np.array([0, 2], np.uint64)
    col = np.array([0, 1], np.uint32)
    val = np.array([1.0 + 2.0 ** -30, -1.0])
    A = spira_amd.DeviceCSR(dev, 1, 2)
    A.upload(off, col, val)
    y0, n0 = A.spmv_mixed(np.ones(2), storage="f32", recompute_ratio=0.0)
    assert y0[0] == 0.0 and n0 == 0
    y1, n1 = A.spmv_mixed(np.ones(2), storage="f32", recompute_ratio=1e6)
    assert y1[0] == 2.0 ** -30 and n1 == 1


def test_mixed_device_pointers_and_errors(dev, powerlaw):
    import torch
    A, csr, x, y_ref = powerlaw
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty(A.n_rows, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    A.spmv_mixed(xt, yt, storage="f32", recompute_ratio=64.0, device=True)
    assert oracle.normwise_rel_err(yt.cpu().numpy(), y_ref) <= 1e-6
    with pytest.raises(spira_amd.InvalidArgument):
        A.spmv_mixed(x[:-1], storage="f32")
    with pytest.raises(spira_amd.InvalidArgument):
        A.spmv_mixed(x, storage="f32", variant=spira_amd.VARIANTS["xstream"])
    B = spira_amd.DeviceCSR(dev, 3, 3)
    with pytest.raises(spira_amd.LogicError):
        B.spmv_mixed(np.ones(3), storage="f32")


def test_storage_copies_dropped_by_lock(dev):
    off, col, val = W.random_uniform_rows(1000, 1000, 1, 20, seed=3)
    A = spira_amd.DeviceCSR(dev, 1000, 1000)
    A.upload(off, col, val)
    x = np.random.default_rng(0).uniform(-1, 1, 1000)
    A.spmv_mixed(x, storage="f32")
    rows = np.arange(1000)
    cols = np.zeros(1000, np.uint32)
    A.lock(*spira_amd.stage_from_coo(1000, rows, cols, np.full(1000, 2.5)))
    exp = oracle.lock(*spira_amd.stage_from_coo(1000, rows, cols, np.full(1000, 2.5)), old=(off, col, val))
    y, _ = A.spmv_mixed(x, storage="f32", recompute_ratio=64.0)
    assert oracle.normwise_rel_err(y, oracle.spmv_f64(*exp, x)) <= 1e-6


def test_mixed_full_size_powerlaw(dev):
    """The full power-law workload (8,388,608 rows, Zipf(1.8) row lengths
    clipped to 50,000, 70 % of columns within +-4096 of the diagonal, values
    N(0, 1)) on the XPHASE layout for every storage, against the FP64 oracle:
    FP64 <= 1e-12, FP32 <= 2e-6 (value and x rounding alone), split BF16 <=
    1e-4, BF16 <= 2e-2; and the row-major FP32 path with the FP64 recompute of
    cancelling rows <= 1e-6."""
    import torch
    n = 1 << 23
    print(f"full-size power-law: {n} rows", flush=True)
    off_t, col_t, val_t = W.powerlaw_rows_device(n, seed=W.BASE_SEED ^ 0xDEADBEEF)
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(off_t, col_t, val_t, device=True)
    off, col, val = off_t.cpu().numpy().astype(np.uint64), col_t.cpu().numpy().astype(np.uint32), val_t.cpu().numpy()
    del off_t, col_t, val_t
    x = np.random.default_rng(4).uniform(-1, 1, n)
    y_ref = oracle.spmv_f64(off, col, val, x)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty(n, dtype=torch.float64, device="cuda")
    for name in ("xphase", "xphase_paced"):
        A.spmv(xt, yt, variant=spira_amd.VARIANTS[name], device=True)
        err = oracle.normwise_rel_err(yt.cpu().numpy(), y_ref)
        assert err <= 1e-12, (name, err)
        for storage, tol in (("f32", 2e-6), ("split_bf16", 1e-4), ("bf16", 2e-2)):
            y, nrec = A.spmv_mixed(xt, yt, storage=storage, variant=spira_amd.VARIANTS[name], device=True)
            err = oracle.normwise_rel_err(yt.cpu().numpy(), y_ref)
            print(f"{name} {storage}: {err:.3e}", flush=True)
            assert err <= tol and nrec == 0, (name, storage, err)
    y, nrec = A.spmv_mixed(xt, yt, storage="f32", variant=spira_amd.VARIANTS["stream1k"], recompute_ratio=16.0,
                           device=True)
    err = oracle.normwise_rel_err(yt.cpu().numpy(), y_ref)
    assert err <= 1e-6 and nrec > 0, (err, nrec)
    A.close()
