"""Batched multi-RHS SpMV (block-diagonal batch, persistent work-queue kernel,
hipGraph replay) against the FP64 oracle.

Inputs are FP32-representable (values N(0,1) rounded to FP32, X FP32), so the
oracle's FP64 product is the exact reference; the bound is normwise relative
error <= 1e-6 for FP32 storage (compensated FP32 accumulation, FP32 output).
Graph replays must equal the eager call bit for bit (each output row is
computed by one wavefront in a fixed order).
"""
import numpy as np
import pytest
import torch

import oracle
import spira_amd
from spira_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


def _make(dev, seeds, n_rows=None, extra=None):
    mats, host = [], []
    for s in seeds:
        off, col, val = W.batch_matrix(s, n_rows=n_rows, min_rows=64, max_rows=3000)
        off, col, val = off.numpy().astype(np.uint64), col.numpy().astype(np.uint32), val.numpy()
        host.append((off, col, val))
    if extra is not None:
        host.append(extra)
    for off, col, val in host:
        n = len(off) - 1
        A = spira_amd.DeviceCSR(dev, n, n)
        A.upload(off, col, val)
        mats.append(A)
    return mats, host


def _reference(host, X, k, col_base):
    ys = []
    for m, (off, col, val) in enumerate(host):
        n = len(off) - 1
        Xm = X[int(col_base[m]):int(col_base[m]) + n].astype(np.float64)
        ys.append(oracle.spmm_f64(off, col, val, Xm))
    return np.concatenate(ys)


@pytest.fixture(scope="module")
def batch(dev):
    rng = np.random.default_rng(9)
    # one matrix with a single very long row exercises the whole-block path
    n = 6000
    lens = np.ones(n, np.int64)
    lens[17] = 5000
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    cols = []
    for r in range(n):
        cols.append(np.sort(rng.choice(n, size=int(lens[r]), replace=False)))
    col = np.concatenate(cols).astype(np.uint32)
    val = rng.standard_normal(col.size).astype(np.float32).astype(np.float64)
    mats, host = _make(dev, range(100, 112), extra=(off, col, val))
    B = spira_amd.Batch(dev, mats, "f32")
    return B, host


@pytest.mark.parametrize("k", [1, 4, 16, 32])
def test_batch_spmm_matches_oracle(batch, k):
    B, host = batch
    B.set_schedule(k)
    X = np.random.default_rng(k).standard_normal((B.n_cols, k)).astype(np.float32)
    Xt = torch.from_numpy(X).cuda()
    Yt = torch.full((B.n_rows, k), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    B.spmm(Xt, Yt)
    ref = _reference(host, X, k, B.col_base)
    y = Yt.cpu().numpy().astype(np.float64)
    assert np.isfinite(y).all()
    assert oracle.normwise_rel_err(y.ravel(), ref.ravel()) <= 1e-6


def _xlds_ok(host, k):
    n = np.array([len(off) - 1 for off, _, _ in host])
    return (n * k <= spira_amd.BATCH_XLDS_FLOATS) | ((k == 1) & (n <= spira_amd.BATCH_XLDS_TWO_PASS_COLS))


def test_k1_two_pass_x_blocks(dev):
    """k = 1 matrices with 36865..65536 columns stage X in two halves: rows
    whose entries straddle the halves, rows entirely in one half, empty rows,
    long rows (whole-block tiles, read from global X) and small items."""
    rng = np.random.default_rng(21)
    host = []
    for n in (40000, 65536):
        lens = rng.integers(0, 60, size=n)
        lens[::11] = 0
        lens[5::97] = 3000  # whole-block rows
        off = np.zeros(n + 1, np.uint64)
        np.cumsum(lens, out=off[1:])
        cols = []
        for r, L in enumerate(lens):
            if r % 3 == 0:  # only in the first half
                cols.append(np.sort(rng.choice(n // 2, size=min(int(L), n // 2), replace=False)))
            elif r % 3 == 1:  # only in the second half
                cols.append(np.sort(rng.choice(n - n // 2, size=int(L), replace=False) + n // 2))
            else:
                cols.append(np.sort(rng.choice(n, size=int(L), replace=False)))
        lens = np.array([c.size for c in cols])
        np.cumsum(lens, out=off[1:])
        col = np.concatenate(cols).astype(np.uint32)
        val = rng.standard_normal(col.size).astype(np.float32).astype(np.float64)
        host.append((off, col, val))
    mats = []
    for off, col, val in host:
        A = spira_amd.DeviceCSR(dev, len(off) - 1, len(off) - 1)
        A.upload(off, col, val)
        mats.append(A)
    B = spira_amd.Batch(dev, mats, "f32")
    X = rng.standard_normal((B.n_cols, 1)).astype(np.float32)
    ref = _reference(host, X, 1, B.col_base)
    for plog2 in (0, 2, 5):
        for item in (4096, 1 << 20):
            B.set_schedule(1, plog2=plog2, xlds=1, item_nnz=item)
            Yt = torch.full((B.n_rows, 1), float("nan"), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            B.spmm(torch.from_numpy(X).cuda(), Yt)
            y = Yt.cpu().numpy().astype(np.float64)
            assert np.isfinite(y).all(), (plog2, item)
            assert oracle.normwise_rel_err(y.ravel(), ref.ravel()) <= 1e-6, (plog2, item)
    with pytest.raises(spira_amd.InvalidArgument):  # two halves are a k = 1 mode only
        B.set_schedule(4, plog2=2, xlds=1, item_nnz=4096)


@pytest.mark.parametrize("k", [1, 4, 16, 32])
def test_every_schedule_gives_the_same_answer(batch, k):
    B, host = batch
    X = np.random.default_rng(7).standard_normal((B.n_cols, k)).astype(np.float32)
    Xt = torch.from_numpy(X).cuda()
    ref = _reference(host, X, k, B.col_base)
    fits = _xlds_ok(host, k)
    for plog2 in range(spira_amd.BATCH_MAX_PLOG2 + 1):
        for xl in (0, 1):
            for item in (64, 4096, 1 << 20):
                B.set_schedule(k, plog2=plog2, xlds=fits.astype(int) * xl, item_nnz=item)
                Yt = torch.full((B.n_rows, k), float("nan"), dtype=torch.float32, device="cuda")
                torch.cuda.synchronize()
                B.spmm(Xt, Yt)
                y = Yt.cpu().numpy().astype(np.float64)
                assert np.isfinite(y).all(), (plog2, xl, item)
                assert oracle.normwise_rel_err(y.ravel(), ref.ravel()) <= 1e-6, (plog2, xl, item)


def test_wide_matrices_use_batch_columns(dev):
    """A matrix with more than 65536 columns switches the batch to u32
    columns; empty matrices and empty rows still produce zeros."""
    rng = np.random.default_rng(4)
    n = 70_000
    lens = rng.integers(0, 6, size=n)
    lens[::7] = 0
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    col = np.concatenate([np.sort(rng.choice(n, size=int(L), replace=False)) for L in lens]).astype(np.uint32)
    val = rng.standard_normal(col.size).astype(np.float32).astype(np.float64)
    empty = (np.zeros(11, np.uint64), np.zeros(0, np.uint32), np.zeros(0))
    mats, host = _make(dev, [400, 401], extra=(off, col, val))
    for o, c, v in [empty]:
        A = spira_amd.DeviceCSR(dev, 10, 10)
        A.upload(o, c, v)
        mats.append(A)
        host.append((o, c, v))
    B = spira_amd.Batch(dev, mats, "f32")
    for k in (1, 4, 32):
        X = rng.standard_normal((B.n_cols, k)).astype(np.float32)
        ref = _reference(host, X, k, B.col_base)
        for xl in (0, 1):
            B.set_schedule(k, plog2=2, xlds=_xlds_ok(host, k).astype(int) * xl, item_nnz=2048)
            Yt = torch.full((B.n_rows, k), float("nan"), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            B.spmm(torch.from_numpy(X).cuda(), Yt)
            y = Yt.cpu().numpy().astype(np.float64)
            assert np.isfinite(y).all()
            assert oracle.normwise_rel_err(y.ravel(), ref.ravel()) <= 1e-6, (k, xl)
            assert (y[-10:] == 0).all()


def test_schedule_arguments_are_checked(batch):
    B, host = batch
    n = len(host)
    with pytest.raises(spira_amd.InvalidArgument):
        B.set_schedule(1, plog2=spira_amd.BATCH_MAX_PLOG2 + 1, xlds=0, item_nnz=1024)
    with pytest.raises(spira_amd.InvalidArgument):
        B.set_schedule(32, plog2=0, xlds=np.ones(n, int), item_nnz=1024)  # 6000 x 32 floats do not fit in LDS
    with pytest.raises(spira_amd.InvalidArgument):
        B.set_schedule(1, plog2=0, xlds=0, item_nnz=0)
    with pytest.raises(spira_amd.InvalidArgument):
        B.set_schedule(1, plog2=0)


def test_graph_replay_equals_eager(batch):
    B, host = batch
    k = 16
    B.set_schedule(k)
    X = torch.from_numpy(np.random.default_rng(3).standard_normal((B.n_cols, k)).astype(np.float32)).cuda()
    Y_eager = torch.zeros((B.n_rows, k), dtype=torch.float32, device="cuda")
    Y_graph = torch.zeros((B.n_rows, k), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    B.spmm(X, Y_eager)
    B.capture(X, Y_graph)
    for rep in range(4):
        Y_graph.fill_(float("nan"))
        torch.cuda.synchronize()
        B.replay()
        assert torch.equal(Y_graph, Y_eager), rep
        # eager work between replays, on the same handle
        Y_other = torch.zeros_like(Y_eager)
        B.spmm(X, Y_other)
        assert torch.equal(Y_other, Y_eager)
        torch.cuda.synchronize()


@pytest.mark.parametrize("storage,tol", [("bf16", 2e-2), ("split_bf16", 1e-4)])
def test_batch_reduced_storage(dev, storage, tol):
    mats, host = _make(dev, range(200, 206))
    B = spira_amd.Batch(dev, mats, storage)
    B.set_schedule(4)
    X = np.random.default_rng(1).standard_normal((B.n_cols, 4)).astype(np.float32)
    Yt = torch.zeros((B.n_rows, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    B.spmm(torch.from_numpy(X).cuda(), Yt)
    ref = _reference(host, X, 4, B.col_base)
    assert oracle.normwise_rel_err(Yt.cpu().numpy().astype(np.float64).ravel(), ref.ravel()) <= tol


def test_batch_errors(dev):
    mats, host = _make(dev, [300, 301])
    B = spira_amd.Batch(dev, mats)
    X = torch.zeros((B.n_cols, 4), dtype=torch.float32, device="cuda")
    Y = torch.zeros((B.n_rows, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(spira_amd.LogicError):
        B.spmm(X, Y)
    with pytest.raises(spira_amd.InvalidArgument):
        B.set_schedule(3)
    B.set_schedule(4)
    with pytest.raises(spira_amd.InvalidArgument):  # Y sized for k = 1: rejected before any launch
        B.spmm(X, Y[:, :1].contiguous())
    with pytest.raises(spira_amd.InvalidArgument):
        B.spmm(X.double(), Y)
    unlocked = spira_amd.DeviceCSR(dev, 5, 5)
    with pytest.raises(spira_amd.LogicError):
        spira_amd.Batch(dev, [mats[0], unlocked])


def test_bench_generator_slice_with_selector_schedules(dev):
    """192 matrices of the batched workload's own generator (rows log-uniform
    in [256, 65536]: X blocks in one LDS load, in two halves, and too wide for
    LDS at k = 4) under the learned selector's schedules, k = 1 and 4, against
    the FP64 oracle at 1e-6; and the built-in schedule gives the same answer
    to 1e-6."""
    from spira_amd import selector as S
    mats, host = [], []
    for s in range(20_000_000, 20_000_192):
        off, col, val = W.batch_matrix(s, device="cuda")
        off, col, val = (off.cpu().numpy().astype(np.uint64), col.cpu().numpy().astype(np.uint32),
                         val.cpu().numpy())
        host.append((off, col, val))
        A = spira_amd.DeviceCSR(dev, len(off) - 1, len(off) - 1)
        A.upload(off, col, val)
        mats.append(A)
    assert max(len(o) - 1 for o, _, _ in host) > spira_amd.BATCH_XLDS_FLOATS  # two-pass matrices present
    B = spira_amd.Batch(dev, mats, "f32")
    sel = S.Selector.load(dev)
    n_cols = np.array([A.n_cols for A in mats])
    feats = S.features_by_k(mats)
    for k in (1, 4):
        X = np.random.default_rng(k + 40).standard_normal((B.n_cols, k)).astype(np.float32)
        Xt = torch.from_numpy(X).cuda()
        ref = _reference(host, X, k, B.col_base)
        outs = []
        for sched in ("selector", "builtin"):
            if sched == "selector":
                S.apply(B, k, sel.predict(feats[k], [k] * len(mats), n_cols), n_cols)
            else:
                B.set_schedule(k)
            Yt = torch.full((B.n_rows, k), float("nan"), dtype=torch.float32, device="cuda")
            B.spmm(Xt, Yt)
            y = Yt.cpu().numpy().astype(np.float64)
            err = oracle.normwise_rel_err(y.ravel(), ref.ravel())
            print(f"k={k} {sched}: {err:.3e}", flush=True)
            assert np.isfinite(y).all() and err <= 1e-6, (k, sched, err)
            outs.append(y)
        assert oracle.normwise_rel_err(outs[0].ravel(), outs[1].ravel()) <= 1e-6
    B.close()
    for A in mats:
        A.close()
