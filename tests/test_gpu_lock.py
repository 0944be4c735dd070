"""Device lock() against the oracle, bit-exact (offsets, cols, vals), through
the C-ABI: first lock (build_csr) and re-locks (merge_csr), dirty masks,
explicit zeros from load_csr, +-0.0 deletes, NaN kept, rows with many staged
inserts, the reference's lifecycle known answers, and error codes."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import spira_amd
from spira_amd import workloads as W
from oracle_model import GOLDEN, OracleMatrix, brute_lock, csr_equal, load_kats, random_stream, run_scenario, stage

pytestmark = pytest.mark.gpu
KATS = load_kats()


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("case", KATS["lifecycle"], ids=lambda c: c["source"])
def test_lifecycle_known_answers_on_device(dev, case):
    if "threads" in case:
        mat = spira_amd.ParallelMatrix(case["n_rows"], case["n_cols"], case["threads"], dev)
    else:
        mat = spira_amd.Matrix(case["n_rows"], case["n_cols"], dev)
    run_scenario(mat, case)


@pytest.mark.parametrize("case", KATS["build_csr"], ids=lambda c: c["source"])
def test_build_csr_known_answers_on_device(dev, case):
    n = len(case["rows"])
    rows, cols, vals = [], [], []
    for r, entries in enumerate(case["rows"]):
        for c, v in entries:
            rows.append(r)
            cols.append(c)
            vals.append(v)
    A = spira_amd.DeviceCSR(dev, n, case["n_cols"])
    A.lock(*stage(n, rows, cols, vals))
    off, col, val = A.download()
    assert off.tolist() == case["offsets"]
    assert col.tolist() == case["cols"]
    assert val.tolist() == case["vals"]


@pytest.mark.parametrize("seed", range(8))
def test_lock_multicycle_random_bit_exact(dev, seed):
    rng = np.random.default_rng(1000 + seed)
    n_rows = int(rng.integers(1, 3000))
    n_cols = int(rng.integers(1, 5000))
    A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    old = None
    for cycle in range(4):
        n_ins = int(rng.integers(0, 20 * n_rows))
        rows, cols, vals = random_stream(rng, n_rows, n_cols, n_ins, zero_frac=0.15,
                                         col_pool=int(rng.integers(1, n_cols + 1)))
        ptr, c, v = stage(n_rows, rows, cols, vals)
        dirty = None
        if old is not None and cycle == 2:
            dirty = (rng.random(n_rows) < 0.3).astype(np.uint8)
            dirty[np.bincount(rows, minlength=n_rows) > 0] = 1
        nnz = A.lock(ptr, c, v, dirty=dirty)
        exp = oracle.lock(ptr, c, v, old=old, dirty=dirty)
        got = A.download()
        assert nnz == exp[1].size
        assert csr_equal(got, exp), f"seed {seed} cycle {cycle}"
        old = exp


def test_lock_matches_bruteforce_small(dev):
    rng = np.random.default_rng(77)
    for trial in range(20):
        n_rows, n_cols = int(rng.integers(1, 12)), int(rng.integers(1, 12))
        A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
        old = None
        for cycle in range(3):
            rows, cols, vals = random_stream(rng, n_rows, n_cols, int(rng.integers(0, 40)))
            ptr, c, v = stage(n_rows, rows, cols, vals)
            A.lock(ptr, c, v)
            exp = brute_lock(n_rows, rows, cols, vals, old=old)
            got = A.download()
            assert csr_equal(got, exp)
            old = got


def test_lock_nan_and_signed_zero(dev):
    rows = [0, 0, 0, 0, 1, 1]
    cols = [4, 2, 4, 7, 1, 1]
    vals = [1.0, float("nan"), -0.0, 2.0, 3.0, 0.0]
    ptr, c, v = stage(2, rows, cols, vals)
    A = spira_amd.DeviceCSR(dev, 2, 8)
    A.lock(ptr, c, v)
    assert csr_equal(A.download(), oracle.lock(ptr, c, v))


@pytest.mark.parametrize("k", [33, 64, 257, 1000, 4096, 8192])
def test_lock_rows_with_many_staged_inserts(dev, k):
    """Rows past the one-lane path (> 32 staged inserts) use the LDS sort."""
    rng = np.random.default_rng(k)
    n_rows, n_cols = 5, 3000
    rows = np.concatenate([np.full(k, 1), np.full(7, 3), np.full(k // 2 + 40, 4)])
    rng.shuffle(rows)
    cols = rng.integers(0, n_cols, rows.size)
    vals = rng.uniform(-1, 1, rows.size)
    vals[rng.random(rows.size) < 0.1] = 0.0
    ptr, c, v = stage(n_rows, rows, cols, vals)
    A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    A.lock(ptr, c, v)
    exp = oracle.lock(ptr, c, v)
    assert csr_equal(A.download(), exp)
    rows2 = np.full(k, 1)
    cols2 = rng.integers(0, n_cols, k)
    vals2 = rng.uniform(-1, 1, k)
    ptr2, c2, v2 = stage(n_rows, rows2, cols2, vals2)
    A.lock(ptr2, c2, v2)
    assert csr_equal(A.download(), oracle.lock(ptr2, c2, v2, old=exp))


def _long_row_stream(rng, n_rows, lens, col_pool):
    """Interleaved inserts: row r gets lens[r] of them, columns from
    [0, col_pool) (so a long row repeats columns: last write wins), 10 % zeros."""
    rows = np.concatenate([np.full(k, r, np.int64) for r, k in lens.items()])
    rng.shuffle(rows)
    cols = rng.integers(0, col_pool, rows.size)
    vals = rng.uniform(-1, 1, rows.size)
    vals[rng.random(rows.size) < 0.1] = 0.0
    return rows, cols, vals


@pytest.mark.parametrize("k", [8193, 50000, 1 << 20])
def test_lock_rows_longer_than_the_lds_sort(dev, k):
    """Rows with more than 8192 staged inserts (the LDS sort's size) take the
    HBM radix sort: first lock and re-lock (with a dirty mask), through both
    entry points (per-row segments and the device COO stage), bit-exact
    against the oracle, next to short and LDS-sorted rows."""
    rng = np.random.default_rng(k)
    n_rows, n_cols = 7, max(2 * k, 4096)
    print(f"long rows: k={k}", flush=True)
    rows, cols, vals = _long_row_stream(rng, n_rows, {1: k, 3: 7, 4: 9000, 5: k // 3 + 40}, k // 2 + 1)
    ptr, c, v = stage(n_rows, rows, cols, vals)
    exp = oracle.lock(ptr, c, v)
    A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    assert A.lock(ptr, c, v) == exp[1].size
    assert csr_equal(A.download(), exp)
    B = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    st = spira_amd.Stage(dev, n_rows, n_cols)
    st.append(rows, cols, vals)
    B.lock_staged(st)
    assert csr_equal(B.download(), exp)
    # re-lock: row 1 again with k inserts, row 0 new and long, row 6 short;
    # row 2 marked dirty without inserts
    rows2, cols2, vals2 = _long_row_stream(rng, n_rows, {0: k, 1: k, 6: 3}, n_cols)
    ptr2, c2, v2 = stage(n_rows, rows2, cols2, vals2)
    dirty = np.zeros(n_rows, np.uint8)
    dirty[[0, 1, 2, 6]] = 1
    exp2 = oracle.lock(ptr2, c2, v2, old=exp, dirty=dirty)
    assert A.lock(ptr2, c2, v2, dirty=dirty) == exp2[1].size
    assert csr_equal(A.download(), exp2)
    st.append(rows2, cols2, vals2)
    B.lock_staged(st, dirty=dirty)
    assert csr_equal(B.download(), exp2)
    st.close()


def test_lock_many_long_rows_device_inputs(dev):
    """Hundreds of rows past the LDS sort among short rows, device-resident
    staged arrays, first lock and re-lock, bit-exact against the oracle."""
    import torch
    rng = np.random.default_rng(51)
    n_rows, n_cols = 3000, 40000
    lens = {int(r): int(rng.integers(8193, 12000)) for r in rng.choice(n_rows, 300, replace=False)}
    for r in range(0, n_rows, 7):
        lens.setdefault(r, int(rng.integers(0, 40)))
    A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    old = None
    for cycle in range(2):
        rows, cols, vals = _long_row_stream(rng, n_rows, lens, n_cols // 2)
        ptr, c, v = stage(n_rows, rows, cols, vals)
        exp = oracle.lock(ptr, c, v, old=old)
        d = [torch.from_numpy(a).cuda() for a in (ptr.astype(np.int64), c.astype(np.int32), v)]
        assert A.lock(*d, device=True) == exp[1].size
        assert csr_equal(A.download(), exp), cycle
        old = exp


def test_powerlaw_8M_built_by_insert_and_lock(dev):
    """The power-law matrix (8,388,608 rows, Zipf(1.8) lengths clipped to
    50,000, so thousands of rows past the LDS sort) built through
    Matrix.insert_many + lock(): every row's entries inserted in reverse
    column order after a first write of 5 % of them with other values, so the
    locked CSR must equal the generator's (last write wins, sorted columns)."""
    import torch
    n = 1 << 23
    off_d, col_d, val_d = W.powerlaw_rows_device(n)
    off = off_d.cpu().numpy().astype(np.uint64)
    col = col_d.cpu().numpy().astype(np.uint32)
    val = val_d.cpu().numpy()
    del off_d, col_d, val_d
    torch.cuda.empty_cache()
    lens = np.diff(off.astype(np.int64))
    print(f"power-law: nnz {col.size}, longest row {lens.max()}, rows > 8192: {(lens > 8192).sum()}", flush=True)
    rows = np.repeat(np.arange(n, dtype=np.int64), lens)
    rev = (np.repeat(off[1:].astype(np.int64), lens) - 1) - (np.arange(col.size) - np.repeat(off[:-1].astype(np.int64), lens))
    first = np.nonzero(np.random.default_rng(5).random(col.size) < 0.05)[0]
    m = spira_amd.Matrix(n, n, dev)
    m.insert_many(rows[first], col[first], -val[first] - 1.0)  # overwritten below
    m.insert_many(rows[rev], col[rev], val[rev])
    del rows, rev, first
    m.lock()
    keep = val != 0.0  # an exact zero would be a delete marker
    exp_off = off
    if not keep.all():
        exp_off = np.zeros(n + 1, np.uint64)
        np.cumsum(np.bincount(np.repeat(np.arange(n), lens)[keep], minlength=n), out=exp_off[1:])
    assert m.nnz() == int(keep.sum())
    assert csr_equal(m.csr(), (exp_off, col[keep], val[keep]))


def test_lock_after_load_csr_with_explicit_zeros(dev):
    rng = np.random.default_rng(5)
    n, nc = 500, 400
    off, col, val = W.random_uniform_rows(n, nc, 0, 20, seed=9)
    val = val.copy()
    val[rng.random(val.size) < 0.25] = 0.0
    A = spira_amd.DeviceCSR(dev, n, nc)
    A.upload(off, col, val)
    old = (off, col, val)
    for cycle in range(3):
        rows, cols, vals = random_stream(rng, n, nc, 800, zero_frac=0.2, col_pool=60)
        ptr, c, v = stage(n, rows, cols, vals)
        dirty = (rng.random(n) < 0.2).astype(np.uint8)
        dirty[np.bincount(rows, minlength=n) > 0] = 1
        A.lock(ptr, c, v, dirty=dirty)
        exp = oracle.lock(ptr, c, v, old=old, dirty=dirty)
        assert csr_equal(A.download(), exp), cycle
        old = exp


def test_open_lock_without_inserts_is_identity(dev):
    off, col, val = W.random_uniform_rows(1000, 1000, 0, 9, seed=4)
    A = spira_amd.DeviceCSR(dev, 1000, 1000)
    A.upload(off, col, val)
    empty = np.zeros(1001, np.uint64)
    A.lock(empty, np.zeros(0, np.uint32), np.zeros(0))
    assert csr_equal(A.download(), (off, col, val))


def test_lock_empty_matrix_and_zero_rows(dev):
    A = spira_amd.DeviceCSR(dev, 4, 5)
    assert A.lock(np.zeros(5, np.uint64), np.zeros(0, np.uint32), np.zeros(0)) == 0
    off, col, val = A.download()
    assert off.tolist() == [0, 0, 0, 0, 0] and col.size == 0
    B = spira_amd.DeviceCSR(dev, 0, 3)
    assert B.lock(np.zeros(1, np.uint64), np.zeros(0, np.uint32), np.zeros(0)) == 0
    assert B.download()[0].tolist() == [0]


def test_lock_errors(dev):
    A = spira_amd.DeviceCSR(dev, 3, 4)
    ptr, c, v = stage(3, [0, 1], [1, 4], [1.0, 2.0])  # column 4 >= n_cols
    with pytest.raises(spira_amd.OutOfRange):
        A.lock(ptr, c, v)
    bad_ptr = np.array([0, 2, 1, 2], np.uint64)
    with pytest.raises(spira_amd.InvalidArgument):
        A.lock(bad_ptr, np.array([0, 1], np.uint32), np.array([1.0, 2.0]))
    # a long row (past the LDS sort) together with a bad column: still the
    # reference's exception, and the handle is untouched
    k = 9000
    ptr, c, v = stage(3, np.zeros(k + 1, np.int64), np.r_[np.zeros(k, np.uint32), 4], np.ones(k + 1))
    with pytest.raises(spira_amd.OutOfRange):
        A.lock(ptr, c, v)
    assert not A.built  # failed locks leave the handle untouched
    m = spira_amd.Matrix(2, 3, dev)
    with pytest.raises(spira_amd.OutOfRange):
        m.insert(2, 0, 1.0)
    with pytest.raises(spira_amd.OutOfRange):
        m.insert(0, 3, 1.0)
    m.lock()
    with pytest.raises(spira_amd.LogicError):
        m.insert(0, 0, 1.0)


def test_load_csr_validation_messages(dev):
    A = spira_amd.DeviceCSR(dev, 2, 4)
    cases = [
        ((np.array([1, 1, 2]), [0, 1], [1.0, 2.0]), r"offsets\[0\] must be zero"),
        ((np.array([0, 1, 3]), [0, 1], [1.0, 2.0]), r"offsets\[n_rows\] does not equal nnz"),
        ((np.array([0, 2, 1]), [0], [1.0]), "non-decreasing"),
        ((np.array([0, 1, 2]), [0, 4], [1.0, 2.0]), "out of bounds"),
        ((np.array([0, 2, 2]), [3, 3], [1.0, 2.0]), "strictly sorted"),
    ]
    for (off, col, val), msg in cases:
        with pytest.raises(spira_amd.InvalidArgument, match=msg):
            A.upload(off, col, val)
    assert not A.built


def test_relock_config3_scaled_then_spmv(dev):
    (off, col, val), (rows, cols, vals) = W.relock_workload(1 << 16)
    A = spira_amd.DeviceCSR(dev, 1 << 16, 1 << 16)
    A.upload(off, col, val)
    ptr, c, v = stage(1 << 16, rows, cols, vals)
    nnz = A.lock(ptr, c, v)
    exp = oracle.lock(ptr, c, v, old=(off, col, val))
    assert nnz == exp[1].size
    assert csr_equal(A.download(), exp)
    x = W.dense_x(1 << 16)
    y = A.spmv(x)
    assert oracle.normwise_rel_err(y, oracle.spmv_f64(*exp, x)) <= 1e-12


def test_relock_config3_full_size(dev):
    """1,048,576 rows x 16 random nnz, 2 inserts/row, 25% overwrite: bit-exact."""
    n = 1 << 20
    (off, col, val), (rows, cols, vals) = W.relock_workload(n)
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(off, col, val)
    ptr, c, v = stage(n, rows, cols, vals)
    nnz = A.lock(ptr, c, v)
    exp = oracle.lock(ptr, c, v, old=(off, col, val))
    assert nnz == exp[1].size == 16 * n + 2 * n - (2 * n) // 4
    assert csr_equal(A.download(), exp)


def test_stress_known_answer_on_device(dev, tmp_path):
    """The reference's stress test (tests/large_matrix_stress_tests.cpp:11-90)
    on the device: nnz == map size, CSR == map, y within 1e-10."""
    exe = tmp_path / "gen_stress"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", os.path.join(GOLDEN, "gen_stress.cpp"), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe), str(tmp_path)], check=True, capture_output=True, text=True)
    rd = lambda name, dt: np.fromfile(str(tmp_path / name), dt)
    m = spira_amd.Matrix(10000, 10000, dev)
    m.insert_many(rd("ins_row.u64", np.uint64).astype(np.int64), rd("ins_col.u64", np.uint64),
                  rd("ins_val.f64", np.float64))
    m.lock()
    assert m.nnz() == int(out.stdout.strip())
    assert csr_equal(m.csr(), (rd("exp_off.u64", np.uint64), rd("exp_col.u64", np.uint64),
                               rd("exp_val.f64", np.float64)))
    x, y_exp = rd("x.f64", np.float64), rd("y_exp.f64", np.float64)
    y = np.empty(10000)
    spira_amd.spmv(m, x, y)
    tol = 1e-10 + 1e-10 * np.maximum(np.abs(y), np.abs(y_exp))
    assert np.all(np.abs(y - y_exp) <= tol)


def test_parallel_matrix_partitions_are_slices(dev):
    pm = spira_amd.ParallelMatrix(10, 10, 3, dev)
    rng = np.random.default_rng(2)
    rows, cols, vals = random_stream(rng, 10, 10, 60, zero_frac=0.0)
    pm.insert_many(rows, cols, vals)
    pm.lock()
    off, col, val = pm.csr()
    assert pm.partition_bounds() == [0, 3, 6, 10]
    for t in range(3):
        po, pc, pv = pm.partition_at(t)
        a, b = pm.partition_bounds()[t:t + 2]
        assert po[0] == 0 and po.size == b - a + 1
        assert np.array_equal(pc, col[int(off[a]):int(off[b])])
    with pytest.raises(spira_amd.OutOfRange):
        pm.partition_at(3)


@pytest.mark.parametrize("case", KATS["partition_balanced"], ids=lambda c: c["source"])
def test_device_partition_boundaries_known_answers(dev, case):
    """compute_partition_boundaries on the device (from a locked CSR's
    offsets) equals the reference's known answers (partition_tests.cpp)."""
    row_nnz = case["row_nnz"]
    n = len(row_nnz)
    if n == 0:
        pytest.skip("no rows")
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(row_nnz, out=off[1:])
    width = max(max(row_nnz), 1)
    col = np.concatenate([np.arange(k, dtype=np.uint32) for k in row_nnz]) if off[-1] else np.zeros(0, np.uint32)
    A = spira_amd.DeviceCSR(dev, n, width)
    A.upload(off, col, np.ones(col.size))
    b = A.partition_boundaries(case["n_threads"]).tolist()
    if "bounds" in case:
        assert b == case["bounds"]
    for k, v in case.get("bounds_at", {}).items():
        assert b[int(k)] == v


def test_device_partition_boundaries_random_and_rebalance(dev):
    rng = np.random.default_rng(8)
    for trial in range(10):
        n = int(rng.integers(1, 5000))
        lens = rng.integers(0, 50, n) * (rng.random(n) < 0.7)
        off = np.zeros(n + 1, np.uint64)
        np.cumsum(lens, out=off[1:])
        col = np.concatenate([np.arange(k, dtype=np.uint32) for k in lens]) if off[-1] else np.zeros(0, np.uint32)
        A = spira_amd.DeviceCSR(dev, n, 64)
        A.upload(off, col, np.ones(col.size))
        for T in (1, 2, 7, 64, n + 3):
            assert np.array_equal(A.partition_boundaries(T), oracle.partition_bounds(lens, T)), (trial, T)
    pm = spira_amd.ParallelMatrix(1000, 100, 4, dev)
    with pytest.raises(spira_amd.LogicError):
        pm.rebalance()
    for r in range(1000):
        for c in range(1 + r // 20):
            pm.insert(r, c, 1.0)
    pm.lock()
    pm.rebalance()
    off = pm.csr()[0]
    assert pm.partition_bounds() == oracle.partition_bounds(np.diff(off.astype(np.int64)), 4).tolist()


def test_parallel_fill_locks_like_inserts_in_order(dev):
    """parallel_fill (parallel_matrix.hpp:166-189) over four partitions, one
    worker thread each, over several open/lock cycles with a rebalance in
    between: every cycle's CSR equals the oracle model fed the same triples
    through insert() in order (duplicates in a row keep their order, so the
    last write still wins), bit for bit."""
    n_rows, n_cols, T = 2000, 300, 4
    pm = spira_amd.ParallelMatrix(n_rows, n_cols, T, dev)
    ref = OracleMatrix(n_rows, n_cols, parallel=True)
    rng = np.random.default_rng(41)
    for cycle in range(4):
        n = int(rng.integers(0, 20000))
        rows = rng.integers(0, n_rows, n)
        cols = rng.integers(0, 40, n)
        vals = np.where(rng.random(n) < 0.1, 0.0, rng.uniform(-1, 1, n))
        bounds = pm.partition_bounds()
        owner = np.searchsorted(bounds[1:-1], rows, side="right")
        by_part = [np.nonzero(owner == t)[0] for t in range(T)]
        seen = []

        def fill(part_rows, r0, r1, t):
            seen.append((t, r0, r1, len(part_rows)))
            for i in by_part[t]:
                part_rows[int(rows[i]) - r0].insert(int(cols[i]), float(vals[i]))

        pm.parallel_fill(fill)
        assert sorted(seen) == [(t, bounds[t], bounds[t + 1], bounds[t + 1] - bounds[t]) for t in range(T)]
        for i in range(n):
            ref.insert(int(rows[i]), int(cols[i]), float(vals[i]))
        assert pm.nnz() == ref.nnz()
        pm.lock()
        ref.lock()
        assert csr_equal(pm.csr(), ref.csr()), f"cycle {cycle}"
        if cycle == 1:
            pm.rebalance()
        pm.open()
        ref.open()
    pm.lock()
    with pytest.raises(spira_amd.LogicError, match="must be open"):
        pm.parallel_fill(lambda *a: None)


def test_parallel_fill_errors(dev):
    """row::insert's column check (row.hpp:162-168) inside parallel_fill, and
    an exception from the fill function re-raised after every worker ended."""
    pm = spira_amd.ParallelMatrix(8, 5, 2, dev)

    def bad_col(rows, r0, r1, t):
        if t == 1:
            rows[0].insert(5, 1.0)

    with pytest.raises(spira_amd.OutOfRange, match="Column index out of range"):
        pm.parallel_fill(bad_col)

    def good_then_raise(rows, r0, r1, t):
        rows[0].insert(t, 2.0 + t)
        if t == 0:
            raise KeyError("from the fill function")

    with pytest.raises(KeyError):
        pm.parallel_fill(good_then_raise)
    pm.lock()  # the inserts staged before the exception are kept
    off, col, val = pm.csr()
    assert off.tolist() == [0, 1, 1, 1, 1, 2, 2, 2, 2]
    assert col.tolist() == [0, 1] and val.tolist() == [2.0, 3.0]


def test_lock_with_the_staged_count_given(dev):
    """SPIRA_COUNT_GIVEN: the caller's staged count replaces the read of
    stage_ptr[n] (device inputs).  The right count gives the same CSR as
    without it, bit for bit; a count that disagrees with stage_ptr[n] is
    refused before any column is read, and the matrix is left as it was."""
    import ctypes

    import torch
    n = 5000
    off, col, val = W.random_uniform_rows(n, n, 4, 20, seed=91)
    rng = np.random.default_rng(92)
    rows = rng.integers(0, n, 30000)
    cols = rng.integers(0, n, 30000)
    vals = rng.uniform(-1, 1, 30000)
    sp, sc, sv = spira_amd.stage_from_coo(n, rows, cols, vals)
    d_sp = torch.from_numpy(sp.astype(np.int64)).cuda()
    d_sc = torch.from_numpy(sc.astype(np.int32)).cuda()
    d_sv = torch.from_numpy(sv).cuda()
    torch.cuda.synchronize()
    ref = oracle.lock(sp, sc, sv, old=(off, col, val))
    L = spira_amd.lib()
    A = spira_amd.DeviceCSR(dev, n, n)
    for given in (False, True):
        A.upload(off, col, val)
        out = ctypes.c_uint64(sc.size if given else 0)
        flags = spira_amd.IN_DEVICE | (spira_amd.COUNT_GIVEN if given else 0)
        rc = L.spira_lock_f64(A.handle, d_sp.data_ptr(), d_sc.data_ptr(), d_sv.data_ptr(), None, flags,
                              ctypes.byref(out))
        assert rc == 0
        assert out.value == ref[1].size
        assert csr_equal(A.download(), ref)
    for wrong in (sc.size + 1, sc.size - 1, 0):
        A.upload(off, col, val)
        out = ctypes.c_uint64(wrong)
        rc = L.spira_lock_f64(A.handle, d_sp.data_ptr(), d_sc.data_ptr(), d_sv.data_ptr(), None,
                              spira_amd.IN_DEVICE | spira_amd.COUNT_GIVEN, ctypes.byref(out))
        assert rc == spira_amd.ERR_INVALID_ARGUMENT, wrong
        assert b"stage_ptr" in L.spira_dev_last_error(dev.handle)
        assert csr_equal(A.download(), (off, col, val))
    rc = L.spira_lock_f64(A.handle, d_sp.data_ptr(), d_sc.data_ptr(), d_sv.data_ptr(), None,
                          spira_amd.IN_DEVICE | spira_amd.COUNT_GIVEN, None)
    assert rc == spira_amd.ERR_INVALID_ARGUMENT
    A.close()


def test_relock_scatter_staged_and_global_blocks(dev):
    """One re-lock whose scatter blocks take both paths: a region of one-entry
    rows (a block of 2,048 positions spans more rows than it stages in LDS),
    regular rows (staged), and a row with thousands of distinct inserts (more
    deduped inserts in a block than it stages), over an old CSR with explicit
    zeros and a dirty mask wider than the inserted rows; bit-exact against the
    oracle."""
    rng = np.random.default_rng(123)
    n, nc = 20000, 50000
    lens = np.concatenate([np.ones(8000, np.int64), np.full(8000, 20, np.int64), rng.integers(0, 40, 4000)])
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    col = np.concatenate([np.sort(rng.choice(nc, int(k), replace=False)) for k in lens]).astype(np.uint32)
    val = rng.uniform(-1, 1, col.size)
    val[rng.random(col.size) < 0.2] = 0.0  # explicit zeros: only load_csr keeps them
    A = spira_amd.DeviceCSR(dev, n, nc)
    A.upload(off, col, val)
    rows = [np.arange(0, 8000, 2), np.repeat(np.arange(8000, 16000), 2), np.full(3000, 17000)]
    rows = np.concatenate(rows)
    cols = rng.integers(0, nc, rows.size)
    own = rows[(rows >= 8000) & (rows < 16000)]
    hit = rng.random(own.size) < 0.25  # a quarter of the regular rows' inserts overwrite
    sel = np.nonzero((rows >= 8000) & (rows < 16000))[0][hit]
    for k in sel:
        r = rows[k]
        cols[k] = col[int(off[r]) + int(rng.integers(0, lens[r]))]
    vals = rng.uniform(-1, 1, rows.size)
    vals[rng.random(rows.size) < 0.1] = 0.0
    ptr, c, v = stage(n, rows, cols, vals)
    dirty = (rng.random(n) < 0.1).astype(np.uint8)
    dirty[np.bincount(rows, minlength=n) > 0] = 1
    A.lock(ptr, c, v, dirty=dirty)
    exp = oracle.lock(ptr, c, v, old=(off, col, val), dirty=dirty)
    assert csr_equal(A.download(), exp)
    A.close()
