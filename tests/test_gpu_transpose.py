"""Device transpose against the reference's known answers and the oracle,
through the C-ABI (spira_transpose_f64) and the Python mirror.

Bar: bit-exact.  Offsets, columns and values of A^T must equal the oracle's
(serial/transpose.hpp:35-157: a column histogram and a fill in input-row
order, explicit zeros kept; with drop_zeros the parallel form,
parallel/algorithms/transpose.hpp:37-69, whose lock() removes them).  At the
full 1M-row size the oracle is checked as well, plus (A^T)^T == A.
"""
import ctypes

import numpy as np
import pytest

import oracle
import spira_amd
from spira_amd import workloads as W
from oracle_model import load_kats, stage

pytestmark = pytest.mark.gpu

KATS = load_kats()


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


def _equal(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


def _bits_equal(a, b):
    """Offsets and columns equal, values equal bit for bit (signed zeros too)."""
    return (np.array_equal(a[0].astype(np.uint64), b[0].astype(np.uint64))
            and np.array_equal(a[1].astype(np.uint32), b[1].astype(np.uint32))
            and np.array_equal(np.asarray(a[2], np.float64).view(np.uint64),
                               np.asarray(b[2], np.float64).view(np.uint64)))


def _build(dev, case):
    n_rows, n_cols = case["n_rows"], case["n_cols"]
    if case.get("mode", "serial") == "serial":
        m = spira_amd.Matrix(n_rows, n_cols, dev)
    else:
        m = spira_amd.ParallelMatrix(n_rows, n_cols, case["n_threads"], dev)
    ins = case["inserts"]
    if ins:
        m.insert_many([t[0] for t in ins], [t[1] for t in ins], [t[2] for t in ins])
    return m


def _check_case(t, case, shape):
    assert t.shape() == shape
    for r, c, v in case["expect"]:
        assert t.get(r, c) == v, (r, c)
    for r, c in case.get("expect_zero", []):
        assert t.get(r, c) == 0.0 and not t.contains(r, c)
    if "nnz" in case:
        assert t.nnz() == case["nnz"]
    off, col, _ = t.csr()
    for r in range(shape[0]):
        seg = col[int(off[r]):int(off[r + 1])].astype(np.int64)
        assert np.all(np.diff(seg) > 0)


@pytest.mark.parametrize("case", KATS["transpose"], ids=lambda c: c["source"])
def test_transpose_known_answers(dev, case):
    m = _build(dev, case)
    m.lock()
    t = spira_amd.transpose(m)
    shape = (case["n_cols"], case["n_rows"])
    assert t.is_locked() and isinstance(t, type(m))
    if case.get("twice"):
        t = spira_amd.transpose(t)
        shape = (case["n_rows"], case["n_cols"])
    _check_case(t, case, shape)
    if case["mode"] == "parallel_vs_serial":
        s = spira_amd.Matrix(case["n_rows"], case["n_cols"], dev)
        ins = case["inserts"]
        s.insert_many([x[0] for x in ins], [x[1] for x in ins], [x[2] for x in ins])
        s.lock()
        assert _equal(spira_amd.transpose(s).csr(), t.csr())
    # the oracle agrees bit for bit
    exp = oracle.transpose(*m.csr(), case["n_cols"], drop_zeros=case["mode"] != "serial")
    if not case.get("twice"):
        assert _bits_equal(t.csr(), exp)


@pytest.mark.parametrize("case", KATS["transpose_itself"], ids=lambda c: c["source"])
def test_transpose_itself_known_answers(dev, case):
    m = _build(dev, case)
    if case.get("error"):
        with pytest.raises(spira_amd.LogicError, match="square"):
            spira_amd.transpose_itself(m)
        return
    spira_amd.transpose_itself(m)
    assert m.is_open()
    if case.get("twice"):
        spira_amd.transpose_itself(m)
    _check_case(m, case, (case["n_rows"], case["n_cols"]))


def test_transpose_preconditions(dev):
    m = spira_amd.Matrix(3, 4, dev)
    with pytest.raises(spira_amd.LogicError, match="must be locked"):
        spira_amd.transpose(m)  # open
    m.lock()
    with pytest.raises(spira_amd.LogicError, match="must be open"):
        spira_amd.transpose_itself(m)
    # C-ABI checks: never-built handle, wrong output shape, unknown flags
    A = spira_amd.DeviceCSR(dev, 3, 4)
    with pytest.raises(spira_amd.LogicError):
        A.transpose()
    A.upload(np.array([0, 1, 1, 2], np.uint64), np.array([2, 0], np.uint32), np.array([1.0, 2.0]))
    with pytest.raises(spira_amd.InvalidArgument):
        A.transpose(spira_amd.DeviceCSR(dev, 3, 4))
    out = spira_amd.DeviceCSR(dev, 4, 3)
    rc = spira_amd.lib().spira_transpose_f64(A.handle, out.handle, ctypes.c_uint32(2))
    assert rc == spira_amd.ERR_INVALID_ARGUMENT
    assert not out.built


def _with_zeros(off, col, val, seed, frac=0.2):
    rng = np.random.default_rng(seed)
    val = val.copy()
    z = rng.random(val.size) < frac
    val[z] = np.where(rng.random(int(z.sum())) < 0.5, 0.0, -0.0)
    return off, col, val


def _long_row_and_column(n):
    """Row 0 holds every column, column 0 holds every row."""
    lens = np.ones(n, np.int64)
    lens[0] = n
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    col = np.zeros(int(off[-1]), np.uint32)
    col[:n] = np.arange(n, dtype=np.uint32)
    val = np.random.default_rng(n).standard_normal(col.size)
    return off, col, val


def _mostly_empty_rows(n, n_cols, frac, seed):
    """About frac of the rows hold one entry: a key block's 2048 entries span
    far more rows than its LDS row cache, so the in-memory search runs."""
    rng = np.random.default_rng(seed)
    lens = (rng.random(n) < frac).astype(np.int64)
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    col = rng.integers(0, n_cols, int(off[-1])).astype(np.uint32)
    val = rng.standard_normal(col.size)
    return off, col, val


CASES = [
    # (name, builder)
    ("empty_1x1", lambda: (np.zeros(2, np.uint64), np.zeros(0, np.uint32), np.zeros(0), 1)),
    ("all_rows_empty", lambda: (np.zeros(5001, np.uint64), np.zeros(0, np.uint32), np.zeros(0), 7000)),
    ("tall_thin", lambda: (*W.random_uniform_rows(50000, 3, 0, 3, seed=31), 3)),
    ("short_wide", lambda: (*W.random_uniform_rows(3, 200000, 0, 5000, seed=32), 200000)),
    ("ragged_empty_rows", lambda: (*W.random_uniform_rows(30000, 20000, 0, 12, seed=33), 20000)),
    ("uniform_2b_like", lambda: (*W.random_uniform_rows(20000, 70000, 8, 32, seed=34), 70000)),
    ("long_row_long_col", lambda: (*_long_row_and_column(40000), 40000)),
    ("mostly_empty_rows", lambda: (*_mostly_empty_rows(400000, 5000, 0.02, 37), 5000)),
    ("explicit_zeros", lambda: (*_with_zeros(*W.random_uniform_rows(9000, 11000, 0, 20, seed=35), 36), 11000)),
]


@pytest.mark.parametrize("drop", [False, True], ids=["serial", "drop_zeros"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_transpose_matches_oracle(dev, case, drop):
    off, col, val, n_cols = case[1]()
    n_rows = len(off) - 1
    A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    A.upload(off, col, val)
    T = A.transpose(drop_zeros=drop)
    got = T.download()
    exp = oracle.transpose(off, col, val, n_cols, drop_zeros=drop)
    assert _bits_equal(got, exp)
    assert T.nnz == len(exp[1])
    # (A^T)^T == A (zeros kept) and the source is left untouched
    if not drop:
        assert _bits_equal(T.transpose().download(), (off, col, val))
    assert _bits_equal(A.download(), (off, col, val))


def test_transpose_in_place_square_and_spmv(dev):
    n = 30000
    off, col, val = _with_zeros(*W.random_uniform_rows(n, n, 0, 24, seed=41), 42)
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(off, col, val)
    x = np.random.default_rng(43).uniform(-1, 1, n)
    A.spmv(x, variant=spira_amd.VARIANTS["xphase"])  # a plan exists before the swap
    assert A.transpose(A) is A
    exp = oracle.transpose(off, col, val, n)
    assert _bits_equal(A.download(), exp)
    for v in sorted(spira_amd.VARIANTS.values()):  # plans rebuilt for the new arrays
        y = A.spmv(x, variant=v)
        y_ref = oracle.spmv_f64(*exp, x)
        assert oracle.normwise_rel_err(y, y_ref) <= 1e-12, v
    A.transpose(A, drop_zeros=True)
    assert _bits_equal(A.download(), oracle.transpose(*exp, n, drop_zeros=True))


def test_transpose_relock_after_transpose(dev):
    """A transposed matrix takes further inserts and locks like any other."""
    m = spira_amd.Matrix(500, 300, dev)
    rows, cols, vals = W.insert_stream_uniform(500, 300, 6)
    m.insert_many(rows, cols, vals)
    m.lock()
    t = spira_amd.transpose(m)
    exp = oracle.transpose(*m.csr(), 300)
    assert _bits_equal(t.csr(), exp)
    t.open()
    t.insert(7, 3, 2.5)
    t.insert(299, 499, 0.0)  # a delete or a no-op
    t.lock()
    ptr, c, v = stage(300, [7, 299], [3, 499], [2.5, 0.0])
    dirty = np.zeros(300, np.uint8)
    dirty[[7, 299]] = 1
    assert _bits_equal(t.csr(), oracle.lock(ptr, c, v, old=exp, dirty=dirty))


def test_transpose_full_size_1m(dev):
    """1,048,576 x 1,048,576, 8-32 entries per row (about 21M entries): bit-exact
    against the oracle, and (A^T)^T == A."""
    n = 1 << 20
    off, col, val = W.random_uniform_rows(n, n, 8, 32, seed=51)
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(off, col, val)
    T = A.transpose()
    assert _bits_equal(T.download(), oracle.transpose(off, col, val, n))
    assert _bits_equal(T.transpose().download(), (off, col, val))


@pytest.mark.parametrize("shape", [(0, 5), (5, 0), (0, 0)], ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("drop", [False, True], ids=["serial", "drop_zeros"])
def test_transpose_degenerate_shapes(dev, shape, drop):
    n_rows, n_cols = shape
    A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    A.upload(np.zeros(n_rows + 1, np.uint64), np.zeros(0, np.uint32), np.zeros(0))
    T = A.transpose(drop_zeros=drop)
    assert T.n_rows == n_cols and T.n_cols == n_rows and T.nnz == 0
    off, col, val = T.download()
    assert off.tolist() == [0] * (n_cols + 1) and col.size == 0
