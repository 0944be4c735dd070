"""Device matrix addition / subtraction and SpGEMM against the reference's known
answers and the oracle, through the C-ABI (spira_add_f64, spira_spgemm_f64)
and the Python mirror.

Bar: bit-exact.  Offsets, columns and values (as bit patterns) of C must equal
the oracle's (serial/matrix_addition.hpp:34-112 with zeros dropped by the
output's lock(); serial/spgemm.hpp:12-50 with each product rounded, then added
from 0.0 in A's and then B's column order, sums equal to 0 dropped).  Full-size
SpGEMM is checked bit-exact on a sample of rows and by (A B) x == A (B x)
within 1e-12 normwise.
"""
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


def _bits_equal(a, b):
    return (np.array_equal(np.asarray(a[0], np.uint64), np.asarray(b[0], np.uint64))
            and np.array_equal(np.asarray(a[1], np.uint32), np.asarray(b[1], np.uint32))
            and np.array_equal(np.asarray(a[2], np.float64).view(np.uint64),
                               np.asarray(b[2], np.float64).view(np.uint64)))


def _mat(dev, n_rows, n_cols, ins, threads=None, lock=True):
    m = spira_amd.Matrix(n_rows, n_cols, dev) if threads is None else \
        spira_amd.ParallelMatrix(n_rows, n_cols, threads, dev)
    if ins:
        m.insert_many([t[0] for t in ins], [t[1] for t in ins], [t[2] for t in ins])
    if lock:
        m.lock()
    return m


def _dense(m):
    r, c = m.shape()
    d = np.zeros((r, c))
    off, col, val = m.csr()
    for i in range(r):
        for k in range(int(off[i]), int(off[i + 1])):
            d[i, col[k]] = val[k]
    return d


def _check(m, case):
    off, col, val = m.csr()
    if "expect_rows" in case:
        for r, exp in enumerate(case["expect_rows"]):
            assert [[int(col[k]), float(val[k])] for k in range(int(off[r]), int(off[r + 1]))] == exp
    for r, c, v in case.get("expect", []):
        assert m.get(r, c) == v
    for r, c in case.get("expect_zero", []):
        assert m.get(r, c) == 0.0 and not m.contains(r, c)
    if "nnz" in case:
        assert m.nnz() == case["nnz"]
    if "expect_dense" in case:
        assert np.array_equal(_dense(m), np.array(case["expect_dense"]))
    assert not np.any(val == 0.0)


@pytest.mark.parametrize("case", KATS["add"], ids=lambda c: f'{c["source"]}:{c.get("op", "")}')
def test_add_known_answers(dev, case):
    threads = case.get("n_threads") if case["mode"] == "parallel" else None
    lock = case.get("lock", True)
    A = _mat(dev, case["n_rows"], case["n_cols"], case["a"], threads, lock)
    br, bc = case.get("b_shape", (case["n_rows"], case["n_cols"]))
    B = _mat(dev, br, bc, case["b"], threads, lock)
    op = case.get("op", "+")
    if "error" in case:
        with pytest.raises(spira_amd.InvalidArgument, match=case["message"].replace("(", r"\(").replace(")", r"\)")):
            _ = (A - B) if op == "-" else (A + B if case["mode"] == "operator" else spira_amd.MatrixAddition(A, B))
        return
    C = (A - B) if op == "-" else (A + B if case["mode"] == "operator" else spira_amd.MatrixAddition(A, B))
    assert C.is_locked() and type(C) is type(A)
    _check(C, case)
    if case.get("compare_threads"):
        t = case["compare_threads"]
        C1 = spira_amd.MatrixAddition(_mat(dev, case["n_rows"], case["n_cols"], case["a"], t),
                                      _mat(dev, case["n_rows"], case["n_cols"], case["b"], t))
        assert all(np.array_equal(x, y) for x, y in zip(C1.csr(), C.csr()))
    assert _bits_equal(C.csr(), oracle.add(A.csr(), B.csr(), subtract=op == "-"))


@pytest.mark.parametrize("case", KATS["spgemm"], ids=lambda c: c["source"])
def test_spgemm_known_answers(dev, case):
    threads = case.get("n_threads") if case["mode"] == "parallel" else None
    lock = case.get("lock", True)
    A = _mat(dev, case["n_rows"], case["n_inner"], case["a"], threads, lock)
    br, bc = case.get("b_shape", (case["n_inner"], case["n_cols"]))
    B = _mat(dev, br, bc, case["b"], threads, lock)
    if "error" in case:
        with pytest.raises(spira_amd.InvalidArgument, match=case["message"].replace("(", r"\(").replace(")", r"\)")):
            spira_amd.spgemm(A, B)
        return
    C = A * B if case["mode"] == "operator" else spira_amd.spgemm(A, B)
    assert C.is_locked() and C.shape() == (case["n_rows"], case["n_cols"]) and type(C) is type(A)
    _check(C, case)
    if case.get("in_place"):
        A2 = A
        A2 *= B
        assert all(np.array_equal(x, y) for x, y in zip(A2.csr(), C.csr()))
    if case.get("swap_differs"):
        assert not np.array_equal(_dense(spira_amd.spgemm(B, A)), _dense(C))
    if case.get("compare_threads"):
        t = case["compare_threads"]
        C1 = spira_amd.spgemm(_mat(dev, case["n_rows"], case["n_inner"], case["a"], t),
                              _mat(dev, case["n_inner"], case["n_cols"], case["b"], t))
        assert all(np.array_equal(x, y) for x, y in zip(C1.csr(), C.csr()))
    assert _bits_equal(C.csr(), oracle.spgemm(A.csr(), B.csr(), case["n_cols"]))


def test_mode_and_shape_checks(dev):
    A = _mat(dev, 3, 4, [[0, 0, 1.0]], lock=False)
    B = _mat(dev, 3, 4, [[0, 0, 1.0]])
    with pytest.raises(spira_amd.LogicError, match="MatrixAddition: A must be locked"):
        spira_amd.MatrixAddition(A, B)
    with pytest.raises(spira_amd.LogicError, match="MatrixAddition: B must be locked"):
        spira_amd.MatrixAddition(B, A)
    with pytest.raises(spira_amd.LogicError, match="multiplication_scaler"):
        B - A  # the scaled copy of the open right operand is refused first
    P1, P2 = (_mat(dev, 4, 4, [], threads=t) for t in (2, 3))
    with pytest.raises(spira_amd.InvalidArgument, match="same thread count"):
        spira_amd.MatrixAddition(P1, P2)
    Bt = _mat(dev, 4, 2, [[0, 0, 1.0]])
    with pytest.raises(spira_amd.LogicError, match="spgemm: A must be locked"):
        spira_amd.spgemm(A, Bt)
    with pytest.raises(spira_amd.LogicError, match="spgemm: B must be locked"):
        spira_amd.spgemm(B, _mat(dev, 4, 2, [], lock=False))
    # C-ABI: never-built handles and wrong output shapes
    X = spira_amd.DeviceCSR(dev, 3, 4)
    Y = B.device_csr()
    with pytest.raises(spira_amd.LogicError):
        X.add(Y)
    with pytest.raises(spira_amd.InvalidArgument):
        Y.add(Y, spira_amd.DeviceCSR(dev, 4, 3))
    with pytest.raises(spira_amd.InvalidArgument):
        Y.spgemm(Y)  # 3x4 * 3x4
    with pytest.raises(spira_amd.InvalidArgument):
        Y.spgemm(Bt.device_csr(), spira_amd.DeviceCSR(dev, 3, 4))


def _upload(dev, csr, n_cols):
    A = spira_amd.DeviceCSR(dev, len(csr[0]) - 1, n_cols)
    A.upload(*csr)
    return A


def _with_zeros(csr, seed, frac=0.15):
    off, col, val = csr
    rng = np.random.default_rng(seed)
    val = val.copy()
    z = rng.random(val.size) < frac
    val[z] = np.where(rng.random(int(z.sum())) < 0.5, 0.0, -0.0)
    return off, col, val


def _mostly_empty(n, n_cols, frac, seed):
    rng = np.random.default_rng(seed)
    lens = (rng.random(n) < frac).astype(np.int64)
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    return off, rng.integers(0, n_cols, int(off[-1])).astype(np.uint32), rng.standard_normal(int(off[-1]))


def _cancel_pair(n, n_cols, seed):
    """B shares about half of A's pattern with the values negated (exact zero
    sums) and adds entries of its own."""
    a = W.random_uniform_rows(n, n_cols, 0, 20, seed=seed)
    b_own = W.random_uniform_rows(n, n_cols, 0, 6, seed=seed + 1)
    rng = np.random.default_rng(seed + 2)
    take = rng.random(a[1].size) < 0.5
    rows = np.repeat(np.arange(n), np.diff(a[0].astype(np.int64)))
    r = np.concatenate([rows[take], np.repeat(np.arange(n), np.diff(b_own[0].astype(np.int64)))])
    c = np.concatenate([a[1][take], b_own[1]])
    v = np.concatenate([-a[2][take], b_own[2]])
    ptr, sc, sv = stage(n, r, c, v)
    return a, oracle.lock(ptr, sc, sv)


ADD_CASES = [
    ("empty_both", lambda: ((np.zeros(4, np.uint64), np.zeros(0, np.uint32), np.zeros(0)),) * 2 + (5,)),
    ("one_side_empty", lambda: (W.random_uniform_rows(3000, 500, 0, 9, seed=61),
                                (np.zeros(3001, np.uint64), np.zeros(0, np.uint32), np.zeros(0)), 500)),
    ("random_overlap", lambda: (W.random_uniform_rows(20000, 3000, 0, 30, seed=62),
                                W.random_uniform_rows(20000, 3000, 0, 30, seed=63), 3000)),
    ("cancellation", lambda: (*_cancel_pair(15000, 4000, 64), 4000)),
    ("explicit_zeros", lambda: (_with_zeros(W.random_uniform_rows(8000, 900, 0, 25, seed=65), 66),
                                _with_zeros(W.random_uniform_rows(8000, 900, 0, 25, seed=67), 68), 900)),
    ("long_rows", lambda: (W.random_uniform_rows(40, 200000, 2000, 9000, seed=69),
                           W.random_uniform_rows(40, 200000, 0, 9000, seed=70), 200000)),
    ("mostly_empty_rows", lambda: (_mostly_empty(400000, 3000, 0.02, 71), _mostly_empty(400000, 3000, 0.03, 72), 3000)),
    ("mixed_lengths", lambda: (W.random_uniform_rows(20000, 6000, 0, 90, seed=73),
                               W.random_uniform_rows(20000, 6000, 0, 90, seed=74), 6000)),
    ("one_row_dense_other_sparse", lambda: (W.random_uniform_rows(3, 50000, 20000, 30000, seed=75),
                                            W.random_uniform_rows(3, 50000, 0, 40, seed=76), 50000)),
]


@pytest.mark.parametrize("subtract", [False, True], ids=["add", "sub"])
@pytest.mark.parametrize("case", ADD_CASES, ids=lambda c: c[0])
def test_add_matches_oracle(dev, case, subtract):
    a, b, n_cols = case[1]()
    A, B = _upload(dev, a, n_cols), _upload(dev, b, n_cols)
    C = A.add(B, subtract=subtract)
    exp = oracle.add(a, b, subtract=subtract)
    assert _bits_equal(C.download(), exp)
    assert C.nnz == len(exp[1])
    # in place (C = A) and the inputs untouched otherwise
    assert _bits_equal(B.download(), b)
    A.add(B, A, subtract=subtract)
    assert _bits_equal(A.download(), exp)


def _rand(n, n_cols, lo, hi, seed):
    return W.random_uniform_rows(n, n_cols, lo, hi, seed=seed)


def _int_valued(csr, seed, lo=-3, hi=4):
    """Small integer values: many products cancel to exactly 0."""
    off, col, _ = csr
    return off, col, np.random.default_rng(seed).integers(lo, hi, col.size).astype(np.float64)


def _wide_a_row(n_inner, seed):
    """One row of A with 3000 entries over mostly empty rows of B: many A
    entries, few products (the HBM path)."""
    rng = np.random.default_rng(seed)
    cols = np.sort(rng.choice(n_inner, 3000, replace=False)).astype(np.uint32)
    off = np.array([0, 3000, 3000, 3001], np.uint64)
    return off, np.concatenate([cols, np.array([5], np.uint32)]), rng.standard_normal(3001)


SPGEMM_CASES = [
    # (name, builder -> (a, b, n_cols of B))
    ("empty", lambda: ((np.zeros(3, np.uint64), np.zeros(0, np.uint32), np.zeros(0)),
                       W.random_uniform_rows(7, 6, 0, 3, seed=80), 6)),
    ("tiny_rect", lambda: (_rand(50, 30, 0, 6, 81), _rand(30, 70, 0, 6, 82), 70)),
    ("small_class", lambda: (_rand(30000, 20000, 0, 16, 83), _rand(20000, 25000, 0, 16, 84), 25000)),
    ("large_class", lambda: (_rand(6000, 5000, 20, 32, 85), _rand(5000, 9000, 20, 32, 86), 9000)),
    ("hbm_class", lambda: (_rand(500, 3000, 60, 120, 87), _rand(3000, 4000, 30, 60, 88), 4000)),
    ("mixed_classes", lambda: (W.powerlaw_rows(20000, clip=3000, seed=89), _rand(20000, 20000, 0, 12, 90), 20000)),
    ("cancellation", lambda: (_int_valued(_rand(8000, 300, 0, 24, 91), 92), _int_valued(_rand(300, 300, 0, 24, 93), 94),
                              300)),
    ("explicit_zeros", lambda: (_with_zeros(_rand(5000, 4000, 0, 20, 95), 96), _with_zeros(_rand(4000, 3000, 0, 20, 97), 98),
                                3000)),
    ("wide_a_row", lambda: (_wide_a_row(100000, 99), _mostly_empty(100000, 8000, 0.001, 100), 8000)),
]


@pytest.mark.parametrize("case", SPGEMM_CASES, ids=lambda c: c[0])
def test_spgemm_matches_oracle(dev, case):
    a, b, n_cols = case[1]()
    A, B = _upload(dev, a, len(b[0]) - 1), _upload(dev, b, n_cols)
    C = A.spgemm(B)
    exp = oracle.spgemm(a, b, n_cols)
    assert _bits_equal(C.download(), exp)
    assert C.nnz == len(exp[1])
    assert _bits_equal(A.download(), a) and _bits_equal(B.download(), b)


def test_spgemm_laplacian_squared_and_in_place(dev):
    """The 512^2 5-point Laplacian squared (integer values: every sum exact),
    bit-exact; C 1 = A (A 1) exactly; then C = A * A into A itself."""
    a = W.laplacian_2d(512)
    n = len(a[0]) - 1
    A = _upload(dev, a, n)
    exp = oracle.spgemm(a, a, n)
    C = A.spgemm(A)
    assert _bits_equal(C.download(), exp)
    ones = np.ones(n)
    assert np.array_equal(C.spmv(ones), A.spmv(A.spmv(ones)))
    A.spgemm(A, A)
    assert _bits_equal(A.download(), exp)


def _rows_of(csr, rows):
    off, col, val = csr
    lens = np.diff(off.astype(np.int64))[rows]
    o = np.zeros(len(rows) + 1, np.uint64)
    np.cumsum(lens, out=o[1:])
    idx = np.concatenate([np.arange(int(off[r]), int(off[r + 1])) for r in rows]) if len(rows) else np.zeros(0, int)
    return o, col[idx], val[idx]


def test_spgemm_full_size_random_1m(dev):
    """1,048,576^2 with 16 random entries per row, squared (268M products,
    about 268M entries): bit-exact on 3000 sampled rows, and (A A) x = A (A x)
    within 1e-12 normwise."""
    (a, _) = W.relock_workload(1 << 20)
    n = len(a[0]) - 1
    A = _upload(dev, a, n)
    C = A.spgemm(A)
    got = C.download()
    rows = np.sort(np.random.default_rng(5).choice(n, 3000, replace=False))
    assert _bits_equal(_rows_of(got, rows), oracle.spgemm(_rows_of(a, rows), a, n))
    x = np.random.default_rng(6).uniform(-1, 1, n)
    y1 = C.spmv(x)
    y2 = A.spmv(A.spmv(x))
    assert oracle.normwise_rel_err(y1, y2) <= 1e-12


def _empty(n_rows):
    return np.zeros(n_rows + 1, np.uint64), np.zeros(0, np.uint32), np.zeros(0)


@pytest.mark.parametrize("shape", [(0, 5), (5, 0), (0, 0)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_add_degenerate_shapes(dev, shape):
    n_rows, n_cols = shape
    A, B = _upload(dev, _empty(n_rows), n_cols), _upload(dev, _empty(n_rows), n_cols)
    for sub in (False, True):
        C = A.add(B, subtract=sub)
        assert C.nnz == 0 and _bits_equal(C.download(), _empty(n_rows))


@pytest.mark.parametrize("dims", [(0, 5, 3), (4, 0, 3), (4, 5, 0), (3, 6, 2)], ids=lambda d: "x".join(map(str, d)))
def test_spgemm_degenerate_shapes(dev, dims):
    """Empty products: no rows, no inner dimension, no columns, or B without entries."""
    n, k, m = dims
    a = W.random_uniform_rows(n, k, 0, min(k, 3), seed=110) if n and k else _empty(n)
    A, B = _upload(dev, a, k), _upload(dev, _empty(k), m)
    C = A.spgemm(B)
    assert C.n_rows == n and C.n_cols == m and C.nnz == 0
    assert _bits_equal(C.download(), oracle.spgemm(a, _empty(k), m))


def test_add_and_spgemm_aliasing_the_right_operand(dev):
    """C = A + B into B's handle and C = A * B into B's handle (square)."""
    a = W.random_uniform_rows(3000, 3000, 0, 12, seed=111)
    b = W.random_uniform_rows(3000, 3000, 0, 12, seed=112)
    A = _upload(dev, a, 3000)
    B = _upload(dev, b, 3000)
    A.add(B, B, subtract=True)
    assert _bits_equal(B.download(), oracle.add(a, b, subtract=True))
    B2 = _upload(dev, b, 3000)
    A.spgemm(B2, B2)
    assert _bits_equal(B2.download(), oracle.spgemm(a, b, 3000))
