"""The C++23 host mirror (include/spira_amd.hpp) against the oracle, through a
line-oriented driver (tests/cpp/host_mirror_driver.cpp, built by
__graft_entry__.build()): the reference's lifecycle known answers, multi-cycle
lock() bit-exact against the oracle model (including row_at_mut-style dirty
rows and +-0.0 deletes), the serial / parallel / operator* spmv overloads
(normwise 1e-12, exact on integer known answers) with the reference's check
order and exception types, partition_at slices, rebalance() boundaries
against the oracle's compute_partition_boundaries, and load_csr validation.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from oracle_model import OracleMatrix, csr_equal, load_kats, run_scenario

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "host_mirror_driver")
KATS = load_kats()

_ERRORS = {"invalid_argument": ValueError, "logic_error": RuntimeError, "out_of_range": IndexError,
           "length_error": OverflowError, "runtime_error": OSError, "bad_alloc": MemoryError}


class Driver:
    def __init__(self):
        if not os.path.exists(DRIVER):
            raise FileNotFoundError(f"{DRIVER} is not built (run __graft_entry__.build())")
        self.p = subprocess.Popen([DRIVER], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True, bufsize=1)

    def cmd(self, line):
        self.p.stdin.write(line + "\n")
        self.p.stdin.flush()
        out = self.p.stdout.readline().strip()
        if out.startswith("error "):
            parts = out.split(" ", 2)
            err = _ERRORS.get(parts[1], AssertionError)(parts[2] if len(parts) > 2 else "")
            err.kind = parts[1]
            raise err
        return out

    def close(self):
        if self.p.poll() is None:
            self.p.stdin.write("quit\n")
            self.p.stdin.flush()
            self.p.wait(timeout=30)


def _h(v):
    return float(v).hex()


def _parse_csr(text):
    head, offs, cols, vals = text.split(";")
    n, nnz = (int(t) for t in head.split())
    return (np.array([int(t) for t in offs.split()], np.uint64), np.array([int(t) for t in cols.split()], np.uint32),
            np.array([float.fromhex(t) for t in vals.split()], np.float64)), n, nnz


class CppMatrix:
    """spira::matrix / parallel_matrix on the C++ mirror, with the method names
    the lifecycle scenarios use."""

    def __init__(self, drv, n_rows, n_cols, threads=None):
        self.d = drv
        self.nr, self.nc = n_rows, n_cols
        if threads is None:
            drv.cmd(f"new serial {n_rows} {n_cols}")
        else:
            drv.cmd(f"new parallel {n_rows} {n_cols} {threads}")

    def insert(self, r, c, v):
        self.d.cmd(f"insert {r} {c} {_h(v)}")

    def mark_dirty(self, r):
        self.d.cmd(f"mark_dirty {r}")

    def lock(self):
        self.d.cmd("lock")

    def open(self):
        self.d.cmd("open")

    def clear(self):
        self.d.cmd("clear")

    def rebalance(self):
        self.d.cmd("rebalance")

    def get(self, r, c):
        return float.fromhex(self.d.cmd(f"get {r} {c}").split()[1])

    def accumulate(self, r):
        return float.fromhex(self.d.cmd(f"accumulate {r}").split()[1])

    def contains(self, r, c):
        return self.d.cmd(f"contains {r} {c}").split()[1] == "1"

    def row_nnz(self, r):
        return int(self.d.cmd(f"row_nnz {r}").split()[1])

    def nnz(self):
        return int(self.d.cmd("nnz").split()[1])

    def is_locked(self):
        return self.d.cmd("is_locked").split()[1] == "1"

    def csr(self):
        out = self.d.cmd("csr")
        if out == "csr none":
            return None
        return _parse_csr(out[4:])[0]

    def spmv(self, x, kind="serial", ny=None):
        ny = self.nr if ny is None else ny
        out = self.d.cmd(f"spmv {kind} {len(x)} {ny} " + " ".join(_h(v) for v in x))
        return np.array([float.fromhex(t) for t in out.split()[1:]], np.float64)

    def spmv32(self, x, kind="serial", ny=None):
        ny = self.nr if ny is None else ny
        out = self.d.cmd(f"spmv32 {kind} {len(x)} {ny} " + " ".join(_h(v) for v in x))
        return np.array([float.fromhex(t) for t in out.split()[1:]])

    def partition(self, t):
        out = self.d.cmd(f"partition {t}")
        _, rest = out.split(" ", 1)
        se, csr = rest.split(";", 1)
        a, b = (int(t) for t in se.split())
        return a, b, _parse_csr(csr)[0]

    def bounds(self):
        return [int(t) for t in self.d.cmd("bounds").split()[1:]]

    def parallel_fill(self, rows, cols, vals):
        self.d.cmd(f"parallel_fill {len(rows)} "
                   + " ".join(f"{int(r)} {int(c)} {_h(v)}" for r, c, v in zip(rows, cols, vals)))

    def load_csr(self, off, col, val):
        self.d.cmd(f"load_csr {len(off) - 1} {len(col)} " + " ".join(str(int(v)) for v in off) + " "
                   + " ".join(str(int(v)) for v in col) + " " + " ".join(_h(v) for v in val))


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_reports_a_missing_device_as_an_exception():
    """Without a GPU the mirror's device constructor throws and the driver
    reports it (the status -> exception path, no device work)."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    d = Driver()
    with pytest.raises((OSError, ValueError)):
        d.cmd("new serial 3 3")
    d.close()


@pytest.fixture(scope="module")
def drv():
    d = Driver()
    yield d
    d.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", KATS["lifecycle"], ids=lambda c: c["source"])
def test_lifecycle_known_answers_on_cpp_mirror(drv, case):
    run_scenario(CppMatrix(drv, case["n_rows"], case["n_cols"], case.get("threads")), case)


@pytest.mark.gpu
@pytest.mark.parametrize("parallel", [False, True])
def test_multicycle_lock_bit_exact_against_oracle_model(drv, parallel):
    rng = np.random.default_rng(5 + parallel)
    n_rows, n_cols = 300, 200
    m = CppMatrix(drv, n_rows, n_cols, threads=4 if parallel else None)
    ref = OracleMatrix(n_rows, n_cols, parallel=parallel)
    for cycle in range(4):
        n_ins = int(rng.integers(0, 1500))
        for _ in range(n_ins):
            r, c = int(rng.integers(0, n_rows)), int(rng.integers(0, 40))
            v = 0.0 if rng.random() < 0.15 else float(rng.uniform(-1, 1))
            if v == 0.0 and rng.random() < 0.5:
                v = -0.0
            m.insert(r, c, v)
            ref.insert(r, c, v)
        if not parallel and cycle == 2:
            for r in rng.choice(n_rows, 20, replace=False):
                m.mark_dirty(int(r))
                ref.dirty[int(r)] = 1
        # open-mode reads see the buffer first
        for _ in range(20):
            r, c = int(rng.integers(0, n_rows)), int(rng.integers(0, 40))
            assert m.get(r, c) == ref.get(r, c)
            assert m.contains(r, c) == ref.contains(r, c)
        assert m.nnz() == ref.nnz()
        m.lock()
        ref.lock()
        assert csr_equal(m.csr(), ref.csr()), f"cycle {cycle}"
        assert m.nnz() == ref.nnz()
        m.open()
        ref.open()


@pytest.mark.gpu
def test_parallel_fill_bit_exact_against_oracle_model(drv):
    """parallel_matrix::parallel_fill (one host thread per partition, each
    inserting its own rows) over open/lock cycles with a rebalance between
    them: the CSR equals the oracle model fed the same triples in order."""
    rng = np.random.default_rng(17)
    n_rows, n_cols = 500, 120
    m = CppMatrix(drv, n_rows, n_cols, threads=6)
    ref = OracleMatrix(n_rows, n_cols, parallel=True)
    for cycle in range(4):
        n = int(rng.integers(0, 3000))
        rows = rng.integers(0, n_rows, n)
        cols = rng.integers(0, 30, n)
        vals = np.where(rng.random(n) < 0.1, -0.0, rng.uniform(-1, 1, n))
        m.parallel_fill(rows, cols, vals)
        for r, c, v in zip(rows, cols, vals):
            ref.insert(int(r), int(c), float(v))
        assert m.nnz() == ref.nnz()
        m.lock()
        ref.lock()
        assert csr_equal(m.csr(), ref.csr()), f"cycle {cycle}"
        if cycle == 1:
            m.rebalance()
        m.open()
        ref.open()
    with pytest.raises(IndexError, match="Column index out of range"):
        m.parallel_fill([0], [n_cols], [1.0])
    m.lock()
    with pytest.raises(RuntimeError, match="must be open"):
        m.parallel_fill([0], [0], [1.0])


@pytest.mark.gpu
def test_spmv_overloads_known_answers_and_checks(drv):
    # spmv_tests.cpp: 3x3 known answer {9, 9, 33}
    for case in KATS["spmv"]:
        m = CppMatrix(drv, case["n_rows"], case["n_cols"])
        for r, c, v in case["inserts"]:
            m.insert(r, c, v)
        m.lock()
        for kind in ("serial", "op"):
            assert m.spmv(case["x"], kind).tolist() == case["y"], case["source"]
    m = CppMatrix(drv, 3, 4)
    with pytest.raises(RuntimeError) as e:  # checks: x size, y size, then mode
        m.spmv([1.0] * 4)
    assert "locked" in str(e.value)
    with pytest.raises(ValueError) as e:
        m.spmv([1.0] * 3)
    assert "input vector x" in str(e.value)
    with pytest.raises(ValueError) as e:
        m.spmv([1.0] * 4, ny=2)
    assert "output vector y" in str(e.value)
    p = CppMatrix(drv, 3, 4, threads=2)
    p.lock()
    with pytest.raises(ValueError) as e:
        p.spmv([1.0] * 3, "parallel")
    assert str(e.value) == "spmv: x size does not match matrix column count"
    assert p.spmv([1.0] * 4, "parallel").tolist() == [0.0, 0.0, 0.0]


@pytest.mark.gpu
@pytest.mark.parametrize("parallel", [False, True])
def test_spmv_random_matches_oracle(drv, parallel):
    rng = np.random.default_rng(11)
    n_rows, n_cols = 500, 400
    m = CppMatrix(drv, n_rows, n_cols, threads=3 if parallel else None)
    for _ in range(6000):
        m.insert(int(rng.integers(0, n_rows)), int(rng.integers(0, n_cols)), float(rng.uniform(-1, 1)))
    m.lock()
    off, col, val = m.csr()
    x = rng.uniform(-1, 1, n_cols)
    y = m.spmv(x, "parallel" if parallel else "serial")
    assert oracle.normwise_rel_err(y, oracle.spmv_f64(off, col, val, x)) <= 1e-12


@pytest.mark.gpu
def test_partitions_and_rebalance_match_reference_boundaries(drv):
    rng = np.random.default_rng(3)
    n_rows, n_cols, T = 1000, 50, 5
    m = CppMatrix(drv, n_rows, n_cols, threads=T)
    for r in range(n_rows):  # skewed: the last rows are long
        for c in rng.choice(n_cols, size=1 + (r * 40) // n_rows, replace=False):
            m.insert(r, int(c), 1.0)
    assert m.bounds() == [t * n_rows // T for t in range(T + 1)]
    with pytest.raises(RuntimeError):
        m.rebalance()  # open
    m.lock()
    off, col, val = m.csr()
    for t in range(T):
        a, b, (po, pc, pv) = m.partition(t)
        base = int(off[a])
        assert po.tolist() == (off[a:b + 1] - off[a]).tolist()
        assert pc.tolist() == col[base:int(off[b])].tolist()
    m.rebalance()
    assert m.bounds() == oracle.partition_bounds(np.diff(off.astype(np.int64)), T).tolist()
    x = rng.uniform(-1, 1, n_cols)
    assert oracle.normwise_rel_err(m.spmv(x, "parallel"), oracle.spmv_f64(off, col, val, x)) <= 1e-12
    with pytest.raises(IndexError):
        m.partition(T)


@pytest.mark.gpu
@pytest.mark.parametrize("case", KATS["partition_balanced"], ids=lambda c: c["source"])
def test_rebalance_on_reference_partition_cases(drv, case):
    row_nnz = case["row_nnz"]
    n = len(row_nnz)
    if n == 0:
        pytest.skip("no rows")
    width = max(max(row_nnz), 1)
    m = CppMatrix(drv, n, width, threads=case["n_threads"])
    for r, k in enumerate(row_nnz):
        for c in range(k):
            m.insert(r, c, 1.0)
    m.lock()
    m.rebalance()
    b = m.bounds()
    if "bounds" in case:
        assert b == case["bounds"], case["source"]
    for k, v in case.get("bounds_at", {}).items():
        assert b[int(k)] == v, case["source"]


@pytest.mark.gpu
def test_load_csr_validates_and_locks(drv):
    m = CppMatrix(drv, 2, 3)
    m.load_csr([0, 1, 2], [2, 0], [5.0, 0.0])  # explicit zero accepted
    assert m.is_locked() and m.get(0, 2) == 5.0 and m.contains(1, 0)
    with pytest.raises(ValueError):
        m.load_csr([0, 2, 2], [1, 1], [1.0, 2.0])  # repeated column
    with pytest.raises(ValueError):
        m.load_csr([0, 1, 1], [3], [1.0])  # column out of bounds
    with pytest.raises(ValueError):
        m.load_csr([1, 1, 1], [0], [1.0])  # offsets[0] != 0
    with pytest.raises(RuntimeError):
        m.insert(0, 0, 1.0)  # locked
    with pytest.raises(IndexError):  # (the driver holds one matrix: this replaces m)
        CppMatrix(drv, 2, 3).get(2, 0)


def _cpp_from_case(drv, case):
    m = CppMatrix(drv, case["n_rows"], case["n_cols"],
                  threads=case.get("n_threads") if case.get("mode", "serial") != "serial" else None)
    for r, c, v in case["inserts"]:
        m.insert(r, c, v)
    return m


@pytest.mark.gpu
@pytest.mark.parametrize("case", KATS["transpose"], ids=lambda c: c["source"])
def test_transpose_known_answers_on_cpp_mirror(drv, case):
    """serial::algorithms::transpose / parallel::algorithms::transpose through
    the C++ mirror: the reference's known answers and the oracle's CSR."""
    m = _cpp_from_case(drv, case)
    m.lock()
    before = m.csr()
    drv.cmd("transpose")
    shape = (case["n_cols"], case["n_rows"])
    if case.get("twice"):
        drv.cmd("transpose")
        shape = (case["n_rows"], case["n_cols"])
    assert m.is_locked()
    for r, c, v in case["expect"]:
        assert m.get(r, c) == v
    for r, c in case.get("expect_zero", []):
        assert not m.contains(r, c)
    if "nnz" in case:
        assert m.nnz() == case["nnz"]
    got = m.csr()
    assert len(got[0]) == shape[0] + 1
    if not case.get("twice"):
        exp = oracle.transpose(*before, case["n_cols"], drop_zeros=case["mode"] != "serial")
        assert all(np.array_equal(a, b) for a, b in zip(got, exp))


@pytest.mark.gpu
@pytest.mark.parametrize("case", KATS["transpose_itself"], ids=lambda c: c["source"])
def test_transpose_itself_known_answers_on_cpp_mirror(drv, case):
    m = _cpp_from_case(drv, case)
    if case.get("error"):
        with pytest.raises(RuntimeError) as e:  # std::logic_error
            drv.cmd("transpose_itself")
        assert str(e.value) == "in-place transpose requires square matrix"
        return
    drv.cmd("transpose_itself")
    assert not m.is_locked()
    if case.get("twice"):
        drv.cmd("transpose_itself")
    for r, c, v in case["expect"]:
        assert m.get(r, c) == v
    for r, c in case.get("expect_zero", []):
        assert not m.contains(r, c)


@pytest.mark.gpu
def test_transpose_checks_and_zeros_on_cpp_mirror(drv):
    m = CppMatrix(drv, 3, 4)
    with pytest.raises(RuntimeError) as e:
        drv.cmd("transpose")
    assert str(e.value) == "transpose: matrix must be locked"
    m.lock()
    with pytest.raises(RuntimeError) as e:
        drv.cmd("transpose_itself")
    assert str(e.value) == "transpose_itself: matrix must be open"
    # load_csr keeps an explicit zero; the serial transpose keeps it too
    m = CppMatrix(drv, 2, 3)
    m.load_csr([0, 2, 3], [0, 2, 1], [0.0, 4.0, -1.5])
    drv.cmd("transpose")
    off, col, val = m.csr()
    assert off.tolist() == [0, 1, 2, 3] and col.tolist() == [0, 1, 0] and val.tolist() == [0.0, -1.5, 4.0]
    assert m.contains(0, 0) and m.get(2, 0) == 4.0


@pytest.mark.gpu
@pytest.mark.parametrize("parallel", [False, True])
def test_transpose_random_matches_oracle_on_cpp_mirror(drv, parallel):
    rng = np.random.default_rng(17)
    n_rows, n_cols = 700, 450
    m = CppMatrix(drv, n_rows, n_cols, threads=4 if parallel else None)
    for _ in range(8000):
        m.insert(int(rng.integers(0, n_rows)), int(rng.integers(0, n_cols)), float(rng.uniform(-1, 1)))
    m.lock()
    before = m.csr()
    drv.cmd("transpose")
    got = m.csr()
    exp = oracle.transpose(*before, n_cols, drop_zeros=parallel)
    assert all(np.array_equal(a, b) for a, b in zip(got, exp))
    if parallel:
        assert m.bounds() == [t * n_cols // 4 for t in range(5)]


def _cpp_pair(drv, case, inner=False):
    """Left operand pushed on the driver's stack, right operand current."""
    threads = case.get("n_threads") if case["mode"] == "parallel" else None
    lock = case.get("lock", True)
    n_cols_a = case["n_inner"] if inner else case["n_cols"]
    A = CppMatrix(drv, case["n_rows"], n_cols_a, threads=threads)
    for r, c, v in case["a"]:
        A.insert(r, c, v)
    if lock:
        A.lock()
    a_csr = A.csr()
    drv.cmd("push")
    br, bc = case.get("b_shape", (case["n_inner"], case["n_cols"]) if inner else (case["n_rows"], case["n_cols"]))
    B = CppMatrix(drv, br, bc, threads=threads)
    for r, c, v in case["b"]:
        B.insert(r, c, v)
    if lock:
        B.lock()
    return B, a_csr, B.csr()


def _cpp_check(m, case):
    off, col, val = m.csr()
    if "expect_rows" in case:
        for r, exp in enumerate(case["expect_rows"]):
            assert [[int(col[k]), float(val[k])] for k in range(int(off[r]), int(off[r + 1]))] == exp
    for r, c, v in case.get("expect", []):
        assert m.get(r, c) == v
    for r, c in case.get("expect_zero", []):
        assert not m.contains(r, c)
    if "nnz" in case:
        assert m.nnz() == case["nnz"]
    assert not np.any(val == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("case", KATS["add"], ids=lambda c: f'{c["source"]}:{c.get("op", "")}')
def test_add_known_answers_on_cpp_mirror(drv, case):
    """serial / parallel MatrixAddition and operator+ / operator- through the C++ mirror."""
    m, a, b = _cpp_pair(drv, case)
    cmd = "sub" if case.get("op") == "-" else "add"
    if "error" in case:
        with pytest.raises(ValueError) as e:
            drv.cmd(cmd)
        assert str(e.value) == case["message"]
        return
    drv.cmd(cmd)
    assert m.is_locked()
    _cpp_check(m, case)
    assert all(np.array_equal(x, y) for x, y in zip(m.csr(), oracle.add(a, b, subtract=cmd == "sub")))


@pytest.mark.gpu
@pytest.mark.parametrize("case", KATS["spgemm"], ids=lambda c: c["source"])
def test_spgemm_known_answers_on_cpp_mirror(drv, case):
    m, a, b = _cpp_pair(drv, case, inner=True)
    if "error" in case:
        with pytest.raises(ValueError) as e:
            drv.cmd("mul")
        assert str(e.value) == case["message"]
        return
    drv.cmd("mul")
    _cpp_check(m, case)
    exp = oracle.spgemm(a, b, case["n_cols"])
    got = m.csr()
    assert all(np.array_equal(x, y) for x, y in zip(got, exp))


@pytest.mark.gpu
def test_in_place_operators_on_cpp_mirror(drv):
    rng = np.random.default_rng(23)
    ins_a = [(int(rng.integers(0, 40)), int(rng.integers(0, 40)), float(rng.integers(-3, 4))) for _ in range(300)]
    ins_b = [(int(rng.integers(0, 40)), int(rng.integers(0, 40)), float(rng.integers(-3, 4))) for _ in range(300)]
    for op in ("iadd", "isub", "imul"):
        A = CppMatrix(drv, 40, 40)
        for t in ins_a:
            A.insert(*t)
        A.lock()
        a = A.csr()
        drv.cmd("push")
        B = CppMatrix(drv, 40, 40)
        for t in ins_b:
            B.insert(*t)
        B.lock()
        b = B.csr()
        drv.cmd(op)
        exp = oracle.spgemm(a, b, 40) if op == "imul" else oracle.add(a, b, subtract=op == "isub")
        assert all(np.array_equal(x, y) for x, y in zip(B.csr(), exp)), op
    # operator- refuses an open right operand first (its scaled copy needs it locked)
    A = CppMatrix(drv, 3, 3)
    A.lock()
    drv.cmd("push")
    CppMatrix(drv, 3, 3)
    with pytest.raises(RuntimeError) as e:
        drv.cmd("sub")
    assert str(e.value) == "multiplication_scaler: input matrix must be locked"


@pytest.mark.gpu
def test_float_spmv_overloads(drv):
    """serial/parallel spmv for the float instantiation (serial/spmv.hpp:90-135,
    parallel/algorithms/spmv.hpp:85-119) through the C++ mirror: the known
    answer (spmv_tests.cpp:121-139) exactly, a random case within 1e-6 of the
    FP64 sum of its FP32 data, and the reference's size checks."""
    for case in KATS["spmv_f32"]:
        for kind, threads in (("serial", None), ("parallel", 2)):
            m = CppMatrix(drv, case["n_rows"], case["n_cols"], threads)
            for r, c, v in case["inserts"]:
                m.insert(r, c, v)
            m.lock()
            assert m.spmv32(case["x"], kind).tolist() == case["y"]
    rng = np.random.default_rng(12)
    m = CppMatrix(drv, 300, 200)
    rows, cols = rng.integers(0, 300, 3000), rng.integers(0, 200, 3000)
    vals = rng.uniform(-1, 1, 3000).astype(np.float32).astype(np.float64)
    for r, c, v in zip(rows, cols, vals):
        m.insert(int(r), int(c), float(v))
    m.lock()
    off, col, val = m.csr()
    x = rng.uniform(-1, 1, 200).astype(np.float32)
    y = m.spmv32(x.astype(np.float64))
    exact = oracle.spmv_f64(off, col, val, x.astype(np.float64))
    assert oracle.normwise_rel_err(y, exact) <= 1e-6
    with pytest.raises(ValueError, match="input vector x"):
        m.spmv32([1.0] * 3)
