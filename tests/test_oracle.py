"""Pins the oracle (CPU restatement) before it is trusted as the checker.

* golden known answers transcribed from the reference's own tests
  (tests/golden/reference_kats.json);
* the reference stress test's own std::map ground truth, regenerated by
  tests/golden/gen_stress.cpp (tests/large_matrix_stress_tests.cpp:11-90);
* independent dict-based formulations of dedup / build / merge on random
  inputs (brute force, different algorithm);
* closed forms for SpMV (dense matmul, exact rational sums for the
  double-double reference);
* planted mistakes in a copy of the oracle's logic must fail those checks.
"""
import json
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import oracle
from oracle_model import (GOLDEN, OracleMatrix, brute_lock, csr_equal, load_kats, random_stream, run_scenario,
                          stage)

KATS = load_kats()


@pytest.mark.parametrize("case", KATS["build_csr"], ids=lambda c: c["source"])
def test_build_csr_golden(case):
    rows = []
    for r, entries in enumerate(case["rows"]):
        for c, v in entries:
            rows.append((r, c, v))
    n = len(case["rows"])
    ptr, c, v = stage(n, [t[0] for t in rows], [t[1] for t in rows], [t[2] for t in rows])
    off, col, val = oracle.lock(ptr, c, v)
    assert off.tolist() == case["offsets"]
    assert col.tolist() == case["cols"]
    assert val.tolist() == case["vals"]


@pytest.mark.parametrize("case", KATS["lifecycle"], ids=lambda c: c["source"])
def test_lifecycle_golden(case):
    mat = OracleMatrix(case["n_rows"], case["n_cols"], parallel="threads" in case)
    run_scenario(mat, case)


@pytest.mark.parametrize("case", KATS["spmv"], ids=lambda c: c["source"])
def test_spmv_golden(case):
    ins = case["inserts"]
    ptr, c, v = stage(case["n_rows"], [t[0] for t in ins], [t[1] for t in ins], [t[2] for t in ins])
    off, col, val = oracle.lock(ptr, c, v)
    y = oracle.spmv_f64(off, col, val, np.array(case["x"]))
    assert y.tolist() == case["y"]


@pytest.mark.parametrize("case", KATS["spmv_f32"], ids=lambda c: c["source"])
def test_spmv_f32_golden(case):
    ins = case["inserts"]
    ptr, c, v = stage(case["n_rows"], [t[0] for t in ins], [t[1] for t in ins], [t[2] for t in ins])
    off, col, val = oracle.lock(ptr, c, v)
    y = oracle.spmv_f32(off, col, val, np.array(case["x"], np.float32))
    assert y.tolist() == [np.float32(t) for t in case["y"]]


@pytest.mark.parametrize("case", KATS["dot"], ids=lambda c: c["source"])
def test_dot_golden(case):
    nnz = len(case["vals"])
    off = np.array([0, nnz], np.uint64)
    y = oracle.spmv_f64(off, np.array(case["cols"], np.uint32), np.array(case["vals"]), np.array(case["x"]))
    assert y[0] == case["expect"]


def test_partition_uniform_golden():
    case = KATS["partition"][0]
    p = oracle.Pool(case["n_threads"], case["n_rows"], 4)
    got = [p.partition(t)[0] for t in range(case["n_threads"])] + [case["n_rows"]]
    p.close()
    assert got == case["bounds"]


@pytest.mark.parametrize("case", KATS["partition_balanced"], ids=lambda c: c["source"])
def test_partition_balanced_golden(case):
    b = oracle.partition_bounds(case["row_nnz"], case["n_threads"]).tolist()
    if "bounds" in case:
        assert b == case["bounds"]
    for k, v in case.get("bounds_at", {}).items():
        assert b[int(k)] == v


def test_is_zero_predicate():
    # traits.hpp:73-88: |x| <= 0
    assert oracle.is_zero(0.0) and oracle.is_zero(-0.0)
    assert not oracle.is_zero(float("nan"))
    assert not oracle.is_zero(5e-324)  # subnormal is not zero
    assert not oracle.is_zero(-1e-300)


# ---------------------------------------------------------------------------
# The reference stress test's ground truth (std::map with zero = erase)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def stress(tmp_path_factory):
    d = tmp_path_factory.mktemp("stress")
    exe = d / "gen_stress"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", os.path.join(GOLDEN, "gen_stress.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(d)], check=True, capture_output=True, text=True)
    rd = lambda name, dt: np.fromfile(str(d / name), dt)
    return {
        "n_expected": int(out.stdout.strip()),
        "rows": rd("ins_row.u64", np.uint64), "cols": rd("ins_col.u64", np.uint64),
        "vals": rd("ins_val.f64", np.float64),
        "off": rd("exp_off.u64", np.uint64), "col": rd("exp_col.u64", np.uint64),
        "val": rd("exp_val.f64", np.float64),
        "x": rd("x.f64", np.float64), "y": rd("y_exp.f64", np.float64),
    }


def test_stress_stream_matches_committed_digest(stress):
    import hashlib
    with open(os.path.join(GOLDEN, "stress_digest.json")) as f:
        dig = json.load(f)
    assert stress["n_expected"] == dig["nnz"]
    for name, key in [("rows", "ins_row"), ("cols", "ins_col"), ("vals", "ins_val"), ("off", "exp_off"),
                      ("col", "exp_col"), ("val", "exp_val"), ("x", "x"), ("y", "y_exp")]:
        assert hashlib.sha256(stress[name].tobytes()).hexdigest()[:16] == dig[key], name


def test_stress_oracle_lock_and_spmv(stress):
    """tests/large_matrix_stress_tests.cpp:62-89: nnz == map size, y within 1e-10."""
    n = 10000
    ptr, c, v = stage(n, stress["rows"].astype(np.int64), stress["cols"], stress["vals"])
    off, col, val = oracle.lock(ptr, c, v)
    assert col.size == stress["n_expected"]
    assert csr_equal((off, col, val), (stress["off"], stress["col"], stress["val"]))
    y = oracle.spmv_f64(off, col, val, stress["x"])
    tol = 1e-10 + 1e-10 * np.maximum(np.abs(y), np.abs(stress["y"]))
    assert np.all(np.abs(y - stress["y"]) <= tol)
    # same summation order as the map walk: bit-identical
    assert np.array_equal(y, stress["y"])


# ---------------------------------------------------------------------------
# Brute force (different formulation) on random inputs
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(6))
def test_lock_vs_bruteforce_multicycle(seed):
    rng = np.random.default_rng(seed)
    n_rows, n_cols = int(rng.integers(1, 40)), int(rng.integers(1, 60))
    old = None
    for cycle in range(4):
        rows, cols, vals = random_stream(rng, n_rows, n_cols, int(rng.integers(0, 200)), col_pool=12)
        ptr, c, v = stage(n_rows, rows, cols, vals)
        dirty = None
        if old is not None and cycle == 2:
            dirty = (rng.random(n_rows) < 0.5).astype(np.uint8)
            dirty[np.bincount(rows, minlength=n_rows) > 0] = 1  # matrix rule: insert sets dirty
        got = oracle.lock(ptr, c, v, old=old, dirty=dirty)
        exp = brute_lock(n_rows, rows, cols, vals, old=old, dirty=dirty)
        assert csr_equal(got, exp), f"cycle {cycle}"
        old = got


def test_merge_keeps_clean_row_zeros_and_drops_dirty_row_zeros():
    # An explicit zero can only enter through load_csr (matrix.hpp:540-575).
    old = (np.array([0, 2, 4], np.uint64), np.array([1, 3, 0, 2], np.uint32), np.array([0.0, 5.0, 0.0, 7.0]))
    ptr, c, v = stage(2, [1], [3], [1.5])
    got = oracle.merge_csr(*old, *oracle.dedup(ptr, c, v))
    # row 0 clean: verbatim incl. zero; row 1 dirty: zero dropped, new col 3 added
    assert got[0].tolist() == [0, 2, 4]
    assert got[1].tolist() == [1, 3, 2, 3]
    assert got[2].tolist() == [0.0, 5.0, 7.0, 1.5]
    # dirty mask marking row 0 dirty with an empty buffer drops its zero
    got2 = oracle.merge_csr(*old, *oracle.dedup(ptr, c, v), dirty=np.array([1, 1], np.uint8))
    assert got2[1].tolist() == [3, 2, 3]


def test_dedup_nan_and_negative_zero():
    ptr, c, v = stage(1, [0, 0, 0, 0], [4, 2, 4, 7], [1.0, float("nan"), -0.0, 2.0])
    dp, dc, dv = oracle.dedup(ptr, c, v)
    assert dc.tolist() == [2, 4, 7]
    assert np.isnan(dv[0]) and dv[1] == 0.0 and np.signbit(dv[1]) and dv[2] == 2.0
    off, col, val = oracle.build_csr(dp, dc, dv)
    assert col.tolist() == [2, 7] and np.isnan(val[0])


# ---------------------------------------------------------------------------
# SpMV closed forms
# ---------------------------------------------------------------------------
def _rand_csr(rng, n_rows, n_cols, max_len):
    rows, cols, vals = [], [], []
    for r in range(n_rows):
        k = int(rng.integers(0, max_len + 1))
        cs = rng.choice(n_cols, size=min(k, n_cols), replace=False)
        rows += [r] * cs.size
        cols += cs.tolist()
        vals += rng.uniform(-1, 1, cs.size).tolist()
    ptr, c, v = stage(n_rows, rows, cols, vals)
    return oracle.lock(ptr, c, v)


@pytest.mark.parametrize("seed", range(4))
def test_spmv_vs_dense(seed):
    rng = np.random.default_rng(100 + seed)
    n_rows, n_cols = int(rng.integers(1, 80)), int(rng.integers(1, 80))
    off, col, val = _rand_csr(rng, n_rows, n_cols, 20)
    dense = np.zeros((n_rows, n_cols))
    for r in range(n_rows):
        for k in range(int(off[r]), int(off[r + 1])):
            dense[r, col[k]] = val[k]
    x = rng.uniform(-1, 1, n_cols)
    y = oracle.spmv_f64(off, col, val, x)
    np.testing.assert_allclose(y, dense @ x, rtol=1e-13, atol=1e-14)


def test_spmv_dd_is_exact_on_small_rows():
    rng = np.random.default_rng(7)
    off, col, val = _rand_csr(rng, 30, 50, 12)
    x = rng.uniform(-1, 1, 50)
    hi, lo = oracle.spmv_dd(off, col, val, x)
    for r in range(30):
        exact = sum((Fraction(float(val[k])) * Fraction(float(x[col[k]])) for k in range(int(off[r]), int(off[r + 1]))),
                    Fraction(0))
        approx = Fraction(float(hi[r])) + Fraction(float(lo[r]))
        assert abs(approx - exact) <= abs(exact) * Fraction(1, 2 ** 100) + Fraction(1, 2 ** 1070)


def test_spmm_matches_columns_of_spmv():
    rng = np.random.default_rng(9)
    off, col, val = _rand_csr(rng, 25, 31, 9)
    X = rng.uniform(-1, 1, (31, 5))
    Y = oracle.spmm_f64(off, col, val, X)
    for j in range(5):
        np.testing.assert_array_equal(Y[:, j], oracle.spmv_f64(off, col, val, X[:, j].copy()))


# ---------------------------------------------------------------------------
# Threaded CPU baseline = serial oracle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("threads", [1, 2, 4])
def test_pool_matches_serial(threads):
    rng = np.random.default_rng(threads)
    n, nc = 103, 77
    pool = oracle.Pool(threads, n, nc)
    old = None
    for cycle in range(3):
        rows, cols, vals = random_stream(rng, n, nc, 400, col_pool=40)
        ptr, c, v = stage(n, rows, cols, vals)
        pool.lock(ptr, c, v)
        exp = oracle.lock(ptr, c, v, old=old)
        got = pool.export_csr()
        assert csr_equal(got, exp), f"cycle {cycle}"
        old = exp
        x = rng.uniform(-1, 1, nc)
        np.testing.assert_array_equal(pool.spmv(x), oracle.spmv_f64(*exp, x))
    pool.close()


@pytest.mark.parametrize("kernel", ["scalar dot", "avx2", "avx512"])
def test_pool_simd_dots_within_the_dot_bound(kernel):
    """The CPU baseline's per-row gather dots (dot_avx.cpp / dot_avx512.cpp
    shape): every row of lengths 0..80 (all block and tail paths) within the
    summation bound n * 2^-53 * sum |a x| of the exact dot (math.fsum)."""
    import math
    rng = np.random.default_rng(11)
    lens = np.tile(np.arange(81), 3)
    n, nc = len(lens), 5000
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    col = rng.integers(0, nc, int(off[-1])).astype(np.uint32)
    val = rng.normal(size=int(off[-1]))
    x = rng.uniform(-1, 1, nc)
    pool = oracle.Pool(3, n, nc)
    pool.load_csr(off, col, val)
    try:
        pool.force_kernel(kernel)
    except RuntimeError:
        pool.close()
        pytest.skip(f"{kernel} not supported on this CPU")
    pool.pin([0, 0, 0])
    pool.flush_llc(1 << 16)
    y = pool.spmv(x)
    pool.close()
    for i in range(n):
        a, b = int(off[i]), int(off[i + 1])
        prods = val[a:b] * x[col[a:b]]
        exact = math.fsum(prods)
        bound = (b - a + 1) * 2.0 ** -53 * float(np.abs(prods).sum())
        assert abs(y[i] - exact) <= bound, (i, b - a)
    assert oracle.Pool(1, 2, 2).set_simd(True) in ("scalar dot", "avx2", "avx512")


def test_pool_rejects_zero_threads_and_unlocked_spmv():
    with pytest.raises(ValueError):
        oracle.Pool(0, 10, 10)
    p = oracle.Pool(2, 10, 10)
    with pytest.raises(RuntimeError):
        p.spmv(np.zeros(10))
    p.close()


# ---------------------------------------------------------------------------
# Planted mistakes: plausible bugs in a restatement must fail the checks above
# ---------------------------------------------------------------------------
def _mutant_lock_first_write_wins(n_rows, rows, cols, vals):
    per = [dict() for _ in range(n_rows)]
    for r, c, v in zip(rows, cols, vals):
        per[int(r)].setdefault(int(c), float(v))
    return per


def test_planted_first_write_wins_is_caught():
    case = KATS["lifecycle"][0]  # (0,1)=1.0 then (0,1)=9.0 -> 9.0 must win
    ins = [op for op in case["ops"] if op[0] == "insert"]
    per = _mutant_lock_first_write_wins(1, [o[1] for o in ins], [o[2] for o in ins], [o[3] for o in ins])
    assert per[0][1] != 9.0  # the golden expectation would reject this mutant


def test_planted_merge_old_wins_is_caught():
    rng = np.random.default_rng(1)
    rows, cols, vals = random_stream(rng, 10, 10, 60, zero_frac=0.0)
    ptr, c, v = stage(10, rows, cols, vals)
    old = oracle.lock(ptr, c, v)
    rows2, cols2, vals2 = random_stream(rng, 10, 10, 60, zero_frac=0.0)
    exp = brute_lock(10, rows2, cols2, vals2, old=old)
    # mutant: old value wins on collision
    per = [dict() for _ in range(10)]
    for r, c_, v_ in zip(rows2, cols2, vals2):
        per[int(r)][int(c_)] = float(v_)
    off, col, val = old
    mut_vals = []
    for r in range(10):
        oldrow = {int(col[k]): float(val[k]) for k in range(int(off[r]), int(off[r + 1]))}
        m = dict(per[r])
        m.update(oldrow)
        mut_vals += [v_ for _, v_ in sorted(m.items())]
    assert not np.array_equal(np.array(mut_vals), exp[2])


def test_planted_transposed_spmv_is_caught():
    case = KATS["spmv"][1]  # {9, 9, 33}
    ins = case["inserts"]
    ptr, c, v = stage(case["n_rows"], [t[0] for t in ins], [t[1] for t in ins], [t[2] for t in ins])
    off, col, val = oracle.lock(ptr, c, v)
    dense = np.zeros((3, 5))
    for r in range(3):
        for k in range(int(off[r]), int(off[r + 1])):
            dense[r, col[k]] = val[k]
    xt = np.array(case["x"][:3])
    assert (dense.T @ xt)[:3].tolist() != case["y"]


# ---- transpose (serial/transpose.hpp:35-157, parallel/algorithms/transpose.hpp:37-69) ----

def _kat_csr(case):
    ins = case["inserts"]
    ptr, c, v = stage(case["n_rows"], [t[0] for t in ins], [t[1] for t in ins], [t[2] for t in ins])
    return oracle.lock(ptr, c, v)


def _dense_get(off, col, val, r, c):
    lo, hi = int(off[r]), int(off[r + 1])
    hit = np.nonzero(col[lo:hi] == c)[0]
    return float(val[lo + hit[0]]) if hit.size else 0.0


@pytest.mark.parametrize("case", KATS["transpose"], ids=lambda c: c["source"])
def test_transpose_golden(case):
    off, col, val = _kat_csr(case)
    drop = case["mode"] != "serial"
    t = oracle.transpose(off, col, val, case["n_cols"], drop_zeros=drop)
    shape = (case["n_cols"], case["n_rows"])
    if case.get("twice"):
        t = oracle.transpose(*t, case["n_rows"], drop_zeros=drop)
        shape = (case["n_rows"], case["n_cols"])
    assert len(t[0]) == shape[0] + 1
    for r, c, v in case["expect"]:
        assert _dense_get(*t, r, c) == v
    for r, c in case.get("expect_zero", []):
        assert _dense_get(*t, r, c) == 0.0
    if "nnz" in case:
        assert int(t[0][-1]) == case["nnz"]
    for r in range(shape[0]):  # output rows are sorted (and hold no duplicates)
        seg = t[1][int(t[0][r]):int(t[0][r + 1])]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
    if case["mode"] == "parallel_vs_serial":
        s = oracle.transpose(off, col, val, case["n_cols"], drop_zeros=False)
        assert csr_equal(t, s)


@pytest.mark.parametrize("seed", range(12))
def test_transpose_matches_scipy(seed):
    """scipy.sparse's CSR transpose (library routine) on random shapes, with
    empty rows and columns and explicit zeros kept by the serial form."""
    import scipy.sparse as sp
    rng = np.random.default_rng(900 + seed)
    n_rows, n_cols = int(rng.integers(1, 300)), int(rng.integers(1, 300))
    dense = rng.standard_normal((n_rows, n_cols)) * (rng.random((n_rows, n_cols)) < rng.uniform(0.0, 0.3))
    dense[rng.random((n_rows, n_cols)) < 0.02] = 0.0
    m = sp.csr_matrix(dense)
    off, col, val = m.indptr.astype(np.uint64), m.indices.astype(np.uint32), m.data.astype(np.float64)
    if val.size:  # explicit zeros (as load_csr can bring in)
        val[rng.random(val.size) < 0.1] = 0.0
    ref = sp.csr_matrix((val, col, off), shape=(n_rows, n_cols)).T.tocsr()
    ref.sort_indices()
    t = oracle.transpose(off, col, val, n_cols)
    assert np.array_equal(t[0], ref.indptr.astype(np.uint64))
    assert np.array_equal(t[1], ref.indices.astype(np.uint32))
    assert np.array_equal(t[2].view(np.uint64), ref.data.astype(np.float64).view(np.uint64))
    # parallel form: the same with explicit zeros removed
    tz = oracle.transpose(off, col, val, n_cols, drop_zeros=True)
    keep = ~((t[2].view(np.uint64) & np.uint64(0x7fffffffffffffff)) == 0)
    rows = np.repeat(np.arange(n_cols), np.diff(t[0].astype(np.int64)))
    exp_off = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n_cols))]).astype(np.uint64)
    assert np.array_equal(tz[0], exp_off) and np.array_equal(tz[1], t[1][keep])
    assert np.array_equal(tz[2], t[2][keep])
    # transposing twice gives the input back
    assert csr_equal(oracle.transpose(*t, n_rows), (off, col, val))


def _transpose_py(off, col, val, n_cols, descending_rows=False):
    """Pure-Python two-pass transpose; descending_rows plants the mistake of
    filling from the last input row (output rows come out reversed)."""
    n = len(off) - 1
    cnt = np.zeros(n_cols + 1, np.int64)
    for k in range(int(off[-1])):
        cnt[int(col[k]) + 1] += 1
    t_off = np.cumsum(cnt).astype(np.uint64)
    cur = t_off[:-1].astype(np.int64).copy()
    t_col = np.zeros(int(off[-1]), np.uint32)
    t_val = np.zeros(int(off[-1]))
    for i in (range(n - 1, -1, -1) if descending_rows else range(n)):
        for k in range(int(off[i]), int(off[i + 1])):
            j = int(col[k])
            t_col[cur[j]] = i
            t_val[cur[j]] = val[k]
            cur[j] += 1
    return t_off, t_col, t_val


def test_planted_transpose_unstable_fill_is_caught():
    """The restatement agrees with the oracle; the same code with the fill
    walking the input rows in descending order fails the scipy check."""
    import scipy.sparse as sp
    rng = np.random.default_rng(5)
    dense = rng.standard_normal((40, 30)) * (rng.random((40, 30)) < 0.3)
    m = sp.csr_matrix(dense)
    off, col, val = m.indptr.astype(np.uint64), m.indices.astype(np.uint32), m.data
    ref = m.T.tocsr()
    ref.sort_indices()
    good = _transpose_py(off, col, val, 30)
    assert csr_equal(good, oracle.transpose(off, col, val, 30))
    bad = _transpose_py(off, col, val, 30, descending_rows=True)
    assert not np.array_equal(bad[1], ref.indices.astype(np.uint32))


def _trip_csr(n_rows, ins):
    """Locked CSR of an insert list (the reference's test helpers insert and lock)."""
    ptr, c, v = stage(n_rows, [t[0] for t in ins], [t[1] for t in ins], [t[2] for t in ins])
    return oracle.lock(ptr, c, v)


def _dense_of(csr, n_cols):
    off, col, val = csr
    d = np.zeros((len(off) - 1, n_cols))
    for r in range(len(off) - 1):
        for k in range(int(off[r]), int(off[r + 1])):
            d[r, col[k]] = val[k]
    return d


def _check_rows(csr, case, n_cols):
    if "expect_rows" in case:
        off, col, val = csr
        for r, exp in enumerate(case["expect_rows"]):
            seg = [[int(col[k]), float(val[k])] for k in range(int(off[r]), int(off[r + 1]))]
            assert seg == exp
    for r, c, v in case.get("expect", []):
        assert _dense_get(*csr, r, c) == v
    for r, c in case.get("expect_zero", []):
        lo, hi = int(csr[0][r]), int(csr[0][r + 1])
        assert c not in csr[1][lo:hi]
    if "nnz" in case:
        assert int(csr[0][-1]) == case["nnz"]
    if "expect_dense" in case:
        assert np.array_equal(_dense_of(csr, n_cols), np.array(case["expect_dense"]))
    assert not np.any(csr[2] == 0.0)  # zeros never stored


@pytest.mark.parametrize("case", [c for c in KATS["add"] if "error" not in c], ids=lambda c: c["source"])
def test_add_golden(case):
    a = _trip_csr(case["n_rows"], case["a"])
    b = _trip_csr(case["n_rows"], case["b"])
    c = oracle.add(a, b, subtract=case.get("op") == "-")
    _check_rows(c, case, case["n_cols"])


@pytest.mark.parametrize("case", [c for c in KATS["spgemm"] if "error" not in c], ids=lambda c: c["source"])
def test_spgemm_golden(case):
    a = _trip_csr(case["n_rows"], case["a"])
    b = _trip_csr(case["n_inner"], case["b"])
    c = oracle.spgemm(a, b, case["n_cols"])
    _check_rows(c, case, case["n_cols"])
    if case.get("swap_differs"):
        assert not np.array_equal(_dense_of(oracle.spgemm(b, a, case["n_cols"]), case["n_cols"]),
                                  _dense_of(c, case["n_cols"]))


def _scipy_csr(M):
    M = M.tocsr()
    M.sort_indices()
    return M.indptr.astype(np.uint64), M.indices.astype(np.uint32), M.data.astype(np.float64)


def _rand_sp(n, m, density, seed, zeros=0.0):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    M = sp.random(n, m, density=density, random_state=rng, format="csr")
    M.data = rng.standard_normal(M.data.size)
    if zeros:
        M.data[rng.random(M.data.size) < zeros] = 0.0  # explicit zeros, as load_csr may store
    M.sort_indices()
    return M


def _same_bits(x, y):
    return (np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
            and np.array_equal(np.asarray(x[2]).view(np.uint64), np.asarray(y[2]).view(np.uint64)))


@pytest.mark.parametrize("seed", range(8))
def test_add_and_spgemm_match_scipy(seed):
    """scipy.sparse's CSR kernels (library routines) are bit-identical here:
    csr_binop_csr merges sorted rows and drops results equal to zero;
    csr_matmat accumulates sums[j] += a * b from 0 in A's then B's column
    order and drops sums equal to zero -- the reference's semantics."""
    rng = np.random.default_rng(seed)
    n, k, m = (int(v) for v in rng.integers(1, 400, 3))
    A = _rand_sp(n, k, 0.04, 10 + seed, zeros=0.05)
    B = _rand_sp(k, m, 0.04, 20 + seed)
    A2 = _rand_sp(n, k, 0.04, 30 + seed)
    a, b, a2 = _scipy_csr(A), _scipy_csr(B), _scipy_csr(A2)
    P = (A @ B).tocsr()
    P.eliminate_zeros()
    assert _same_bits(oracle.spgemm(a, b, m), _scipy_csr(P))
    for sub in (False, True):
        S = (A - A2 if sub else A + A2).tocsr()
        S.eliminate_zeros()
        assert _same_bits(oracle.add(a, a2, subtract=sub), _scipy_csr(S))


def test_add_cancellation_and_identities():
    """A + (-A) is empty, A - A is empty, A + 0 = A without its zeros."""
    A = _rand_sp(200, 150, 0.05, 5, zeros=0.1)
    a = _scipy_csr(A)
    neg = (a[0], a[1], -a[2])
    assert int(oracle.add(a, neg)[0][-1]) == 0
    assert int(oracle.add(a, a, subtract=True)[0][-1]) == 0
    z = (np.zeros(201, np.uint64), np.zeros(0, np.uint32), np.zeros(0))
    A.eliminate_zeros()
    assert _same_bits(oracle.add(a, z), _scipy_csr(A))


def _spgemm_reversed_k(a, b, n_cols):
    """A planted mistake: products summed in descending k order."""
    ao, ac, av = a
    bo, bc, bv = b
    rows = []
    for i in range(len(ao) - 1):
        acc = {}
        for ka in reversed(range(int(ao[i]), int(ao[i + 1]))):
            k = int(ac[ka])
            for kb in range(int(bo[k]), int(bo[k + 1])):
                j = int(bc[kb])
                acc[j] = acc.get(j, 0.0) + av[ka] * bv[kb]
        rows.append(sorted((j, v) for j, v in acc.items() if v != 0.0))
    off = np.zeros(len(rows) + 1, np.uint64)
    np.cumsum([len(r) for r in rows], out=off[1:])
    return off, np.array([j for r in rows for j, _ in r], np.uint32), np.array([v for r in rows for _, v in r])


def test_planted_spgemm_order_mistake_is_caught():
    A = _rand_sp(60, 80, 0.3, 7)
    B = _rand_sp(80, 50, 0.3, 8)
    a, b = _scipy_csr(A), _scipy_csr(B)
    P = (A @ B).tocsr()
    P.eliminate_zeros()
    assert _same_bits(oracle.spgemm(a, b, 50), _scipy_csr(P))
    assert not _same_bits(_spgemm_reversed_k(a, b, 50), _scipy_csr(P))
