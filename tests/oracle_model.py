"""Test helpers: an oracle-backed model of spira::matrix and the interpreter of
the lifecycle scenarios in tests/golden/reference_kats.json.

The model drives oracle.dedup / build_csr / merge_csr exactly as
matrix::lock() does (matrix.hpp:277-342); the same scenarios are replayed on
spira_amd.Matrix (device) by the GPU tests.
"""
from __future__ import annotations

import json
import os

import numpy as np

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_kats():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as f:
        return json.load(f)


def stage(n_rows, rows, cols, vals):
    rows = np.asarray(rows, np.int64)
    order = np.argsort(rows, kind="stable")
    ptr = np.zeros(n_rows + 1, np.uint64)
    if rows.size:
        np.cumsum(np.bincount(rows, minlength=n_rows), out=ptr[1:])
    return ptr, np.asarray(cols, np.uint32)[order], np.asarray(vals, np.float64)[order]


class OracleMatrix:
    def __init__(self, n_rows, n_cols, parallel=False):
        self.nr, self.nc = n_rows, n_cols
        self.parallel = parallel
        self.locked = False
        self.ins = []
        self.dirty = np.zeros(n_rows, np.uint8)
        self.csr_ = None

    def is_locked(self):
        return self.locked

    def insert(self, r, c, v):
        if self.locked:
            raise RuntimeError("matrix::insert() requires open mode")
        if not (0 <= r < self.nr) or not (0 <= c < self.nc):
            raise IndexError("out of range")
        self.ins.append((r, c, float(v)))
        self.dirty[r] = 1

    def lock(self):
        if self.locked:
            return
        rows = [t[0] for t in self.ins]
        cols = [t[1] for t in self.ins]
        vals = [t[2] for t in self.ins]
        ptr, c, v = stage(self.nr, rows, cols, vals)
        dirty = None if (self.parallel or self.csr_ is None) else self.dirty
        self.csr_ = oracle.lock(ptr, c, v, old=self.csr_, dirty=dirty)
        self.ins = []
        self.dirty[:] = 0
        self.locked = True

    def open(self):
        self.locked = False

    def clear(self):
        self.ins = []
        self.dirty[:] = 0

    def _committed(self, r, c):
        if self.csr_ is None:
            return None
        off, col, val = self.csr_
        a, b = int(off[r]), int(off[r + 1])
        for k in range(a, b):
            if int(col[k]) == c:
                return float(val[k])
        return None

    def get(self, r, c):
        if not self.locked:
            for rr, cc, vv in reversed(self.ins):
                if rr == r and cc == c:
                    return vv
        v = self._committed(r, c)
        return 0.0 if v is None else v

    def contains(self, r, c):
        if not self.locked and any(rr == r and cc == c for rr, cc, _ in self.ins):
            return True
        return self._committed(r, c) is not None

    def row_nnz(self, r):
        committed = 0 if self.csr_ is None else int(self.csr_[0][r + 1] - self.csr_[0][r])
        return committed + (0 if self.locked else sum(1 for t in self.ins if t[0] == r))

    def nnz(self):
        committed = 0 if self.csr_ is None else int(self.csr_[0][-1])
        return committed + (0 if self.locked else len(self.ins))

    def csr(self):
        return self.csr_


def run_scenario(mat, case):
    """Replay one lifecycle case; assert every expectation it lists."""
    src = case["source"]
    for op in case["ops"]:
        kind = op[0]
        if kind == "insert":
            mat.insert(op[1], op[2], op[3])
        elif kind == "lock":
            mat.lock()
        elif kind == "open":
            mat.open()
        elif kind == "clear":
            mat.clear()
        elif kind == "get":
            got = mat.get(op[1], op[2])
            assert got == op[3], f"{src}: get{tuple(op[1:3])} = {got}, expected {op[3]}"
        elif kind == "contains":
            got = mat.contains(op[1], op[2])
            assert got == op[3], f"{src}: contains{tuple(op[1:3])} = {got}, expected {op[3]}"
        elif kind == "row_nnz":
            got = mat.row_nnz(op[1])
            assert got == op[2], f"{src}: row_nnz({op[1]}) = {got}, expected {op[2]}"
        elif kind == "nnz":
            got = mat.nnz()
            assert got == op[1], f"{src}: nnz() = {got}, expected {op[1]}"
        elif kind == "is_locked":
            assert mat.is_locked() == op[1], f"{src}: is_locked"
        else:
            raise AssertionError(f"unknown op {kind}")


# ---------------------------------------------------------------------------
# Independent brute-force formulations (dict based), used to pin the oracle:
# they restate the semantics as "apply updates to a per-row ordered map",
# not as the reference's reverse+stable_sort / two-pointer merge.
# ---------------------------------------------------------------------------
def brute_dedup(n_rows, rows, cols, vals):
    per = [dict() for _ in range(n_rows)]
    for r, c, v in zip(rows, cols, vals):
        per[int(r)][int(c)] = float(v)  # later writes replace earlier ones
    return per


def brute_lock(n_rows, rows, cols, vals, old=None, dirty=None):
    per = brute_dedup(n_rows, rows, cols, vals)
    out = []
    for r in range(n_rows):
        if old is None:
            m = {c: v for c, v in per[r].items() if abs(v) > 0.0 or v != v}
        else:
            off, col, val = old
            a, b = int(off[r]), int(off[r + 1])
            oldrow = {int(col[k]): float(val[k]) for k in range(a, b)}
            is_dirty = (len(per[r]) > 0) if dirty is None else bool(dirty[r])
            if not is_dirty:
                m = oldrow  # verbatim, zeros included
            else:
                m = dict(oldrow)
                m.update(per[r])
                m = {c: v for c, v in m.items() if not (v == 0.0)}
        out.append(sorted(m.items()))
    off = np.zeros(n_rows + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in out]) if n_rows else []
    col = np.array([c for row in out for c, _ in row], np.uint32)
    val = np.array([v for row in out for _, v in row], np.float64)
    return off, col, val


def random_stream(rng, n_rows, n_cols, n_ins, zero_frac=0.1, col_pool=None):
    rows = rng.integers(0, n_rows, n_ins)
    pool = n_cols if col_pool is None else min(col_pool, n_cols)
    cols = rng.integers(0, pool, n_ins)
    vals = rng.uniform(-1, 1, n_ins)
    z = rng.random(n_ins) < zero_frac
    vals[z] = 0.0
    neg = rng.random(n_ins) < 0.5
    vals[z & neg] = -0.0
    return rows, cols, vals


def csr_equal(a, b):
    """Bit-exact comparison of (offsets, cols, vals)."""
    oa, ca, va = a
    ob, cb, vb = b
    return (np.array_equal(np.asarray(oa, np.uint64), np.asarray(ob, np.uint64))
            and np.array_equal(np.asarray(ca, np.uint32), np.asarray(cb, np.uint32))
            and np.array_equal(np.asarray(va, np.float64).view(np.uint64),
                               np.asarray(vb, np.float64).view(np.uint64)))
