"""Rows built to cancel, and the Dot2 error bound they are held to (see
test_gpu_dot2.py and test_dot2_bound.py)."""
import math

import numpy as np

U = 2.0 ** -24


def cancelling_matrix(rng, n_rows, n_cols, x, max_len=60):
    """CSR (u64 offsets, sorted u32 columns, FP32-exact values) whose rows of
    two or more entries sum to nearly zero against x; also the exact row sums
    and sums of |a x|."""
    lens = rng.integers(0, max_len, size=n_rows)
    off = np.zeros(n_rows + 1, np.uint64)
    cols, vals, exact, absum = [], [], np.zeros(n_rows), np.zeros(n_rows)
    for r, L in enumerate(lens):
        c = np.sort(rng.choice(n_cols, size=int(L), replace=False)).astype(np.uint32)
        v = rng.standard_normal(int(L)).astype(np.float32).astype(np.float64)
        if L >= 2:
            j = int(np.argmax(np.abs(x[c])))  # the cancelling entry: the largest |x|
            rest = math.fsum(np.delete(v * x[c], j))
            v[j] = float(np.float32(-rest / x[c[j]]))
        p = v * x[c]
        exact[r] = math.fsum(p)
        absum[r] = math.fsum(np.abs(p))
        cols.append(c)
        vals.append(v)
    np.cumsum(lens, out=off[1:])
    return (off, np.concatenate(cols) if cols else np.zeros(0, np.uint32),
            np.concatenate(vals) if vals else np.zeros(0)), exact, absum, lens


def dot2_bound(exact, absum, lens):
    n = lens.astype(np.float64) + 64
    gamma = n * U / (1 - n * U)
    return 2 * U * np.abs(exact) + 2 * gamma ** 2 * absum


def assert_dot2_bound(y, exact, absum, lens, what):
    bound = dot2_bound(exact, absum, lens) + 1e-300
    err = np.abs(y - exact)
    bad = np.flatnonzero(err > bound)
    assert bad.size == 0, (what, bad[:5], err[bad[:5]], bound[bad[:5]], absum[bad[:5]])
    # the rows do cancel: plain FP32 summation would be off by ~u sum |a x|
    assert np.median(absum[lens >= 8] / np.maximum(np.abs(exact[lens >= 8]), 1e-300)) > 1e4
