"""The Dot2 bound check of test_gpu_dot2.py separates compensated from
uncompensated accumulation (CPU only).

The same cancelling rows are summed in numpy FP32 three ways, one lane per
row in entry order: Dot2 as written (TwoProduct by FMA, TwoSum, correction
c += q + e), Dot2 with s + a x fused into one FMA (the compiler's contraction
of the written code), and plain FP32 summation.  The bound must hold for the
first and fail on most cancelling rows for the other two, so a GPU path that
passes it cannot be one of them.
"""
import numpy as np

from dot2_cases import U, cancelling_matrix, dot2_bound

f32 = np.float32


def fma32(a, b, c):
    """FP32 fma: a*b of two FP32 numbers is exact in FP64, so a*b + c rounded
    once to FP64 and then to FP32 differs from the single FP32 rounding only
    in double-rounding ties, which are immaterial to the bound."""
    return f32(float(a) * float(b) + float(c))


def row_dot2(v, x, fused):
    s = f32(0)
    c = f32(0)
    for a, b in zip(v.astype(np.float32), x.astype(np.float32)):
        p = f32(a * b)
        e = fma32(a, b, -p)
        t = fma32(a, b, s) if fused else f32(s + p)
        bp = f32(t - s)
        d = fma32(a, b, -bp) if fused else f32(p - bp)
        q = f32(f32(s - f32(t - bp)) + d)
        c = f32(c + f32(q + e))
        s = t
    return float(f32(s + c))


def row_plain(v, x):
    s = f32(0)
    for a, b in zip(v.astype(np.float32), x.astype(np.float32)):
        s = f32(s + f32(a * b))
    return float(s)


def test_bound_separates_dot2_from_fused_and_plain_sums():
    rng = np.random.default_rng(33)
    n = 3000
    x = rng.uniform(-1, 1, n).astype(np.float32).astype(np.float64)
    (off, col, val), exact, absum, lens = cancelling_matrix(rng, n, n, x)
    y_dot2, y_fused, y_plain = np.zeros(n), np.zeros(n), np.zeros(n)
    for r in range(n):
        a, b = int(off[r]), int(off[r + 1])
        xv = x[col[a:b]]
        y_dot2[r] = row_dot2(val[a:b], xv, fused=False)
        y_fused[r] = row_dot2(val[a:b], xv, fused=True)
        y_plain[r] = row_plain(val[a:b], xv)
    bound = dot2_bound(exact, absum, lens)
    long_rows = lens >= 8
    assert (np.abs(y_dot2 - exact) <= bound).all()
    assert np.mean(np.abs(y_fused - exact)[long_rows] > bound[long_rows]) > 0.5
    assert np.mean(np.abs(y_plain - exact)[long_rows] > bound[long_rows]) > 0.5
    # and the margin is wide: the fused sum is off by about u sum |a x|
    assert np.median(np.abs(y_fused - exact)[long_rows] / absum[long_rows]) > U / 16
