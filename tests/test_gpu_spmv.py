"""Device SpMV (FP64) against the oracle, through the C-ABI.

Tolerance: normwise relative error ||y - y_ref||_2 / ||y_ref||_2 <= 1e-12
against the oracle's sequential FP64 loop, plus the reference stress test's
per-row bound |y - y_ref| <= 1e-10 + 1e-10 * max(|y|, |y_ref|)
(tests/large_matrix_stress_tests.cpp:80-88).  Small integer-valued known
answers must match exactly.
"""
import numpy as np
import pytest

import oracle
import spira_amd
from spira_amd import workloads as W
from oracle_model import load_kats, stage

pytestmark = pytest.mark.gpu

KATS = load_kats()
VARIANTS = sorted(spira_amd.VARIANTS.values())
NORMWISE_TOL = 1e-12


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


def _upload(dev, off, col, val, n_cols):
    A = spira_amd.DeviceCSR(dev, len(off) - 1, n_cols)
    A.upload(off, col, val)
    return A


def _check(y, off, col, val, x):
    y_ref = oracle.spmv_f64(off, col, val, x)
    err = oracle.normwise_rel_err(y, y_ref)
    assert err <= NORMWISE_TOL, err
    tol = 1e-10 + 1e-10 * np.maximum(np.abs(y), np.abs(y_ref))
    assert np.all(np.abs(y - y_ref) <= tol)
    return err


@pytest.mark.parametrize("case", KATS["spmv"], ids=lambda c: c["source"])
@pytest.mark.parametrize("variant", VARIANTS)
def test_spmv_known_answers(dev, case, variant):
    ins = case["inserts"]
    ptr, c, v = stage(case["n_rows"], [t[0] for t in ins], [t[1] for t in ins], [t[2] for t in ins])
    A = spira_amd.DeviceCSR(dev, case["n_rows"], case["n_cols"])
    A.lock(ptr, c, v)
    y = A.spmv(np.array(case["x"]), variant=variant)
    assert y.tolist() == case["y"]


@pytest.mark.parametrize("case", KATS["dot"], ids=lambda c: c["source"])
def test_single_row_dot_known_answers(dev, case):
    nnz = len(case["vals"])
    if len(set(case["cols"])) != nnz:
        pytest.skip("repeated column indices are not a valid CSR row (kernel-level case only)")
    order = np.argsort(case["cols"])
    off = np.array([0, nnz], np.uint64)
    col = np.array(case["cols"], np.uint32)[order]
    val = np.array(case["vals"])[order]
    A = _upload(dev, off, col, val, len(case["x"]))
    for variant in VARIANTS:
        y = A.spmv(np.array(case["x"]), variant=variant)
        assert y[0] == case["expect"]


SHAPES = [
    # (n_rows, n_cols, min_len, max_len, seed)
    (1, 1, 0, 1, 1),
    (7, 5, 0, 5, 2),
    (1000, 1000, 0, 0, 3),          # every row empty
    (3000, 2000, 0, 12, 4),         # many empty rows
    (4096, 4096, 5, 5, 5),          # uniform short rows
    (5000, 70000, 8, 32, 6),        # 2b-like
    (2000, 10000, 60, 70, 7),       # rows just above a wavefront
    (300, 200000, 900, 1100, 8),    # rows around the LDS chunk
]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}_{s[2]}-{s[3]}")
@pytest.mark.parametrize("variant", VARIANTS)
def test_spmv_random_shapes(dev, shape, variant):
    n_rows, n_cols, lo, hi, seed = shape
    off, col, val = W.random_uniform_rows(n_rows, n_cols, lo, hi, seed=seed)
    A = _upload(dev, off, col, val, n_cols)
    x = np.random.default_rng(seed).uniform(-1, 1, n_cols)
    y = A.spmv(x, variant=variant)
    _check(y, off, col, val, x)


def _powerlaw_csr(n, n_cols, seed, long_rows=()):
    rng = np.random.default_rng(seed)
    lens = np.minimum(rng.zipf(1.8, n), 3000).astype(np.int64)
    for i, L in long_rows:
        lens[i] = L
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    cols = np.concatenate([np.sort(rng.choice(n_cols, size=int(L), replace=False)) for L in lens]).astype(np.uint32)
    vals = rng.standard_normal(cols.size)
    return off, cols, vals


@pytest.mark.parametrize("variant", VARIANTS)
def test_spmv_power_law_and_very_long_rows(dev, variant):
    off, col, val = _powerlaw_csr(4000, 100000, 11, long_rows=[(0, 50000), (1777, 20000), (3999, 9000)])
    A = _upload(dev, off, col, val, 100000)
    x = np.random.default_rng(3).uniform(-1, 1, 100000)
    _check(A.spmv(x, variant=variant), off, col, val, x)


def test_spmv_auto_and_device_pointers(dev):
    import torch
    off, col, val = W.random_uniform_rows(20000, 30000, 4, 40, seed=12)
    A = _upload(dev, off, col, val, 30000)
    x = np.random.default_rng(5).uniform(-1, 1, 30000)
    xt = torch.from_numpy(x).cuda()
    yt = torch.full((20000,), 7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    A.spmv(xt, yt, device=True)
    _check(yt.cpu().numpy(), off, col, val, x)
    v = A.choose_variant()
    assert 0 <= v < spira_amd.N_VARIANTS


def test_xphase_many_slices_and_blocks(dev):
    """XPHASE with 16 column slices (x of 24 MB, slices of at most 2 MiB) and 256 row
    blocks of 782 rows: every block walks all slices in its XCD's order and
    adds into its LDS rows; both launch widths share the plan, and a plan is
    reused by a later call."""
    n_rows, n_cols = 200000, 3000000
    off, col, val = W.random_uniform_rows(n_rows, n_cols, 8, 32, seed=21)
    A = _upload(dev, off, col, val, n_cols)
    rng = np.random.default_rng(22)
    for name in ("xphase", "xphase_paced", "xphase"):
        x = rng.uniform(-1, 1, n_cols)
        _check(A.spmv(x, variant=spira_amd.VARIANTS[name]), off, col, val, x)


def test_xphase_more_row_blocks_than_cus(dev):
    """4,294,304 rows (more than 256 blocks of 16,384 rows): the grid is not
    resident at once, so XPHASE_PACED runs its phases unpaced; both widths must
    still match the oracle, the last block holding a partial row range."""
    n_rows, n_cols = (1 << 22) + 100000, 2500000
    off, col, val = W.random_uniform_rows(n_rows, n_cols, 0, 4, seed=23)
    A = _upload(dev, off, col, val, n_cols)
    rng = np.random.default_rng(24)
    for name in ("xphase_paced", "xphase"):
        x = rng.uniform(-1, 1, n_cols)
        _check(A.spmv(x, variant=spira_amd.VARIANTS[name]), off, col, val, x)


def test_spmv_laplacian_full_size_exact_row_sums(dev):
    """2048^2 5-point Laplacian (4,194,304 rows): A @ 1 = 4 - (#neighbours),
    exactly, for every variant; and linearity on random vectors."""
    off, col, val = W.laplacian_2d(2048)
    n = len(off) - 1
    A = _upload(dev, off, col, val, n)
    deg = np.diff(off.astype(np.int64)) - 1
    expect = 4.0 - deg
    ones = np.ones(n)
    rng = np.random.default_rng(0)
    x1, x2 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    for variant in VARIANTS:
        assert np.array_equal(A.spmv(ones, variant=variant), expect), variant
    y1, y2, y12 = A.spmv(x1), A.spmv(x2), A.spmv(x1 + x2)
    assert oracle.normwise_rel_err(y1 + y2, y12) <= 1e-14
    _check(y1, off, col, val, x1)


def test_spmv_errors_follow_reference_order(dev):
    case = KATS["spmv_errors"][0]
    A = spira_amd.DeviceCSR(dev, case["n_rows"], case["n_cols"])  # never locked
    for c in case["cases"]:
        with pytest.raises(spira_amd.InvalidArgument):
            A.spmv(np.ones(c["x_len"]), np.zeros(c["y_len"]))
    with pytest.raises(spira_amd.LogicError):
        A.spmv(np.ones(case["n_cols"]), np.zeros(case["n_rows"]))
    m = spira_amd.Matrix(3, 4, dev)
    with pytest.raises(spira_amd.LogicError):
        spira_amd.spmv(m, np.ones(4), np.zeros(3))
    with pytest.raises(spira_amd.InvalidArgument):
        spira_amd.spmv(m, np.ones(3), np.zeros(3))


def test_spmv_config1_parity(dev):
    """10,000 x 10,000, 10 distinct uniform columns per row, built in open mode
    via insert then lock(); CSR arrays bit-exact and y within tolerance."""
    rows, cols, vals = W.insert_stream_uniform(10000, 10000, 10)
    m = spira_amd.Matrix(10000, 10000, dev)
    m.insert_many(rows, cols, vals)
    m.lock()
    ptr, c, v = stage(10000, rows, cols, vals)
    exp = oracle.lock(ptr, c, v)
    got = m.csr()
    assert all(np.array_equal(a, b) for a, b in zip(got, exp))
    x = W.dense_x(10000)
    y = np.empty(10000)
    spira_amd.spmv(m, x, y)
    _check(y, *exp, x)


def test_device_vectors_from_torch_stream_without_sync(dev):
    """x written by a kernel on torch's current stream, then handed to the
    library (whose handle owns another stream) with no synchronize: the call
    waits for the producer.  Wrong tensors are refused before any launch."""
    import torch
    n = 200_000
    off, col, val = W.random_uniform_rows(n, n, 8, 32, seed=11)
    A = _upload(dev, off, col, val, n)
    torch.cuda.synchronize()
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    for rep in range(3):
        x = (torch.arange(n, device="cuda", dtype=torch.float64) * (0.37 + rep)).sin_()
        x = x * x + torch.cos(x)  # several kernels on torch's stream
        A.spmv(x, y, device=True)
        _check(y.cpu().numpy(), off, col, val, x.cpu().numpy())
    for bad_x, msg in ((x.float(), "float64"), (x.cpu(), "device tensor"), (x[::2], "columns")):
        with pytest.raises(spira_amd.InvalidArgument, match=msg):
            A.spmv(bad_x, y, device=True)
    with pytest.raises(spira_amd.InvalidArgument, match="float64"):
        A.spmv(x, y.float(), device=True)
    with pytest.raises(spira_amd.InvalidArgument, match="float64"):
        A.spmv_mixed(x.float(), y, device=True)


def test_spmv_random_4M_full_size(dev):
    """The HBM-scale random matrix at full size (4,194,304 rows, 8-32 uniform
    distinct columns per row, seed 43): the XPHASE variants -- XPHASE_PACED
    with its whole grid resident and its phases paced -- and AUTO, against the
    oracle at 1e-12 normwise plus the per-row bound."""
    n = 1 << 22
    off, col, val = W.random_uniform_rows(n, n, 8, 32, seed=43)
    A = _upload(dev, off, col, val, n)
    x = W.dense_x(n)
    for variant in (spira_amd.VARIANTS["xphase_paced"], spira_amd.VARIANTS["xphase"], spira_amd.SPMV_AUTO):
        print(f"random 4M: variant {variant}", flush=True)
        _check(A.spmv(x, variant=variant), off, col, val, x)


def test_autotune_pins_the_fastest_and_skips_what_cannot_run(dev):
    """spira_spmv_autotune_f64: every variant is timed, the fastest is pinned
    and AUTO runs it afterwards; a variant whose size limits the matrix
    exceeds (XPHASE slices hold at most 2^18 columns, 256 slices: 67,108,864
    columns) is reported as -1 and never chosen."""
    import torch
    for n_rows, n_cols, seed in ((300_000, 300_000, 31), (3000, 70_000_000, 32)):
        print(f"autotune: {n_rows} x {n_cols}", flush=True)
        off, col, val = W.random_uniform_rows(n_rows, n_cols, 4, 24, seed=seed)
        A = _upload(dev, off, col, val, n_cols)
        xh = np.random.default_rng(seed).uniform(-1, 1, n_cols)
        x = torch.from_numpy(xh).cuda()
        y = torch.empty(n_rows, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        best, ms = A.autotune(x, y, reps=3)
        ran = {v: t for v, t in ms.items() if t >= 0}
        assert best == min(ran, key=ran.get)
        skipped = {v for v, t in ms.items() if t < 0}
        if n_cols > (1 << 26):
            assert skipped == {spira_amd.VARIANTS["xphase"], spira_amd.VARIANTS["xphase_paced"]}
        else:
            assert not skipped
        assert A.choose_variant() == best
        y_auto = A.spmv(x, torch.empty_like(y), device=True).cpu().numpy()
        y_best = A.spmv(x, torch.empty_like(y), variant=best, device=True).cpu().numpy()
        if best not in (spira_amd.VARIANTS["xphase"], spira_amd.VARIANTS["xphase_paced"]):
            assert np.array_equal(y_auto.view(np.uint64), y_best.view(np.uint64))
        _check(y_auto, off, col, val, xh)
        A.close()


def test_reproducible_mode_and_xphase_call_to_call(dev):
    """SPIRA_REPRODUCIBLE: AUTO keeps to the variants whose summation order is
    fixed (the fastest of them by the autotuner's times once it has run), y
    is bit-identical on every call, an explicit XPHASE variant is refused.
    Without it the XPHASE variants may differ between calls in the last
    bits only (normwise within 1e-15 of each other)."""
    import torch
    n = 1 << 20
    off, col, val = W.random_uniform_rows(n, n, 8, 32, seed=61)
    A = _upload(dev, off, col, val, n)
    xh = W.dense_x(n)
    x = torch.from_numpy(xh).cuda()
    ys = []
    for _ in range(3):
        y = torch.empty(n, dtype=torch.float64, device="cuda")
        A.spmv(x, y, device=True, reproducible=True)
        ys.append(y.cpu().numpy())
    for y in ys[1:]:
        assert np.array_equal(y.view(np.uint64), ys[0].view(np.uint64))
    _check(ys[0], off, col, val, xh)
    with pytest.raises(spira_amd.InvalidArgument, match="varying order"):
        A.spmv(x, torch.empty_like(x), variant=spira_amd.VARIANTS["xphase_paced"], device=True, reproducible=True)
    best, ms = A.autotune(x, torch.empty_like(x), reps=3)
    y = torch.empty_like(x)
    A.spmv(x, y, device=True, reproducible=True)
    fixed = {v: t for v, t in ms.items() if t >= 0 and v not in (spira_amd.VARIANTS["xphase"],
                                                                  spira_amd.VARIANTS["xphase_paced"])}
    fastest_fixed = min(fixed, key=fixed.get)
    y_ref = A.spmv(x, torch.empty_like(x), variant=fastest_fixed, device=True)
    assert torch.equal(y, y_ref)
    for name in ("xphase", "xphase_paced"):
        a = A.spmv(x, torch.empty_like(x), variant=spira_amd.VARIANTS[name], device=True).cpu().numpy()
        b = A.spmv(x, torch.empty_like(x), variant=spira_amd.VARIANTS[name], device=True).cpu().numpy()
        assert oracle.normwise_rel_err(a, b) <= 1e-15
        _check(a, off, col, val, xh)


def test_xphase_paced_counters_shared_across_calls_and_kernels(dev):
    """A plan's pacing counters run on across calls (zeroed before its first
    paced call only): FP64, float-instantiation and mixed-precision paced
    calls interleaved on one matrix keep the oracle's results, and no wait
    runs to its spin bound (each would cost milliseconds: the 60 calls here
    take about 10 ms)."""
    import time
    import torch
    n = 2_000_000
    off, col, val = W.random_uniform_rows(n, n, 8, 16, seed=61)
    val = val.astype(np.float32).astype(np.float64)  # FP32-exact, so every path has an exact reference
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(off, col, val)
    xh = np.random.default_rng(62).uniform(-1, 1, n).astype(np.float32).astype(np.float64)
    y_ref = oracle.spmv_f64(off, col, val, xh)
    x64 = torch.from_numpy(xh).cuda()
    x32 = x64.float()
    y64 = torch.empty(n, dtype=torch.float64, device="cuda")
    y32 = torch.empty(n, dtype=torch.float32, device="cuda")
    v = spira_amd.VARIANTS["xphase_paced"]
    A.spmv(x64, y64, variant=v, device=True)  # plan, first paced call
    A.spmv_f32(x32, y32, variant=v, device=True)
    A.spmv_mixed(x64, y64, storage="f32", variant=v, device=True, count=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(20):
        A.spmv(x64, y64, variant=v, device=True)
        assert oracle.normwise_rel_err(y64.cpu().numpy(), y_ref) <= 1e-14, i
        A.spmv_f32(x32, y32, variant=v, device=True)
        assert oracle.normwise_rel_err(y32.cpu().numpy().astype(np.float64), y_ref) <= 1e-7, i
        A.spmv_mixed(x64, y64, storage="f32", variant=v, device=True, count=False)
        assert oracle.normwise_rel_err(y64.cpu().numpy(), y_ref) <= 1e-14, i
    dt = time.perf_counter() - t0
    print(f"60 interleaved paced calls (and their checks): {1e3 * dt:.1f} ms")
    assert dt < 5.0  # the checks dominate; 60 waits run to their bound would add ~0.3 s of device time
    t0 = time.perf_counter()
    for _ in range(20):
        A.spmv(x64, y64, variant=v, device=True, sync=False)
        A.spmv_f32(x32, y32, variant=v, device=True, sync=False)
        A.spmv_mixed(x64, y64, storage="f32", variant=v, device=True, sync=False, count=False)
    A.dev.synchronize()
    dt = time.perf_counter() - t0
    print(f"60 interleaved paced calls: {1e3 * dt:.1f} ms")
    assert dt < 0.15, dt


_GIVE_UP_SCRIPT = r"""
import sys, time
import numpy as np
sys.path.insert(0, {root!r})
import oracle, spira_amd
from spira_amd import workloads as W
n = 1 << 20
off, col, val = W.random_uniform_rows(n, n, 8, 32, seed=71)
A = spira_amd.DeviceCSR(spira_amd.Device(0), n, n)
A.upload(off, col, val)
x = W.dense_x(n)
v = spira_amd.VARIANTS["xphase_paced"]
y = A.spmv(x, variant=v)
t = time.perf_counter()
for _ in range(5):
    y = A.spmv(x, variant=v)
dt = (time.perf_counter() - t) / 5
err = oracle.normwise_rel_err(y, oracle.spmv_f64(off, col, val, x))
print(f"{{err}} {{dt}}")
"""


def test_xphase_paced_waits_that_give_up_stop_pacing():
    """XPHASE_PACED with the pacing wait bound forced to one load from the host
    (SPIRA_PHASE_SPIN_MAX=1, read when the handle is created): waits give up,
    each block stops pacing after its first give-up, the kernel exits
    normally, y matches the oracle and a call stays well under 50 ms."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, SPIRA_PHASE_SPIN_MAX="1")
    out = subprocess.run([sys.executable, "-c", _GIVE_UP_SCRIPT.format(root=root)], env=env, capture_output=True,
                         text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    err, dt = (float(t) for t in out.stdout.split()[-2:])
    print(f"spin max 1: error {err:.2e}, {1e3 * dt:.3f} ms per call")
    assert err <= 1e-12
    assert dt < 0.05
