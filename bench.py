#!/usr/bin/env python3
"""Benchmark: plain FP64 CSR SpMV at HBM scale on one MI355X.

Workload (one step = one SpMV on each matrix, inputs resident in HBM):
  * 5-point 2-D Laplacian on a 2048 x 2048 grid (4,194,304 rows, 20,963,328 nnz)
  * 4,194,304-row random matrix, nnz/row uniform in 8..32, uniform distinct
    columns (seed 43), values U(-1, 1)
FP64 values, u32 columns, u32 row offsets on the device; x U(-1, 1).

value = whole-step algorithmic GB/s = (B_lap + B_rand) / step time, with
B = nnz*(8+4) + (n_rows+1)*4 + n_cols*8 + n_rows*8 (x read once, y written
once; the reference's traffic model, BENCHMARKS.md:173-187, at device widths).

Prints one JSON line (the last line of stdout).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="bound on the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-lock", action="store_true", help="skip the lock() measurement")
    p.add_argument("--variant", default="autotune", help="autotune | auto | <variant name>")
    p.add_argument("--matrices", default="laplacian_2048,random_4M_8to32",
                   help="comma list of the SpMV matrices to run (profiling runs pick one)")
    p.add_argument("--lock-only", action="store_true", help="only the lock() measurement (profiling)")
    p.add_argument("--no-transpose", action="store_true", help="skip the transpose measurement")
    p.add_argument("--transpose-only", action="store_true", help="only the transpose measurement (profiling)")
    p.add_argument("--no-algebra", action="store_true", help="skip the addition / SpGEMM measurement")
    p.add_argument("--algebra-only", action="store_true", help="only the addition / SpGEMM measurement (profiling)")
    p.add_argument("--calib-copy-mb", type=int, default=0,
                   help="also run a plain device copy of this many MiB (PMC traffic calibration)")
    p.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision power-law section")
    p.add_argument("--no-batch", action="store_true", help="skip the batched / learned-selector section")
    p.add_argument("--mixed-rows", type=int, default=1 << 23)
    p.add_argument("--mixed-reps", type=int, default=50)
    p.add_argument("--recompute-ratio", type=float, default=16.0,
                   help="cancellation-ratio threshold of the FP64 row recompute in the mixed section")
    p.add_argument("--no-auto", action="store_true", help="skip the AUTO kernel-selector section")
    p.add_argument("--auto-seeds", type=int, default=100)
    p.add_argument("--auto-seed", type=int, default=960_000)
    p.add_argument("--batch-mats", type=int, default=4096)
    p.add_argument("--batch-seed", type=int, default=20_000_000,
                   help="first seed of the batch (disjoint from the selector's training and held-out seeds)")
    p.add_argument("--batch-reps", type=int, default=64)
    p.add_argument("--no-batch-exhaustive", dest="batch_exhaustive", action="store_false",
                   help="skip the exhaustive autotune of the whole batch (the selector's judge)")
    return p.parse_args()


def profile_traffic(workload_key):
    """HBM bytes per launch of the dominant kernel, from the committed rocprof
    PMC summary (profiles/spmv_traffic.json), if it covers this workload."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        e = d.get(workload_key)
        return None if e is None else e.get("traffic_bytes")
    except (OSError, ValueError):
        return None


def run_lock(args, np, torch, spira_amd, W, dev, stream):
    """lock(): re-lock merge on 1,048,576^2 with 16 random nnz/row, then 2
    inserts/row of which 25% overwrite; staged inputs resident in HBM (and,
    separately, handed over from host memory: PCIe-inclusive)."""
    n3 = 1 << 20
    (o3, c3, v3), (r3, ic3, iv3) = W.relock_workload(n3)
    sp3, sc3, sv3 = spira_amd.stage_from_coo(n3, r3, ic3, iv3)
    L = spira_amd.DeviceCSR(dev, n3, n3)
    with torch.cuda.stream(stream):
        d_sp = torch.from_numpy(sp3.astype(np.int64)).to("cuda")
        d_sc = torch.from_numpy(sc3.astype(np.int32)).to("cuda")
        d_sv = torch.from_numpy(sv3).to("cuda")
    stream.synchronize()
    t_dev, t_host = [], []
    nnz_new = 0
    for rep in range(6):
        L.upload(o3, c3, v3)
        dev.synchronize()
        with torch.cuda.stream(stream):  # inputs were produced on the library's stream
            t = time.perf_counter()
            nnz_new = L.lock(d_sp, d_sc, d_sv, device=True)
            t_dev.append(time.perf_counter() - t)
        if args.lock_only:
            continue
        L.upload(o3, c3, v3)
        dev.synchronize()
        t = time.perf_counter()
        L.lock(sp3, sc3, sv3)
        t_host.append(time.perf_counter() - t)
    ref_csr = L.download()
    # the same inserts as a device COO stream: append + sort by row + lock
    st = spira_amd.Stage(dev, n3, n3)
    with torch.cuda.stream(stream):
        d_r = torch.from_numpy(r3.astype(np.int32)).to("cuda")
        d_c = torch.from_numpy(ic3.astype(np.int32)).to("cuda")
        d_v = torch.from_numpy(iv3).to("cuda")
    stream.synchronize()
    t_stg = []
    for rep in range(6):
        L.upload(o3, c3, v3)
        dev.synchronize()
        with torch.cuda.stream(stream):
            t = time.perf_counter()
            st.append(d_r, d_c, d_v, device=True)
            L.lock_staged(st)
            t_stg.append(time.perf_counter() - t)
    staged_same = all(np.array_equal(a.view(np.uint8), b.view(np.uint8)) for a, b in zip(L.download(), ref_csr))
    st.close()
    # lock() then SpMV on the result (AUTO, no autotune): the first call pays
    # the row statistics and the selector and runs a plan-free variant; the
    # second builds the chosen variant's plan
    x3 = torch.from_numpy(W.dense_x(n3)).to("cuda")
    y3 = torch.empty(n3, dtype=torch.float64, device="cuda")
    t_chain, t_first, t_second, t_steady, auto_v = [], [], [], [], None
    for rep in range(4):
        L.upload(o3, c3, v3)
        dev.synchronize()
        with torch.cuda.stream(stream):
            t = time.perf_counter()
            L.lock(d_sp, d_sc, d_sv, device=True)
            t1 = time.perf_counter()
            L.spmv(x3, y3, device=True, sync=True)
            t2 = time.perf_counter()
            L.spmv(x3, y3, device=True, sync=True)
            t2b = time.perf_counter()
            for _ in range(10):
                L.spmv(x3, y3, device=True, sync=False)
            stream.synchronize()
            t3 = time.perf_counter()
        t_chain.append(t2 - t)
        t_first.append(t2 - t1)
        t_second.append(t2b - t2)
        t_steady.append((t3 - t2b) / 10)
        auto_v = L.choose_variant()
    best3, ms3 = L.autotune(x3, y3, reps=5)
    fixed = {v: ms for v, ms in ms3.items() if ms >= 0 and spira_amd.VARIANT_NAMES[v] not in ("xphase", "xphase_paced")}
    v_fix = min(fixed, key=fixed.get)
    t_ph = min(ms3[spira_amd.VARIANTS["xphase"]], ms3[spira_amd.VARIANTS["xphase_paced"]])
    first_ms = 1e3 * float(np.median(t_first[1:]))
    second_ms = 1e3 * float(np.median(t_second[1:]))
    steady_ms = 1e3 * float(np.median(t_steady[1:]))
    relock_spmv = {"relock_then_spmv_ms": round(1e3 * float(np.median(t_chain[1:])), 3),
                   "first_spmv_after_lock_ms": round(first_ms, 3),
                   "second_spmv_after_lock_ms": round(second_ms, 3),
                   "steady_spmv_ms": round(steady_ms, 4),
                   "auto_variant": spira_amd.VARIANT_NAMES[auto_v],
                   "best_fixed_order_variant": spira_amd.VARIANT_NAMES[v_fix],
                   "best_fixed_order_ms": round(fixed[v_fix], 4),
                   "xphase_ms": round(t_ph, 4)}
    if fixed[v_fix] > t_ph:
        # calls after which XPHASE's plan (paid by the second AUTO call after a
        # lock) has paid for itself
        relock_spmv["xphase_plan_break_even_calls"] = round((second_ms - steady_ms) / (fixed[v_fix] - t_ph), 1)
    lbytes = W.lock_bytes(n3, int(c3.size), int(sc3.size), nnz_new)
    tdev = float(np.median(t_dev[1:]))
    detail = {"workload": "re-lock merge 1,048,576^2: 16 random nnz/row, then 2 inserts/row, 25% overwrite",
              "nnz_old": int(c3.size), "staged": int(sc3.size), "nnz_new": int(nnz_new),
              "bytes_per_lock": lbytes, "ms_device_inputs": round(1e3 * tdev, 3),
              "GB_s": round(lbytes / tdev / 1e9, 1),
              "ms_device_coo_stream": round(1e3 * float(np.median(t_stg[1:])), 3),
              "device_coo_stream_equals_segments": bool(staged_same)}
    detail.update(relock_spmv)
    if t_host:
        detail["ms_host_inputs_pcie_inclusive"] = round(1e3 * float(np.median(t_host[1:])), 3)
    L.close()
    return detail, (o3, c3, v3, sp3, sc3, sv3, n3, lbytes)


def run_transpose(args, np, torch, spira_amd, W, dev, stream, src=None):
    """transpose (serial/transpose.hpp:35-157) of the 1,048,576^2 matrix with 16
    random nnz/row, resident in HBM; ms per call from HIP events on the
    library's stream around 10 back-to-back calls (each call builds a new A^T,
    allocation included, and ends with one readback of its nnz)."""
    if src is None:
        (o, c, v), _ = W.relock_workload(1 << 20)
    else:
        o, c, v = src
    n = len(o) - 1
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(o, c, v)
    T = spira_amd.DeviceCSR(dev, n, n)
    dev.synchronize()
    detail = {"workload": "transpose of 1,048,576^2 with 16 random nnz/row (serial form: explicit zeros kept)",
              "nnz": int(c.size), "bytes_per_call": W.transpose_bytes(n, n, int(c.size))}
    for key, drop in (("serial", False), ("drop_zeros", True)):
        for _ in range(2):
            A.transpose(T, drop_zeros=drop)
        dev.synchronize()
        ms = _time_calls(torch, stream, lambda: A.transpose(T, drop_zeros=drop), 10)
        detail[f"ms_{key}"] = round(ms, 3)
    detail["GB_s"] = round(detail["bytes_per_call"] / (detail["ms_serial"] * 1e-3) / 1e9, 1)
    A.transpose(T)
    got = T.download()
    A.close()
    T.close()
    return detail, (o, c, v, got)


def run_algebra(args, np, torch, spira_amd, W, dev, stream, base=None):
    """Matrix addition (serial/matrix_addition.hpp:88-112) and SpGEMM
    (serial/spgemm.hpp:12-50) on device-resident matrices: A + B for two
    1,048,576^2 matrices with 16 random nnz/row, A * A for the same A (268M
    products) and for the 2048^2 5-point Laplacian.  ms per call from HIP
    events around back-to-back calls (allocation and the readbacks of the
    output size included)."""
    if base is None:
        (base, _) = W.relock_workload(1 << 20)
    n = len(base[0]) - 1
    other, _ = W.relock_workload(n, seed=W.BASE_SEED + 7)
    lap = W.laplacian_2d(2048)
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(*base)
    B = spira_amd.DeviceCSR(dev, n, n)
    B.upload(*other)
    L = spira_amd.DeviceCSR(dev, len(lap[0]) - 1, len(lap[0]) - 1)
    L.upload(*lap)
    C = spira_amd.DeviceCSR(dev, n, n)
    CL = spira_amd.DeviceCSR(dev, L.n_rows, L.n_cols)
    dev.synchronize()
    out = {}
    A.add(B, C)
    dev.synchronize()
    ms = _time_calls(torch, stream, lambda: A.add(B, C), 10)
    nnz_c = C.nnz
    byt = 2 * 8 * (n + 1) + 12 * (int(base[1].size) + int(other[1].size)) + 8 * (n + 1) + 12 * nnz_c
    out["add"] = {"workload": "A + B, 1,048,576^2, 16 random nnz/row each", "nnz_a": int(base[1].size),
                  "nnz_b": int(other[1].size), "nnz_c": int(nnz_c), "ms": round(ms, 3), "bytes_per_call": byt,
                  "GB_s": round(byt / (ms * 1e-3) / 1e9, 1)}
    for key, M, Cm, src in (("spgemm_random_1m", A, C, base), ("spgemm_laplacian_2048", L, CL, lap)):
        M.spgemm(M, Cm)
        dev.synchronize()
        ms = _time_calls(torch, stream, lambda: M.spgemm(M, Cm), 3)
        blen = np.diff(src[0].astype(np.int64))
        prods = int(blen[src[1]].sum())
        out[key] = {"workload": ("A * A, 1,048,576^2 with 16 random nnz/row" if key == "spgemm_random_1m"
                                 else "A * A, 2048^2 5-point Laplacian"),
                    "products": prods, "nnz_c": int(Cm.nnz), "ms": round(ms, 3),
                    "GFLOP_s": round(2 * prods / (ms * 1e-3) / 1e9, 1)}
    for M in (A, B, L, C, CL):
        M.close()
    return out, (base, other, lap)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


XPHASE_ROW_BLOCK = 16384  # rows per XPHASE row block (the plan's block: 128 KiB of FP64 y in LDS)
X_LINE_DOUBLES = 16       # FP64 elements per 128-byte x line


def xphase_ceiling(np, torch, dev, mat, kern_ms):
    """What the XPHASE kernels' requests can cost at best on this GPU.
    They issue two kinds from one per-CU memory pipeline: one L1-to-L2
    request per distinct x line per row block (the entries of a block's
    (block, slice) group are sorted by column, so lanes on one line share a
    request; counted exactly here) and the entry stream (12 bytes per entry,
    y written once, x brought into each of the 8 XCDs' L2).  Counters on the
    part show the two queue behind one another in that pipeline (the same
    requests in flight per CU with and without the gathers), so the ceiling
    is the sum of their times at the probed rates."""
    off, col, _ = mat
    n_rows, n_cols, nnz = len(off) - 1, int(col.max()) + 1 if len(col) else 1, len(col)
    g_rate, s_rate = dev.probe_ceilings()
    o = torch.from_numpy(np.asarray(off, np.int64)).cuda()
    rows = torch.repeat_interleave(torch.arange(n_rows, device="cuda"), o[1:] - o[:-1])
    c = torch.from_numpy(np.asarray(col, np.int64)).cuda()
    key = (rows // XPHASE_ROW_BLOCK) * (n_cols // X_LINE_DOUBLES + 1) + c // X_LINE_DOUBLES
    lines = int(torch.unique(key).numel())
    del o, rows, c, key
    stream_bytes = 12 * nnz + 8 * n_rows + 8 * 8 * n_cols
    t_g = lines / g_rate * 1e3
    t_s = stream_bytes / s_rate * 1e3
    return {"model": "x-line requests at the probed L2 gather rate + entry stream at the probed read rate",
            "gather_Greq_s": round(g_rate / 1e9, 1), "stream_read_TB_s": round(s_rate / 1e12, 3),
            "x_line_requests": lines, "requests_per_entry": round(lines / max(nnz, 1), 4),
            "stream_bytes": stream_bytes, "gather_ms": round(t_g, 4), "stream_ms": round(t_s, 4),
            "ceiling_ms": round(t_g + t_s, 4), "frac_of_ceiling": round((t_g + t_s) / kern_ms, 4)}


def _physical_cpus():
    """One CPU per physical core among the CPUs this process may run on (the
    lowest-numbered SMT sibling of each core), as bench/run_benchmark.sh
    pins its threads away from SMT siblings."""
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except AttributeError:
        allowed = list(range(os.cpu_count() or 1))
    seen, out = set(), []
    for c in allowed:
        key = None
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/thread_siblings_list") as f:
                key = f.read().strip()
        except OSError:
            key = str(c)
        if key not in seen:
            seen.add(key)
            out.append(c)
    return out


def _llc_bytes():
    try:
        v = os.sysconf("SC_LEVEL3_CACHE_SIZE")
        if v and v > 0:
            return int(v)
    except (ValueError, OSError):
        pass
    try:
        with open("/sys/devices/system/cpu/cpu0/cache/index3/size") as f:
            t = f.read().strip().upper()
        mult = {"K": 1 << 10, "M": 1 << 20, "G": 1 << 30}
        return int(t[:-1]) * mult[t[-1]] if t[-1] in mult else int(t)
    except (OSError, ValueError, KeyError):
        return 32 << 20


def _time_calls(torch, stream, fn, reps):
    """Mean ms per call of fn() over reps back-to-back calls on `stream`."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def _nrm_err(torch, y, y_ref):
    d = torch.linalg.vector_norm((y.double() - y_ref.double()))
    r = torch.linalg.vector_norm(y_ref.double())
    return float(d / r) if float(r) > 0 else float(d)


def run_mixed(args, np, torch, spira_amd, W, dev, stream):
    """Mixed-precision SpMV on power-law rows: 8,388,608 rows, lengths
    Zipf(1.8) clipped to [1, 50,000], 70% of columns within +-4096 of the
    diagonal, values N(0, 1).  FP64 plain path vs FP32 / BF16 / split-BF16
    value storage with compensated FP32 accumulation, each with its error
    against the FP64 result."""
    n = args.mixed_rows
    t0 = time.time()
    with torch.cuda.stream(stream):
        off, col, val = W.powerlaw_rows_device(n, device="cuda")
        A = spira_amd.DeviceCSR(dev, n, n)
        A.upload(off, col, val, device=True)
        x = torch.from_numpy(W.dense_x(n)).to("cuda")
        y64 = torch.empty(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(y64)
    stream.synchronize()
    nnz = A.nnz
    log(f"bench: mixed: {n} rows, {nnz} nnz generated and uploaded in {time.time() - t0:.1f}s")
    reps = args.mixed_reps
    auto_v = A.choose_variant()  # AUTO (learned selector), before the autotuner pins a variant
    best, at = A.autotune(x, y64, reps=3)
    A.spmv(x, y64, device=True, sync=True)
    t64 = _time_calls(torch, stream, lambda: A.spmv(x, y64, device=True, sync=False), reps)
    b64 = W.spmv_bytes(n, n, nnz, s_val=8)
    out = {"workload": f"power-law rows: {n} rows, Zipf(1.8) lengths clipped to [1, 50000], 70% of columns "
                       f"within +-4096 of the diagonal, values N(0,1), x U(-1,1)",
           "n_rows": n, "nnz": nnz, "mean_row": round(nnz / n, 2),
           "fp64": {"variant": spira_amd.VARIANT_NAMES[best], "ms": round(t64, 4), "bytes_per_call": b64,
                    "GB_s": round(b64 / t64 / 1e6, 1), "GFLOP_s": round(2 * nnz / t64 / 1e6, 1),
                    "auto_variant": spira_amd.VARIANT_NAMES[auto_v],
                    "auto_fraction_of_autotuned": round(at[best] / at[auto_v], 4)},
           "storage": {}}
    sval = {"f32": 4, "bf16": 2, "split_bf16": 4}
    for st in ("f32", "bf16", "split_bf16"):
        # fastest of the mixed kernels for this storage (the FP64 recompute
        # needs a row-major kernel)
        times = {}
        for v in spira_amd.MIXED_VARIANTS:
            A.spmv_mixed(x, y, storage=st, variant=v, device=True, sync=True, count=False)
            times[v] = _time_calls(torch, stream, lambda: A.spmv_mixed(x, y, storage=st, variant=v, device=True,
                                                                       sync=False, count=False), 5)
        bts = W.spmv_bytes(n, n, nnz, s_val=sval[st])
        rec = {"bytes_per_call": bts}
        for ratio in (0.0, args.recompute_ratio):
            allowed = spira_amd.MIXED_VARIANTS if ratio == 0.0 else spira_amd.ROW_MAJOR_MIXED_VARIANTS
            v = min(allowed, key=times.get)
            _, nrec = A.spmv_mixed(x, y, storage=st, variant=v, recompute_ratio=ratio, device=True, sync=True)
            err = _nrm_err(torch, y, y64)
            t = _time_calls(torch, stream, lambda: A.spmv_mixed(x, y, storage=st, variant=v, recompute_ratio=ratio,
                                                                device=True, sync=False, count=False), reps)
            rec["recompute_off" if ratio == 0.0 else f"recompute_ratio_{ratio:g}"] = {
                "variant": spira_amd.VARIANT_NAMES[v], "ms": round(t, 4), "GB_s": round(bts / t / 1e6, 1), "GFLOP_s": round(2 * nnz / t / 1e6, 1),
                "speedup_vs_fp64": round(t64 / t, 3), "rows_recomputed": nrec,
                "normwise_rel_err_vs_fp64": err}
        out["storage"][st] = rec
        log(f"bench: mixed {st}: {rec}")
    host = None
    if not args.no_cpu:
        host = (off.cpu().numpy().astype(np.uint64), col.cpu().numpy().astype(np.uint32), val.cpu().numpy(),
                x.cpu().numpy())
    del off, col, val
    return out, (A, x, y64, y, host)


def run_auto_selector(args, np, torch, spira_amd, dev, stream):
    """AUTO (the learned kernel selector, no autotune call) against the
    exhaustive autotuner on fresh corpus seeds (disjoint from the training
    and held-out seeds), next to the row-length rule it replaces; plus the
    500-seed held-out score recorded when the model was trained."""
    from spira_amd import spmv_selector as S
    from spira_amd import workloads as W
    auto, rule, T, fams = [], [], [], []
    t0 = time.time()
    with torch.cuda.stream(stream):
        for seed in range(args.auto_seed, args.auto_seed + args.auto_seeds):
            off, col, val, n_cols, fam = W.spmv_corpus_matrix(seed, "cuda", 1 << 23)
            n = off.numel() - 1
            A = spira_amd.DeviceCSR(dev, n, n_cols)
            A.upload(off, col, val, device=True)
            rule.append(S.rule_variant(*S.row_stats(off)))
            del off, col, val
            auto.append(A.choose_variant())
            x = torch.rand(n_cols, dtype=torch.float64, device="cuda")
            y = torch.empty(n, dtype=torch.float64, device="cuda")
            _, ms = A.autotune(x, y, reps=3)
            T.append([ms[v] if ms[v] >= 0 else np.nan for v in range(spira_amd.N_VARIANTS)])
            fams.append(fam)
            A.close()
    T = np.array(T)
    rows = np.arange(len(auto))
    best = np.nanmin(T, axis=1)
    f_auto, f_rule = S.fraction_of_best(T, auto), S.fraction_of_best(T, rule)
    with open(S.MODEL_PATH) as f:
        trained = json.load(f)
    out = {"workload": f"{len(auto)} corpus matrices (seeds {args.auto_seed}..{args.auto_seed + len(auto) - 1}; "
                       f"uniform, banded, stencil, power-law, block-diagonal, mostly-local; up to 2^23 rows)",
           "auto_mean_fraction_of_best": round(float(f_auto.mean()), 4),
           "auto_total_time_fraction": round(float(best.sum() / T[rows, auto].sum()), 4),
           "rule_mean_fraction_of_best": round(float(f_rule.mean()), 4),
           "rule_total_time_fraction": round(float(best.sum() / T[rows, rule].sum()), 4),
           "auto_fraction_below_0.9": round(float((f_auto < 0.9).mean()), 4),
           "per_family_auto": {fm: round(float(f_auto[[f == fm for f in fams]].mean()), 4) for fm in sorted(set(fams))},
           "heldout_500_at_training": {k: trained[k] for k in ("heldout_mean_fraction_of_best",
                                                              "heldout_total_time_fraction", "corpus")},
           "seconds": round(time.time() - t0, 1)}
    log(f"bench: auto selector: {out}")
    return out


def _batch_x(torch, n_cols, k, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return (torch.rand((n_cols, k), generator=g, device="cuda", dtype=torch.float32) * 2 - 1).contiguous()


def run_batch(args, np, torch, spira_amd, W, dev, stream):
    """Batched SpMV with learned scheduling: args.batch_mats square matrices
    (rows log-uniform in [256, 65536], nnz/row a 50/50 mix of uniform 4..64
    and Zipf(1.6) clipped to min(n, 1024), FP32 values N(0, 1)), k = 1, 4, 16,
    32 right-hand sides, one persistent launch per product replayed from a
    hipGraph, against the plain FP64 path (the drop-in spira_spmv_f64 on the
    same matrices as one block-diagonal FP64 CSR, one call per right-hand
    side)."""
    from spira_amd import selector as S
    seed0 = args.batch_seed
    t0 = time.time()
    mats, offs, cols, vals = [], [], [], []
    with torch.cuda.stream(stream):
        for i in range(args.batch_mats):
            o, c, v = W.batch_matrix(seed0 + i, device="cuda")
            nr = o.numel() - 1
            A = spira_amd.DeviceCSR(dev, nr, nr)
            A.upload(o, c, v, device=True)
            mats.append(A)
            offs.append(o)
            cols.append(c)
            vals.append(v)
        # block-diagonal FP64 copy for the plain path
        rows = np.array([A.n_rows for A in mats], np.int64)
        nnzs = np.array([A.nnz for A in mats], np.int64)
        rb = np.concatenate([[0], np.cumsum(rows)])
        nb = np.concatenate([[0], np.cumsum(nnzs)])
        big_off = torch.cat([o[:-1] + int(nb[i]) for i, o in enumerate(offs)]
                            + [torch.tensor([int(nb[-1])], dtype=torch.int64, device="cuda")])
        big_col = torch.cat([c + int(rb[i]) for i, c in enumerate(cols)])
        big_val = torch.cat(vals)
        del offs, cols, vals
        N = int(rb[-1])
        P = spira_amd.DeviceCSR(dev, N, N)
        P.upload(big_off, big_col, big_val, device=True)
        del big_off, big_col, big_val
        batch = spira_amd.Batch(dev, mats, "f32")
    stream.synchronize()
    nnz = int(batch.nnz)
    n_cols = np.array([A.n_cols for A in mats])
    log(f"bench: batch: {len(mats)} matrices, {N} rows, {nnz} nnz built in {time.time() - t0:.1f}s")
    col_bytes = 2 if max(A.n_cols for A in mats) <= 65536 else 4
    sel = None
    if os.path.exists(S.MODEL_PATH):
        try:
            sel = S.Selector.load(dev)
        except ValueError as e:
            log(f"bench: selector model not used: {e}")
    t1 = time.time()
    feats = S.features_by_k(mats)
    choice = {k: sel.predict(feats[k], [k] * len(mats), n_cols) for k in S.K_LIST} if sel is not None else {}
    t_select = time.time() - t1
    out = {"workload": f"{len(mats)} square matrices (seeds {seed0}..{seed0 + len(mats) - 1}), rows log-uniform in "
                       f"[256, 65536], nnz/row 50/50 uniform 4..64 / Zipf(1.6) clipped to min(n, 1024), "
                       f"FP32 values N(0,1), X U(-1,1) FP32",
           "n_matrices": len(mats), "rows": N, "nnz": nnz,
           "plain_fp64_path": "spira_spmv_f64 (autotuned variant) on the block-diagonal FP64 CSR, one call per "
                              "right-hand side",
           "batched_path": "spira_batch: FP32 values, FP32 X/Y, Dot2 compensated accumulation, one persistent "
                           "work-queue launch per product, replayed from a hipGraph",
           "per_k": {}}
    y64 = torch.empty(N, dtype=torch.float64, device="cuda")
    x0 = torch.empty(N, dtype=torch.float64, device="cuda").uniform_(-1, 1)
    best_v, _ = P.autotune(x0, y64, reps=2)
    out["plain_fp64_variant"] = spira_amd.VARIANT_NAMES[best_v]
    tot_flops, tot_t_batch, tot_t_plain = 0.0, 0.0, 0.0
    worst_err = 0.0
    for k in S.K_LIST:
        X = _batch_x(torch, N, k, 1000 + k)
        Y = torch.empty((N, k), dtype=torch.float32, device="cuda")
        xs = [X[:, j].double().contiguous() for j in range(k)]
        ys = [torch.empty(N, dtype=torch.float64, device="cuda") for _ in range(k)]
        stream.synchronize()
        reps = max(3, args.batch_reps // k)

        def plain():
            for j in range(k):
                P.spmv(xs[j], ys[j], device=True, sync=False)

        plain()
        t_plain = _time_calls(torch, stream, plain, reps)
        rec = {"plain_fp64_ms": round(t_plain, 4), "plain_fp64_GFLOP_s": round(2 * nnz * k / t_plain / 1e6, 1)}
        scheds = {"heuristic": None}
        if args.batch_exhaustive:
            # exhaustive autotune of the whole batch (in-batch costs per matrix)
            t1 = time.time()
            T = S.exhaustive(batch, k, n_cols, X, Y, reps=1)
            rec["exhaustive_autotune_s"] = round(time.time() - t1, 2)
            scheds["exhaustive"] = np.nanargmin(T, axis=1)
            if sel is not None:
                rec["selector_cost_fraction_of_best"] = round(float(np.nansum(np.nanmin(T, 1)) / np.nansum(
                    T[np.arange(len(mats)), choice[k]])), 4)
        if sel is not None:
            scheds["selector"] = choice[k]
        for name, idx in scheds.items():
            if idx is None:
                batch.set_schedule(k)
            else:
                S.apply(batch, k, idx, n_cols)
            batch.capture(X, Y)
            batch.replay(sync=True)
            t = _time_calls(torch, stream, lambda: batch.replay(sync=False), reps)
            rec[f"{name}_ms"] = round(t, 4)
            rec[f"{name}_GFLOP_s"] = round(2 * nnz * k / t / 1e6, 1)
            # bytes at the batch's device widths: FP32 values, u16 columns
            # (every matrix has at most 65,536 columns), u32 row offset and
            # u32 row id per stored row, FP32 X and Y
            rec[f"{name}_GB_s"] = round(W.spmv_bytes(N, N, nnz, s_val=4, s_col=col_bytes, s_ptr=8, s_x=4, s_y=4,
                                                     k=k) / t / 1e6, 1)
        # error of the last schedule's result against the FP64 path
        batch.replay(sync=True)
        plain()
        stream.synchronize()
        Y64 = torch.stack(ys, dim=1)
        err = _nrm_err(torch, Y, Y64)
        worst_err = max(worst_err, err)
        t_b = rec.get("selector_ms", rec["heuristic_ms"])
        rec["speedup_vs_plain_fp64"] = round(t_plain / t_b, 3)
        if "exhaustive_ms" in rec:
            for name in ("selector", "heuristic"):
                if f"{name}_ms" in rec:
                    rec[f"{name}_fraction_of_exhaustive"] = round(rec["exhaustive_ms"] / rec[f"{name}_ms"], 4)
        rec["normwise_rel_err_vs_fp64"] = err
        out["per_k"][str(k)] = rec
        tot_flops += 2.0 * nnz * k
        tot_t_batch += t_b
        tot_t_plain += t_plain
        log(f"bench: batch k={k}: {rec}")
        del X, Y, xs, ys, Y64
    out["GFLOP_s"] = round(tot_flops / tot_t_batch / 1e6, 1)
    out["plain_fp64_GFLOP_s"] = round(tot_flops / tot_t_plain / 1e6, 1)
    out["speedup_vs_plain_fp64"] = round(tot_t_plain / tot_t_batch, 3)
    out["normwise_rel_err_vs_fp64_max"] = worst_err
    out["schedule"] = "selector" if sel is not None else "heuristic"
    if sel is not None:
        out["selector_features_and_inference_s"] = round(t_select, 3)
    batch.close()
    P.close()
    for A in mats:
        A.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (got --gpus {args.gpus}); "
              "multi-GPU sharding is a design note only.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import spira_amd
    from spira_amd import workloads as W

    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(3)

    stream = torch.cuda.Stream()
    dev = spira_amd.Device(0, stream=stream.cuda_stream)

    if args.calib_copy_mb:
        # A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE per
        # dispatch calibrate the PMC byte counters against true bytes.
        nb = args.calib_copy_mb << 20
        with torch.cuda.stream(stream):
            a = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
            b = torch.empty_like(a)
            for _ in range(3):
                b.copy_(a)
        stream.synchronize()
        log(f"bench: calibration copy of {args.calib_copy_mb} MiB done")
        del a, b

    if args.lock_only:
        out = run_lock(args, np, torch, spira_amd, W, dev, stream)
        print(json.dumps({"lock": out[0]}), flush=True)
        return
    if args.transpose_only:
        out = run_transpose(args, np, torch, spira_amd, W, dev, stream)
        print(json.dumps({"transpose": out[0]}), flush=True)
        return
    if args.algebra_only:
        out = run_algebra(args, np, torch, spira_amd, W, dev, stream)
        print(json.dumps({"algebra": out[0]}), flush=True)
        return

    generators = {
        "laplacian_2048": lambda: W.laplacian_2d(2048),
        "random_4M_8to32": lambda: W.random_uniform_rows(1 << 22, 1 << 22, 8, 32, seed=43),
    }
    t0 = time.time()
    wanted = [m for m in args.matrices.split(",") if m]
    log(f"bench: generating {', '.join(wanted)}")
    mats = {name: generators[name]() for name in wanted}
    log(f"bench: generated in {time.time() - t0:.1f}s")
    dA, xs, ys, info = {}, {}, {}, {}
    for name, (off, col, val) in mats.items():
        n = len(off) - 1
        A = spira_amd.DeviceCSR(dev, n, n)
        A.upload(off, col, val)
        x_host = W.dense_x(n)
        with torch.cuda.stream(stream):
            xs[name] = torch.from_numpy(x_host).to("cuda", non_blocking=False)
            ys[name] = torch.empty(n, dtype=torch.float64, device="cuda")
        stream.synchronize()
        dA[name] = A
        nnz = int(col.size)
        info[name] = {"n_rows": n, "nnz": nnz,
                      "bytes": W.spmv_bytes(n, n, nnz, s_val=8, s_col=4, s_ptr=4, s_x=8, s_y=8),
                      "x_host": x_host}

    # XPHASE plan build (once per locked matrix, on the first XPHASE call):
    # the first call's time minus a steady call's, row statistics built before
    for name, A in dA.items():
        A.spmv(xs[name], ys[name], variant=spira_amd.VARIANTS["vec32"], device=True, sync=True)
        t = time.perf_counter()
        A.spmv(xs[name], ys[name], variant=spira_amd.VARIANTS["xphase_paced"], device=True, sync=True)
        first = time.perf_counter() - t
        t = time.perf_counter()
        for _ in range(5):
            A.spmv(xs[name], ys[name], variant=spira_amd.VARIANTS["xphase_paced"], device=True, sync=True)
        steady = (time.perf_counter() - t) / 5
        info[name]["plan_build_ms"] = round(1e3 * (first - steady), 3)
        log(f"bench: {name}: XPHASE plan built in {info[name]['plan_build_ms']} ms")

    # variant choice: exhaustive autotune (default), heuristic, or fixed
    for name, A in dA.items():
        info[name]["auto_variant"] = A.choose_variant()  # AUTO (learned selector), before any pin
        if args.variant == "autotune":
            best, ms = A.autotune(xs[name], ys[name], reps=5)
            info[name]["autotune_ms"] = {spira_amd.VARIANT_NAMES[v]: round(t, 4) for v, t in ms.items()}
        elif args.variant == "auto":
            best = A.choose_variant()
        else:
            best = spira_amd.VARIANTS[args.variant]
            A.set_variant(best)
        info[name]["variant"] = spira_amd.VARIANT_NAMES[best]
        log(f"bench: {name}: variant {info[name]['variant']}")

    names = list(dA)

    def step(ev=None):
        for name in names:
            if ev is not None and name == dominant:
                ev[0].record(stream)
            dA[name].spmv(xs[name], ys[name], device=True, sync=False)
            if ev is not None and name == dominant:
                ev[1].record(stream)

    # dominant kernel = the one with more bytes per step (the random matrix)
    dominant = max(names, key=lambda k: info[k]["bytes"])

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    torch.cuda.synchronize()

    kev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for i in range(args.steps):
        step(kev[i])
    stop.record(stream)
    stop.synchronize()
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    ms_step = total_ms / args.steps
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in kev]))

    bytes_step = sum(info[k]["bytes"] for k in names)
    flops_step = sum(2 * info[k]["nnz"] for k in names)
    value = bytes_step / (ms_step * 1e-3) / 1e9

    # per-matrix times (separate short timing, same stream)
    detail = {}
    for name in names:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 50
        e0.record(stream)
        for _ in range(reps):
            dA[name].spmv(xs[name], ys[name], device=True, sync=False)
        e1.record(stream)
        e1.synchronize()
        t = e0.elapsed_time(e1) / reps
        gbs = info[name]["bytes"] / (t * 1e-3) / 1e9
        detail[name] = {"variant": info[name]["variant"], "n_rows": info[name]["n_rows"], "nnz": info[name]["nnz"],
                        "bytes_per_call": info[name]["bytes"], "ms": round(t, 5), "GB_s": round(gbs, 1),
                        "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
                        "GFLOP_s": round(2 * info[name]["nnz"] / (t * 1e-3) / 1e9, 1)}
        if "autotune_ms" in info[name]:
            detail[name]["autotune_ms"] = info[name]["autotune_ms"]
        detail[name]["plan_build_ms"] = info[name]["plan_build_ms"]
        av = info[name]["auto_variant"]
        detail[name]["auto_variant"] = spira_amd.VARIANT_NAMES[av]
        if "autotune_ms" in info[name]:
            at = info[name]["autotune_ms"]
            detail[name]["auto_fraction_of_autotuned"] = round(
                min(t for t in at.values() if t >= 0) / at[spira_amd.VARIANT_NAMES[av]], 4)

    # The reference's float instantiation (serial/spmv.hpp:90-135) on the same
    # matrices: spira_spmv_f32 with AUTO's choice, A's values and x rounded to
    # FP32, FP32 y (row sums in FP64), against the FP64 result above.
    for name in names:
        A, n = dA[name], info[name]["n_rows"]
        with torch.cuda.stream(stream):
            x32 = xs[name].float()
            y32 = torch.empty(n, dtype=torch.float32, device="cuda")
        stream.synchronize()
        A.spmv_f32(x32, y32, device=True, sync=True)  # builds the FP32 copy of the values
        t32 = _time_calls(torch, stream, lambda: A.spmv_f32(x32, y32, device=True, sync=False), 50)
        A.spmv(xs[name], ys[name], device=True, sync=True)
        b32 = W.spmv_bytes(n, n, info[name]["nnz"], s_val=4, s_col=4, s_ptr=4, s_x=4, s_y=4)
        detail[name]["float_instantiation"] = {
            "path": "spira_spmv_f32 (AUTO): FP32 values, x and y, FP64 row sums",
            "ms": round(t32, 5), "GB_s": round(b32 / (t32 * 1e-3) / 1e9, 1), "bytes_per_call": b32,
            "speedup_vs_fp64": round(detail[name]["ms"] / t32, 3),
            "normwise_rel_err_vs_fp64": _nrm_err(torch, y32, ys[name])}
        log(f"bench: {name}: float instantiation {detail[name]['float_instantiation']}")
        del x32, y32

    lock_detail, lock_src = (None, None) if args.no_lock else run_lock(args, np, torch, spira_amd, W, dev, stream)
    tr_detail, tr_src = (None, None) if args.no_transpose else run_transpose(
        args, np, torch, spira_amd, W, dev, stream, None if lock_src is None else lock_src[:3])
    if tr_detail is not None:
        log(f"bench: transpose {tr_detail['ms_serial']} ms ({tr_detail['GB_s']} GB/s)")
    alg_detail, alg_src = (None, None) if args.no_algebra else run_algebra(
        args, np, torch, spira_amd, W, dev, stream, None if lock_src is None else lock_src[:3])
    if alg_detail is not None:
        log(f"bench: add {alg_detail['add']['ms']} ms, spgemm {alg_detail['spgemm_random_1m']['ms']} ms")
    mixed_detail, mixed_src = (None, None) if args.no_mixed else run_mixed(args, np, torch, spira_amd, W, dev, stream)
    batch_detail = None if args.no_batch else run_batch(args, np, torch, spira_amd, W, dev, stream)
    auto_detail = None if args.no_auto else run_auto_selector(args, np, torch, spira_amd, dev, stream)

    achieved = info[dominant]["bytes"] / (kern_ms * 1e-3) / 1e9
    traffic = profile_traffic(dominant + ":" + info[dominant]["variant"])
    two = info[dominant]["variant"].startswith("xs")
    roofline = {"bound": "hbm", "kernel": f"spmv {info[dominant]['variant']} on {dominant}"
                + (" (slice kernel + partial-sum kernel, timed together)" if two else ""),
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_launch": info[dominant]["bytes"], "avg_launch_ms": round(kern_ms, 5)}
    if info[dominant]["variant"].startswith("xphase"):
        roofline["ceiling"] = xphase_ceiling(np, torch, dev, mats[dominant], kern_ms)
        log(f"bench: {dominant}: {roofline['ceiling']}")

    # ---- CPU baseline leg (oracle: threaded parallel_matrix restatement) ----
    cpu = None
    if not args.no_cpu:
        import oracle
        # parity of this run's y against the oracle (checker use)
        for name in names:
            off, col, val = mats[name]
            y_ref = oracle.spmv_f64(off, col, val, info[name]["x_host"])
            dA[name].spmv(xs[name], ys[name], device=True, sync=True)
            y = ys[name].cpu().numpy()
            detail[name]["normwise_rel_err_vs_fp64_oracle"] = oracle.normwise_rel_err(y, y_ref)
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        phys = _physical_cpus()
        threads = max(1, min(16, len(phys)))
        llc = _llc_bytes()
        log(f"bench: CPU baseline with {threads} pinned threads (one per physical core), LLC {llc >> 20} MiB")
        cpu_bytes = sum(W.spmv_bytes(info[k]["n_rows"], info[k]["n_rows"], info[k]["nnz"], s_ptr=8) for k in names)
        ybuf = {name: np.empty(info[name]["n_rows"]) for name in names}

        def step_times(pools, budget, min_steps, max_steps, cold):
            """Per-step seconds (one SpMV on each matrix) until budget s and min_steps."""
            for name in names:  # warm-up step
                pools[name].spmv(info[name]["x_host"], ybuf[name])
            ts, spent = [], 0.0
            while (spent < budget or len(ts) < min_steps) and len(ts) < max_steps:
                t = 0.0
                for name in names:
                    if cold:  # every worker evicts its share of the LLC, untimed
                        pools[name].flush_llc(llc)
                    t0 = time.perf_counter()
                    pools[name].spmv(info[name]["x_host"], ybuf[name])
                    t += time.perf_counter() - t0
                ts.append(t)
                spent += t
            return ts

        def make_pools(nt, simd):
            pools, kern = {}, None
            for name in names:
                off, col, val = mats[name]
                n = len(off) - 1
                p = oracle.Pool(nt, n, n)
                p.pin(phys[:nt])
                p.load_csr(off, col, val)  # each pinned worker first-touches its partition
                kern = p.set_simd(simd)
                pools[name] = p
            return pools, kern

        def rate(ts):
            return round(cpu_bytes / float(np.median(ts)) / 1e9, 2)

        pools, kern = make_pools(threads, True)
        warm = step_times(pools, args.cpu_seconds * 0.4, 5, 400, cold=False)
        cold = step_times(pools, args.cpu_seconds * 0.2, 5, 100, cold=True)
        for p in pools.values():
            p.close()
        pools, _ = make_pools(threads, False)
        loop = step_times(pools, args.cpu_seconds * 0.1, 5, 200, cold=False)
        others = {}
        for other in ("avx2", "avx512"):  # every SIMD dot this CPU has, beside the dispatched one
            if other == kern:
                continue
            try:
                for p in pools.values():
                    p.force_kernel(other)
            except RuntimeError:
                continue
            others[other] = step_times(pools, args.cpu_seconds * 0.05, 5, 100, cold=False)
        for p in pools.values():
            p.close()
        cpu = {"value": rate(warm), "unit": "GB/s", "cores": threads, "kind": "port",
               "sample": f"median of {len(warm)} warm steps (one SpMV on each matrix, full size); oracle/ "
                         f"threaded parallel_matrix restatement: uniform row split, one partition CSR per "
                         f"pinned worker (one per physical core), u64 offsets, per-row {kern} gather dot "
                         f"(the dot dispatch.cpp selects on this CPU), -O3 -march=x86-64-v3",
               "ms_per_step": round(1e3 * float(np.median(warm)), 3),
               "dot": kern,
               "cold_llc": {"value": rate(cold), "unit": "GB/s", "steps": len(cold),
                            "ms_per_step": round(1e3 * float(np.median(cold)), 3),
                            "flush": f"each worker reads {llc >> 20} MiB before each SpMV (untimed)"},
               "scalar_loop": {"value": rate(loop), "unit": "GB/s", "steps": len(loop),
                               "ms_per_step": round(1e3 * float(np.median(loop)), 3)},
               "other_dots": {k: {"value": rate(v), "unit": "GB/s", "steps": len(v)} for k, v in others.items()},
               "host": {"cpu_model": _cpu_model(), "nproc": os.cpu_count(), "affinity": cores,
                        "physical_cores": len(phys), "pinned_cpus": phys[:threads], "llc_bytes": llc}}
        # one thread, same matrices and dot (a bounded sample of steps)
        pools, _ = make_pools(1, True)
        one = step_times(pools, args.cpu_seconds * 0.25, 3, 50, cold=False)
        for p in pools.values():
            p.close()
        cpu["one_thread"] = {"value": rate(one), "unit": "GB/s",
                             "ms_per_step": round(1e3 * float(np.median(one)), 3), "steps": len(one)}
        if lock_detail is not None:
            o3, c3, v3, sp3, sc3, sv3, n3, lbytes = lock_src
            tl = []
            for rep in range(5):
                p = oracle.Pool(threads, n3, n3)
                p.pin(phys[:threads])
                p.load_csr(o3, c3, v3)  # install the locked matrix as the partitions' CSR
                t = time.perf_counter()
                p.lock(sp3, sc3, sv3)
                tl.append(time.perf_counter() - t)
                if rep == 0:
                    got = p.export_csr()
                    exp = oracle.lock(sp3, sc3, sv3, old=(o3, c3, v3))
                    lock_detail["bit_exact_vs_oracle"] = bool(
                        all(np.array_equal(a.view(np.uint8), b.view(np.uint8)) for a, b in zip(got, exp)))
                p.close()
            lock_detail["cpu_baseline_ms"] = round(1e3 * float(np.median(tl)), 2)
            lock_detail["cpu_baseline_threads"] = threads
            lock_detail["cpu_baseline_reps"] = len(tl)
        if tr_detail is not None:
            o, c, v, got = tr_src
            tt = []
            for rep in range(3):
                t = time.perf_counter()
                exp = oracle.transpose(o, c, v, len(o) - 1)
                tt.append(time.perf_counter() - t)
            tr_detail["bit_exact_vs_oracle"] = bool(
                all(np.array_equal(a.view(np.uint8), b.view(np.uint8)) for a, b in zip(got, exp)))
            tr_detail["cpu_baseline_ms"] = round(1e3 * float(np.median(tt)), 2)
            tr_detail["cpu_baseline_threads"] = 1
        if alg_detail is not None:
            base, other, lap = alg_src
            t = time.perf_counter()
            oracle.add(base, other)
            alg_detail["add"]["cpu_baseline_ms"] = round(1e3 * (time.perf_counter() - t), 2)
            # SpGEMM on the CPU: a bounded sample of rows (the first 65,536 rows
            # of A times all of A), scaled per product
            rows = 65536
            for key, src in (("spgemm_random_1m", base), ("spgemm_laplacian_2048", lap)):
                off = src[0][:rows + 1]
                sub = (off, src[1][:int(off[-1])], src[2][:int(off[-1])])
                nprod = oracle.spgemm_products(sub, src)
                t = time.perf_counter()
                oracle.spgemm(sub, src, len(src[0]) - 1)
                dt = time.perf_counter() - t
                alg_detail[key]["cpu_baseline_GFLOP_s"] = round(2 * nprod / dt / 1e9, 3)
                alg_detail[key]["cpu_baseline_sample"] = f"first {rows} rows of A times A ({nprod} products)"
            alg_detail["cpu_baseline_threads"] = 1
        if mixed_detail is not None:
            # errors of every storage against the FP64 oracle and its
            # double-double reference (checker use)
            A, x, y64, y, (ho, hc, hv, hx) = mixed_src
            y_or = oracle.spmv_f64(ho, hc, hv, hx)
            dd_hi, dd_lo = oracle.spmv_dd(ho, hc, hv, hx)
            A.spmv(x, y64, device=True, sync=True)
            g64 = y64.cpu().numpy()
            mixed_detail["fp64"]["normwise_rel_err_vs_fp64_oracle"] = oracle.normwise_rel_err(g64, y_or)
            mixed_detail["fp64"]["normwise_rel_err_vs_double_double"] = oracle.normwise_rel_err(g64, dd_hi, dd_lo)
            for st, rec in mixed_detail["storage"].items():
                for key, ratio in (("recompute_off", 0.0), (f"recompute_ratio_{args.recompute_ratio:g}",
                                                             args.recompute_ratio)):
                    v = spira_amd.VARIANTS[rec[key]["variant"]]
                    A.spmv_mixed(x, y, storage=st, variant=v, recompute_ratio=ratio, device=True, sync=True)
                    got = y.cpu().numpy()
                    rec[key]["normwise_rel_err_vs_fp64_oracle"] = oracle.normwise_rel_err(got, y_or)
                    rec[key]["normwise_rel_err_vs_double_double"] = oracle.normwise_rel_err(got, dd_hi, dd_lo)
            log("bench: mixed-precision errors checked against the FP64 oracle")
    if mixed_src is not None:
        mixed_src[0].close()

    out = {
        "metric": "SpMV GB/s & % HBM peak; batched mixed-prec GFLOP/s vs plain FP64 at set error",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain FP64 CSR SpMV at HBM scale: 2048^2 5-point Laplacian + 4,194,304-row random "
                               "matrix with 8-32 nnz/row; one step = one SpMV on each",
                   "pct_hbm_peak": round(100 * value / HBM_PEAK_GBS, 2),
                   "GFLOP_s": round(flops_step / (ms_step * 1e-3) / 1e9, 1),
                   "bytes_per_step": bytes_step},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": detail,
        "lock": lock_detail,
        "transpose": tr_detail,
        "algebra": alg_detail,
        "mixed": mixed_detail,
        "batched": batch_detail,
        "auto_selector": auto_detail,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
