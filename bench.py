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
    p.add_argument("--calib-copy-mb", type=int, default=0,
                   help="also run a plain device copy of this many MiB (PMC traffic calibration)")
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
    lbytes = W.lock_bytes(n3, int(c3.size), int(sc3.size), nnz_new)
    tdev = float(np.median(t_dev[1:]))
    detail = {"workload": "re-lock merge 1,048,576^2: 16 random nnz/row, then 2 inserts/row, 25% overwrite",
              "nnz_old": int(c3.size), "staged": int(sc3.size), "nnz_new": int(nnz_new),
              "bytes_per_lock": lbytes, "ms_device_inputs": round(1e3 * tdev, 3),
              "GB_s": round(lbytes / tdev / 1e9, 1)}
    if t_host:
        detail["ms_host_inputs_pcie_inclusive"] = round(1e3 * float(np.median(t_host[1:])), 3)
    L.close()
    return detail, (o3, c3, v3, sp3, sc3, sv3, n3, lbytes)


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

    # variant choice: exhaustive autotune (default), heuristic, or fixed
    for name, A in dA.items():
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

    lock_detail, lock_src = (None, None) if args.no_lock else run_lock(args, np, torch, spira_amd, W, dev, stream)

    achieved = info[dominant]["bytes"] / (kern_ms * 1e-3) / 1e9
    traffic = profile_traffic(dominant + ":" + info[dominant]["variant"])
    two = info[dominant]["variant"].startswith("xs")
    roofline = {"bound": "hbm", "kernel": f"spmv {info[dominant]['variant']} on {dominant}"
                + (" (slice kernel + partial-sum kernel, timed together)" if two else ""),
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_launch": info[dominant]["bytes"], "avg_launch_ms": round(kern_ms, 5)}

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
        threads = max(1, min(16, cores))
        log(f"bench: CPU baseline with {threads} threads")
        pools = {}
        for name in names:
            off, col, val = mats[name]
            n = len(off) - 1
            p = oracle.Pool(threads, n, n)
            p.load_csr(off, col, val)
            pools[name] = p
        ybuf = {name: np.empty(info[name]["n_rows"]) for name in names}
        for name in names:  # warm
            pools[name].spmv(info[name]["x_host"], ybuf[name])
        reps, t_cpu = 0, 0.0
        tstart = time.perf_counter()
        while True:
            for name in names:
                pools[name].spmv(info[name]["x_host"], ybuf[name])
            reps += 1
            t_cpu = time.perf_counter() - tstart
            if t_cpu >= args.cpu_seconds or reps >= 200:
                break
        cpu_bytes = sum(W.spmv_bytes(info[k]["n_rows"], info[k]["n_rows"], info[k]["nnz"], s_ptr=8) for k in names)
        cpu = {"value": round(cpu_bytes * reps / t_cpu / 1e9, 2), "unit": "GB/s", "cores": threads, "kind": "port",
               "sample": f"{reps} warm steps (both matrices, full size) in {t_cpu:.1f}s; oracle/ threaded "
                         f"parallel_matrix restatement, uniform row split, u64 offsets, scalar loop -O3 "
                         f"-march=x86-64-v3",
               "ms_per_step": round(1e3 * t_cpu / reps, 3)}
        for p in pools.values():
            p.close()
        if lock_detail is not None:
            o3, c3, v3, sp3, sc3, sv3, n3, lbytes = lock_src
            tl = []
            for rep in range(3):
                p = oracle.Pool(threads, n3, n3)
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
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
