"""Learned schedule selector for batched SpMV.

For every matrix of a batch and every k (1, 4, 16, 32) a schedule is one of
SCHEDULES: lanes per row (1 << elog2, at most 64 / k: this fixes the rows per
wavefront, 64 / (lanes * k)) and the work-tile size (nonzeros per tile of the
persistent kernel).  The exhaustive autotuner (`time_schedules`) times every
schedule on the device; the selector is a small MLP (hand-written HIP
forward/backward/Adam in libspira_amd) that maps the matrix features
(`matrix_features`, computed on the device) plus log2(k) to the predicted
log2 slowdown of every schedule relative to the best, and picks the argmin.
It is trained on autotuner timings over a seeded corpus (`train`) and scored
as the fraction of the best achievable throughput on held-out matrices.
The built-in heuristic (`heuristic_schedule`, the batch default) is the
method without it.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import Batch, Device, DeviceCSR, MLP, N_FEATURES, matrix_features

K_LIST = (1, 4, 16, 32)
ELOG2 = tuple(range(7))            # lanes per row 1..64
TILES = ("small", "default")        # work-tile size class
SCHEDULES = [(e, t) for e in ELOG2 for t in range(len(TILES))]
N_OUT = len(SCHEDULES)
HIDDEN = 64
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_model.json")


def max_elog2(k: int) -> int:
    e = 0
    while (1 << (e + 1)) * k <= 64:
        e += 1
    return e


def tile_nnz(k: int, t: int) -> int:
    """Nonzeros per work tile: 'default' matches the batch's built-in choice."""
    return max(64, 2048 // k) if t == 0 else max(256, 16384 // k)


def valid_mask(k: int) -> np.ndarray:
    return np.array([e <= max_elog2(k) for e, t in SCHEDULES], dtype=bool)


def heuristic_schedule(mean_len: float, k: int) -> int:
    """The batch's built-in default (lanes per row ~ half the mean row
    length, default tile), as an index into SCHEDULES."""
    e = 0
    while e < max_elog2(k) and (1 << e) * 2.0 < mean_len:
        e += 1
    return SCHEDULES.index((e, 1))


def features_for(A: DeviceCSR, k: int) -> np.ndarray:
    f = matrix_features(A).copy()
    f[N_FEATURES - 1] = math.log2(k) / 5.0
    return f


def time_schedules(dev: Device, A: DeviceCSR, k: int, X, Y, stream, reps: int = 3) -> np.ndarray:
    """Exhaustive autotune of one matrix at one k: mean ms per call for every
    schedule (NaN where the schedule is not valid for k)."""
    import torch
    B = Batch(dev, [A], "f32")
    out = np.full(N_OUT, np.nan)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for i, (e, t) in enumerate(SCHEDULES):
        if e > max_elog2(k):
            continue
        B.set_schedule(k, elog2=[e], tile_nnz=[tile_nnz(k, t)])
        B.spmm(X, Y, sync=False)  # warm
        e0.record(stream)
        for _ in range(reps):
            B.spmm(X, Y, sync=False)
        e1.record(stream)
        e1.synchronize()
        out[i] = e0.elapsed_time(e1) / reps
    B.close()
    return out


def targets_from_times(T: np.ndarray):
    """log2(t / t_best) per schedule, mask of valid entries."""
    best = np.nanmin(T, axis=1, keepdims=True)
    M = np.isfinite(T)
    Y = np.where(M, np.log2(np.where(M, T, 1.0) / best), 0.0)
    return Y.astype(np.float32), M.astype(np.float32)


def fraction_of_best(T: np.ndarray, choice: np.ndarray) -> np.ndarray:
    best = np.nanmin(T, axis=1)
    return best / T[np.arange(T.shape[0]), choice]


class Selector:
    """Trained network + the rule that turns its outputs into a schedule."""

    def __init__(self, dev: Device, params=None, seed: int = 0):
        self.dev = dev
        self.net = MLP(dev, N_FEATURES, HIDDEN, N_OUT, seed)
        self.meta = {}
        if params is not None:
            self.net.set_params(params)

    @classmethod
    def load(cls, dev: Device, path: str = MODEL_PATH):
        with open(path) as f:
            d = json.load(f)
        s = cls(dev, np.array(d["params"], np.float32))
        s.meta = {k: v for k, v in d.items() if k != "params"}
        return s

    def save(self, path: str, **meta):
        d = dict(meta)
        d.update({"n_in": N_FEATURES, "hidden": HIDDEN, "n_out": N_OUT, "schedules": SCHEDULES,
                  "params": [float(v) for v in self.net.get_params()]})
        with open(path, "w") as f:
            json.dump(d, f)

    def predict(self, F: np.ndarray, ks) -> np.ndarray:
        """Schedule index per row of F (features incl. the k slot) for k in ks."""
        import torch
        X = torch.from_numpy(np.ascontiguousarray(F, np.float32)).cuda()
        Y = torch.empty((F.shape[0], N_OUT), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self.net.forward(X, Y)
        P = Y.cpu().numpy()
        for i, k in enumerate(ks):
            P[i, ~valid_mask(k)] = np.inf
        return np.argmin(P, axis=1)

    def fit(self, F: np.ndarray, T: np.ndarray, epochs: int = 2000, lr: float = 3e-3, log=None):
        """Full-batch Adam on the masked squared error of log2 slowdowns."""
        import torch
        Y, M = targets_from_times(T)
        Xd = torch.from_numpy(np.ascontiguousarray(F, np.float32)).cuda()
        Yd = torch.from_numpy(Y).cuda()
        Md = torch.from_numpy(M).cuda()
        torch.cuda.synchronize()
        cnt = float(M.sum())
        hist = []
        for ep in range(epochs):
            loss, _ = self.net.loss_grad(Xd, Yd, Md, cnt)
            self.net.adam_step(lr=lr)
            if ep % 100 == 0 or ep == epochs - 1:
                hist.append((ep, loss))
                if log:
                    log(f"selector: epoch {ep} loss {loss:.4f}")
        return hist

    def schedule_batch(self, batch: Batch, mats, k: int):
        F = np.stack([features_for(A, k) for A in mats])
        idx = self.predict(F, [k] * len(mats))
        el = [SCHEDULES[i][0] for i in idx]
        tl = [tile_nnz(k, SCHEDULES[i][1]) for i in idx]
        batch.set_schedule(k, elog2=el, tile_nnz=tl)
        return idx


def heuristic_batch_schedule(batch: Batch, mats, k: int):
    batch.set_schedule(k)


def build_corpus(dev: Device, n: int, seed0: int, min_rows=256, max_rows=65536):
    """Seeded matrices of the batched workload, uploaded as locked DeviceCSRs."""
    import torch
    from . import workloads as W
    mats, means = [], []
    for i in range(n):
        off, col, val = W.batch_matrix(seed0 + i, min_rows=min_rows, max_rows=max_rows, device="cuda")
        nr = off.numel() - 1
        A = DeviceCSR(dev, nr, nr)
        A.upload(off, col, val, device=True)
        mats.append(A)
        means.append(float(col.numel()) / max(nr, 1))
        del off, col, val
    torch.cuda.synchronize()
    return mats, np.array(means)


def collect(dev: Device, mats, stream, reps: int = 3, log=None):
    """Features and exhaustive timings for every (matrix, k)."""
    import torch
    nmax = max(A.n_rows for A in mats)
    X = torch.randn((nmax * 32,), dtype=torch.float32, device="cuda")
    Y = torch.empty((nmax * 32,), dtype=torch.float32, device="cuda")
    F, T, KS, MI = [], [], [], []
    for i, A in enumerate(mats):
        for k in K_LIST:
            F.append(features_for(A, k))
            T.append(time_schedules(dev, A, k, X, Y, stream, reps))
            KS.append(k)
            MI.append(i)
        if log and i % 200 == 0:
            log(f"selector: timed {i}/{len(mats)} matrices")
    return np.stack(F), np.stack(T), np.array(KS), np.array(MI)
