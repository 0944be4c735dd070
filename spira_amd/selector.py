"""Learned schedule selector for batched SpMV.

For every matrix of a batch and every k (1, 4, 16, 32) a schedule is one of
SCHEDULES: plog2 (target nonzeros per lane, which sets the lanes per row of
every row-length class and so the rows per wavefront, 64 / (lanes * k)),
xlds (stage the matrix's X block in LDS: the LDS tile), the work-item size
class and stream (read the entries from global memory instead of copying
each item's entries into LDS beside the X block).  The exhaustive autotuner (`exhaustive`) runs every schedule on
the whole batch with device timestamps on every work item
(`Batch.profile`), which gives every matrix's cost inside its batch from
one launch per schedule.  The selector is a small MLP (hand-written HIP
forward/backward/Adam in libspira_amd) that maps the matrix features
(`matrix_features`, computed on the device) plus log2(k) to the predicted
log2 slowdown of every schedule relative to the best, and picks the argmin
among the schedules valid for the matrix.  It is trained on autotuner costs
over a seeded corpus (`train_selector`) and scored as the fraction of the
best achievable throughput on held-out matrices.  The built-in schedule
(`heuristic_schedule`, the batch default) is the method without it.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import (BATCH_MAX_PLOG2, BATCH_XLDS_FLOATS, BATCH_XLDS_TWO_PASS_COLS, Batch, Device, DeviceCSR, MLP,
               N_FEATURES, matrix_features)

K_LIST = (1, 4, 16, 32)
PLOG2 = tuple(range(BATCH_MAX_PLOG2 + 1))
ITEMS = ("default", "large")        # work-item size class
SCHEDULES = [(p, x, t, st) for p in PLOG2 for x in (0, 1) for t in range(len(ITEMS)) for st in (0, 1)]
N_OUT = len(SCHEDULES)
HIDDEN = 64
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_model.json")


def item_nnz(k: int, t: int) -> int:
    """Work-item size (nonzeros + rows): 'default' is the batch's built-in."""
    return max(4096, 131072 // k) if t == 0 else max(16384, 524288 // k)


def xlds_fits(n_cols, k: int):
    """The X block can be staged in LDS: in one load, or at k = 1 in two halves."""
    n_cols = np.asarray(n_cols)
    return (n_cols * k <= BATCH_XLDS_FLOATS) | ((k == 1) & (n_cols <= BATCH_XLDS_TWO_PASS_COLS))


def valid_mask(n_cols: int, k: int) -> np.ndarray:
    fits = bool(xlds_fits(n_cols, k))
    return np.array([(x == 0) or fits for p, x, t, st in SCHEDULES], dtype=bool)


def heuristic_schedule(n_cols: int, k: int) -> int:
    """The batch's built-in schedule (plog2 2, X block in LDS at k = 1 when
    it fits in one or two loads, default items, entries staged), as an index
    into SCHEDULES."""
    x = 1 if (k == 1 and bool(xlds_fits(n_cols, k))) else 0
    return SCHEDULES.index((2, x, 0, 0))


def features_for(A: DeviceCSR, k: int) -> np.ndarray:
    f = matrix_features(A).copy()
    f[N_FEATURES - 1] = math.log2(k) / 5.0
    return f


def features_by_k(mats, ks=K_LIST) -> dict:
    """{k: features of every matrix at k}: one device feature pass per matrix,
    the k slot filled per k."""
    base = np.stack([matrix_features(A) for A in mats])
    out = {}
    for k in ks:
        f = base.copy()
        f[:, N_FEATURES - 1] = math.log2(k) / 5.0
        out[k] = f
    return out


def apply(batch: Batch, k: int, choice, n_cols):
    """Install per-matrix schedule indices into the batch."""
    choice = np.asarray(choice)
    p = np.array([SCHEDULES[i][0] for i in choice])
    x = np.array([SCHEDULES[i][1] for i in choice]) * xlds_fits(n_cols, k)
    t = np.array([item_nnz(k, SCHEDULES[i][2]) for i in choice])
    st = np.array([SCHEDULES[i][3] for i in choice])
    batch.set_schedule(k, plog2=p, xlds=x.astype(np.int64), item_nnz=t, stream=st)


def exhaustive(batch: Batch, k: int, n_cols, X, Y, reps: int = 2) -> np.ndarray:
    """Exhaustive autotune of a whole batch at one k: every matrix's mean
    cost (CU-microseconds inside the batch) under every schedule; NaN where
    the schedule is not valid for the matrix."""
    n = len(batch.mats)
    fits = xlds_fits(n_cols, k)
    out = np.full((n, N_OUT), np.nan)
    for i, (p, x, t, st) in enumerate(SCHEDULES):
        if x and not fits.any():
            continue
        xl = (fits if x else np.zeros(n, bool)).astype(np.int64)
        batch.set_schedule(k, plog2=p, xlds=xl, item_nnz=item_nnz(k, t), stream=st)
        batch.spmm(X, Y, sync=True)  # warm
        c = np.mean([batch.profile(X, Y) for _ in range(reps)], axis=0)
        out[:, i] = np.where(fits | (x == 0), c, np.nan)
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
        if d.get("n_out") != N_OUT or [tuple(s) for s in d.get("schedules", [])] != SCHEDULES:
            raise ValueError(f"{path}: model was trained for another schedule space")
        s = cls(dev, np.array(d["params"], np.float32))
        s.meta = {k: v for k, v in d.items() if k != "params"}
        return s

    def save(self, path: str, **meta):
        d = dict(meta)
        d.update({"n_in": N_FEATURES, "hidden": HIDDEN, "n_out": N_OUT, "schedules": SCHEDULES,
                  "params": [float(v) for v in self.net.get_params()]})
        with open(path, "w") as f:
            json.dump(d, f)

    def predict(self, F: np.ndarray, ks, n_cols) -> np.ndarray:
        """Schedule index per row of F (features incl. the k slot), for the
        k and the column count of each row's matrix."""
        import torch
        X = torch.from_numpy(np.ascontiguousarray(F, np.float32)).cuda()
        Y = torch.empty((F.shape[0], N_OUT), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self.net.forward(X, Y)
        P = Y.cpu().numpy()
        for i, (k, nc) in enumerate(zip(ks, n_cols)):
            P[i, ~valid_mask(nc, k)] = np.inf
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

    def schedule_batch(self, batch: Batch, mats, k: int, feats=None):
        """Predict and install a schedule for every matrix of the batch."""
        F = np.stack([features_for(A, k) for A in mats]) if feats is None else feats
        n_cols = [A.n_cols for A in mats]
        idx = self.predict(F, [k] * len(mats), n_cols)
        apply(batch, k, idx, n_cols)
        return idx


def build_corpus(dev: Device, n: int, seed0: int, min_rows=256, max_rows=65536):
    """Seeded matrices of the batched workload, uploaded as locked DeviceCSRs."""
    import torch
    from . import workloads as W
    mats = []
    for i in range(n):
        off, col, val = W.batch_matrix(seed0 + i, min_rows=min_rows, max_rows=max_rows, device="cuda")
        nr = off.numel() - 1
        A = DeviceCSR(dev, nr, nr)
        A.upload(off, col, val, device=True)
        mats.append(A)
        del off, col, val
    torch.cuda.synchronize()
    return mats


def collect(dev: Device, mats, reps: int = 2, log=None, ks=K_LIST):
    """Features and exhaustive in-batch costs for every (matrix, k), the
    matrices forming one FP32 batch.  Returns F, T, K, MI, n_cols."""
    import torch
    B = Batch(dev, mats, "f32")
    X = torch.randn((B.n_cols * max(ks),), dtype=torch.float32, device="cuda")
    Y = torch.empty((B.n_rows * max(ks),), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    n_cols = np.array([A.n_cols for A in mats])
    feats = features_by_k(mats, ks)
    F, T, KS, MI = [], [], [], []
    for k in ks:
        F.append(feats[k])
        T.append(exhaustive(B, k, n_cols, X, Y, reps))
        KS.append(np.full(len(mats), k))
        MI.append(np.arange(len(mats)))
        if log:
            log(f"selector: k={k} exhaustive costs collected for {len(mats)} matrices")
    B.close()
    return np.concatenate(F), np.concatenate(T), np.concatenate(KS), np.concatenate(MI), n_cols
