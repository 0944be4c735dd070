"""Learned kernel selector for single-matrix FP64 SpMV (the AUTO variant).

Replaces the reference's one-time kernel pick (dispatch.cpp:34-67, a CPUID
test choosing the per-row SIMD dot for every matrix) with a per-matrix choice
among the 15 FP64 kernel variants of spira_spmv_f64 (lanes per row of the VEC
kernels, the LDS tile of the STREAM kernels, the XCD column-slice and XPHASE
layouts).  The exhaustive autotuner (spira_spmv_autotune_f64: every variant
timed on the matrix) is the judge and the fallback; the selector is a small
MLP (the project's HIP forward/backward/Adam kernels, as for the batched
schedules) mapping the device-computed matrix features (matrix_features) to
the predicted log2 slowdown of every variant relative to the best, whose
argmin among the variants the matrix fits is AUTO's choice.  The trained
parameters are compiled into libspira_amd (csrc/spmv_selector_weights.inc),
so the C-ABI's AUTO path (spira_spmv_f64 with SPIRA_SPMV_AUTO, no autotune
call) runs the same network on the device once per locked matrix.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import MLP, N_FEATURES, N_VARIANTS, VARIANT_NAMES, Device, DeviceCSR, matrix_features

HIDDEN = 64
N_OUT = N_VARIANTS
_HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(_HERE, "spmv_selector_model.json")
INC_PATH = os.path.join(_HERE, "csrc", "spmv_selector_weights.inc")


def rule_variant(mean: float, max_len: float, cv: float) -> int:
    """The library's choice without a trained model (spmv_rule_variant, row
    lengths only): irregular or short rows -> STREAM, else VEC with lanes
    close to the mean row length."""
    if max_len > 4.0 * max(mean, 8.0) or cv > 0.5:
        return 8  # STREAM2K
    if mean <= 6.0:
        return 7  # STREAM1K
    g = 2
    while g < 64 and g < mean:
        g <<= 1
    return {2: 1, 4: 2, 8: 3, 16: 4, 32: 5}.get(g, 6)


def row_stats(off) -> tuple:
    """(mean, max, coefficient of variation) of the row lengths."""
    import torch
    lens = (off[1:] - off[:-1]).double()
    n = max(lens.numel(), 1)
    mean = float(lens.sum()) / n
    var = float((lens * lens).sum()) / n - mean * mean
    return mean, float(lens.max()) if lens.numel() else 0.0, (max(var, 0.0) ** 0.5 / mean) if mean > 0 else 0.0


def collect(dev: Device, seeds, reps: int = 3, log=None, max_rows: int = 1 << 23):
    """Features and exhaustive autotune times (ms per call; NaN for a variant
    whose size limits the matrix exceeds) of the corpus matrices of `seeds`.
    Returns (F [n x 18], T [n x 15], meta list; meta carries the row-length
    rule's choice)."""
    import torch
    from . import workloads as W
    F, T, meta = [], [], []
    for i, seed in enumerate(seeds):
        off, col, val, n_cols, fam = W.spmv_corpus_matrix(int(seed), "cuda", max_rows)
        n = off.numel() - 1
        A = DeviceCSR(dev, n, n_cols)
        A.upload(off, col, val, device=True)
        nnz = int(col.numel())
        stats = row_stats(off)
        del off, col, val
        x = torch.rand(n_cols, dtype=torch.float64, device="cuda")
        y = torch.empty(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        F.append(matrix_features(A))
        _, ms = A.autotune(x, y, reps=reps)
        T.append([ms[v] if ms[v] >= 0 else np.nan for v in range(N_OUT)])
        meta.append({"seed": int(seed), "family": fam, "n_rows": n, "n_cols": int(n_cols), "nnz": nnz,
                     "rule": rule_variant(*stats)})
        A.close()
        del x, y
        if log and (i + 1) % 50 == 0:
            log(f"spmv selector: {i + 1}/{len(seeds)} matrices timed")
    return np.stack(F).astype(np.float32), np.array(T, np.float64), meta


def targets_from_times(T: np.ndarray):
    """log2(t / t_best) per variant, and the mask of the variants that ran."""
    best = np.nanmin(T, axis=1, keepdims=True)
    M = np.isfinite(T)
    Y = np.where(M, np.log2(np.where(M, T, 1.0) / best), 0.0)
    return Y.astype(np.float32), M.astype(np.float32)


def fraction_of_best(T: np.ndarray, choice) -> np.ndarray:
    """Throughput of the chosen variant as a fraction of the best one's."""
    choice = np.asarray(choice)
    return np.nanmin(T, axis=1) / T[np.arange(T.shape[0]), choice]


class SpmvSelector:
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
        if d.get("n_out") != N_OUT or d.get("variants") != [VARIANT_NAMES[v] for v in range(N_OUT)]:
            raise ValueError(f"{path}: model was trained for another variant list")
        s = cls(dev, np.array(d["params"], np.float32))
        s.meta = {k: v for k, v in d.items() if k != "params"}
        return s

    def scores(self, F: np.ndarray) -> np.ndarray:
        """Predicted log2 slowdown of every variant (network forward on the device)."""
        import torch
        X = torch.from_numpy(np.ascontiguousarray(F, np.float32)).cuda()
        Y = torch.empty((F.shape[0], N_OUT), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self.net.forward(X, Y)
        return Y.cpu().numpy()

    def predict(self, F: np.ndarray, ran: np.ndarray) -> np.ndarray:
        """Variant per row of F among those marked runnable in `ran`."""
        P = self.scores(F)
        P[~ran] = np.inf
        return np.argmin(P, axis=1)

    def fit(self, F, T, epochs: int = 4000, lr: float = 3e-3, log=None):
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
            if ep % 250 == 0 or ep == epochs - 1:
                hist.append((ep, loss))
                if log:
                    log(f"spmv selector: epoch {ep} loss {loss:.4f}")
        return hist

    def save(self, path: str = MODEL_PATH, **meta):
        d = dict(meta)
        d.update({"n_in": N_FEATURES, "hidden": HIDDEN, "n_out": N_OUT,
                  "variants": [VARIANT_NAMES[v] for v in range(N_OUT)],
                  "params": [float(v) for v in self.net.get_params()]})
        with open(path, "w") as f:
            json.dump(d, f)

    def export_inc(self, path: str = INC_PATH):
        write_inc(self.net.get_params(), path)


def write_inc(params, path: str = INC_PATH, source: str = "spira_amd/spmv_selector_model.json"):
    """The parameters as the HIP include the library compiles in (layout
    W1[hidden][in], b1, W2, b2, W3[out][hidden], b3): a device array, read in
    place by the selector's forward kernel."""
    p = np.asarray(params, np.float32)
    lines = ["// Parameters of the single-matrix SpMV kernel selector, written by",
             f"// spira_amd/spmv_selector.py (write_inc) from {source}.",
             f"constexpr unsigned kSpmvSelIn = {N_FEATURES}, kSpmvSelHidden = {HIDDEN}, kSpmvSelOut = {N_OUT};",
             "constexpr int kSpmvSelTrained = 1;",
             f"constexpr unsigned kSpmvSelNParams = {p.size};",
             f"__device__ const float kSpmvSelParams[{p.size}] = {{"]
    for i in range(0, p.size, 6):
        lines.append("    " + ", ".join(f"{float(v):.9e}f" for v in p[i:i + 6]) + ",")
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":  # regenerate the include from the committed model
    with open(MODEL_PATH) as f:
        write_inc(json.load(f)["params"])
