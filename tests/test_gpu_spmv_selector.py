"""The learned kernel selector of single-matrix FP64 SpMV (AUTO) on the
device: the library's AUTO choice (features + the compiled-in network,
spira_spmv_f64 with SPIRA_SPMV_AUTO and no autotune call) is the trained
network's argmin; on fresh corpus seeds (disjoint from the training and
held-out seeds) it reaches >= 90% of the exhaustive autotuner's best
throughput and gains over the row-length rule used without a trained model;
on the HBM-scale Laplacian and random matrices AUTO runs within 90% of the
autotuned time; AUTO's results match the oracle."""
import numpy as np
import pytest
import torch

import oracle
import spira_amd
from spira_amd import spmv_selector as S
from spira_amd import workloads as W

pytestmark = pytest.mark.gpu
FRESH = range(950_000, 950_060)


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


def test_auto_is_the_networks_choice_and_beats_the_rule(dev):
    sel = S.SpmvSelector.load(dev)
    auto, F, T, rule = [], [], [], []
    for seed in FRESH:
        off, col, val, n_cols, fam = W.spmv_corpus_matrix(seed, "cuda", 1 << 22)
        n = off.numel() - 1
        A = spira_amd.DeviceCSR(dev, n, n_cols)
        A.upload(off, col, val, device=True)
        rule.append(S.rule_variant(*S.row_stats(off)))
        del off, col, val
        auto.append(A.choose_variant())  # AUTO resolved in the library
        F.append(spira_amd.matrix_features(A))
        x = torch.rand(n_cols, dtype=torch.float64, device="cuda")
        y = torch.empty(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _, ms = A.autotune(x, y, reps=3)
        T.append([ms[v] if ms[v] >= 0 else np.nan for v in range(spira_amd.N_VARIANTS)])
        A.close()
    F, T = np.stack(F), np.array(T)
    pred = sel.predict(F, np.isfinite(T))
    auto = np.array(auto)
    assert np.array_equal(auto, pred), (auto, pred)
    f_auto = S.fraction_of_best(T, auto)
    f_rule = S.fraction_of_best(T, np.array(rule))
    print(f"AUTO {f_auto.mean():.3f} (total time {np.nanmin(T, 1).sum() / T[np.arange(len(auto)), auto].sum():.3f}), "
          f"rule {f_rule.mean():.3f} of the exhaustive best")
    assert f_auto.mean() >= 0.9
    assert f_auto.mean() >= f_rule.mean() + 0.15


@pytest.mark.parametrize("name", ["laplacian_2048", "random_4M_8to32"])
def test_auto_on_the_hbm_scale_matrices(dev, name):
    """AUTO (no autotune call) on the two headline matrices: within 90% of the
    autotuned time, and its y within 1e-12 of the oracle."""
    if name == "laplacian_2048":
        off, col, val = W.laplacian_2d(2048)
    else:
        off, col, val = W.random_uniform_rows(1 << 22, 1 << 22, 8, 32, seed=43)
    n = len(off) - 1
    A = spira_amd.DeviceCSR(dev, n, n)
    A.upload(off, col, val)
    xh = W.dense_x(n)
    x = torch.from_numpy(xh).cuda()
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    A.spmv(x, y, device=True)  # AUTO
    v = A.choose_variant()
    y_auto = y.cpu().numpy()
    assert oracle.normwise_rel_err(y_auto, oracle.spmv_f64(off, col, val, xh)) <= 1e-12
    best, ms = A.autotune(x, y, reps=5)
    print(f"{name}: AUTO {spira_amd.VARIANT_NAMES[v]} {ms[v]:.4f} ms, best {spira_amd.VARIANT_NAMES[best]} "
          f"{ms[best]:.4f} ms")
    assert ms[best] / ms[v] >= 0.9


def test_auto_builds_a_copying_plan_on_the_second_call(dev):
    """AUTO whose choice copies A into its own layout (here XPHASE on a
    1,048,576-row random matrix) runs a plan-free variant on the first call
    after an upload and builds the copy on the second; both calls match the
    oracle.  Seen through the allocation-free entry point, which refuses a
    variant without a plan: after the first call the choice has none and
    AUTO there reruns the first call's variant bit for bit; after the second
    the choice runs.  A re-upload starts over."""
    n = 1 << 20
    off, col, val = W.random_uniform_rows(n, n, 8, 32, seed=77)
    A = spira_amd.DeviceCSR(dev, n, n)
    xh = W.dense_x(n)
    x = torch.from_numpy(xh).cuda()
    y_ref = oracle.spmv_f64(off, col, val, xh)
    for cycle in range(2):
        A.upload(off, col, val)
        y1 = torch.empty(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        A.spmv(x, y1, device=True)  # AUTO, first call
        full = A.choose_variant()
        assert full >= spira_amd.VARIANTS["xslice1"], spira_amd.VARIANT_NAMES[full]
        with pytest.raises(spira_amd.LogicError, match="no SpMV plan"):
            A.spmv_noalloc(x, torch.empty_like(y1), variant=full)
        y1n = torch.empty_like(y1)
        A.spmv_noalloc(x, y1n)  # AUTO: the first call's plan-free variant
        assert np.array_equal(y1.cpu().numpy().view(np.uint64), y1n.cpu().numpy().view(np.uint64))
        assert oracle.normwise_rel_err(y1.cpu().numpy(), y_ref) <= 1e-12
        y2 = torch.empty_like(y1)
        A.spmv(x, y2, device=True)  # AUTO, second call: builds the choice's plan
        assert A.choose_variant() == full
        assert oracle.normwise_rel_err(y2.cpu().numpy(), y_ref) <= 1e-12
        y3 = torch.empty_like(y1)
        A.spmv_noalloc(x, y3, variant=full)  # the plan exists now
        A.spmv_noalloc(x, y3)
        assert oracle.normwise_rel_err(y3.cpu().numpy(), y_ref) <= 1e-12
    A.close()
