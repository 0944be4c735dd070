"""Device open-mode staging (spira_stage + spira_lock_staged_f64) against the
oracle, bit-exact: inserts appended in several chunks from host and device
arrays (so a row's inserts arrive in different appends and must keep their
insertion order for last-write-wins), first lock and re-locks with dirty
masks, +-0.0 deletes, empty stages, the reference's lifecycle-style cases,
and the error behaviour (a bad index rejects the whole append)."""
import numpy as np
import pytest
import torch

import oracle
import spira_amd
from oracle_model import csr_equal, random_stream, stage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


def _append_chunks(rng, st, rows, cols, vals):
    """Append in 1..4 chunks, alternating host and device arrays."""
    cuts = np.sort(rng.choice(np.arange(1, max(rows.size, 2)), size=min(3, max(rows.size - 1, 0)), replace=False)) \
        if rows.size > 1 else []
    bounds = [0, *[int(c) for c in cuts], rows.size]
    for i, (a, b) in enumerate(zip(bounds[:-1], bounds[1:])):
        if b <= a:
            continue
        if i % 2:
            st.append(torch.from_numpy(rows[a:b].astype(np.int32)).cuda(),
                      torch.from_numpy(cols[a:b].astype(np.int32)).cuda(),
                      torch.from_numpy(vals[a:b]).cuda(), device=True)
        else:
            st.append(rows[a:b], cols[a:b], vals[a:b])


@pytest.mark.parametrize("seed", range(6))
def test_staged_lock_multicycle_bit_exact(dev, seed):
    rng = np.random.default_rng(300 + seed)
    n_rows = int(rng.integers(1, 4000))
    n_cols = int(rng.integers(1, 3000))
    A = spira_amd.DeviceCSR(dev, n_rows, n_cols)
    st = spira_amd.Stage(dev, n_rows, n_cols)
    old = None
    for cycle in range(4):
        n_ins = int(rng.integers(0, 15 * n_rows))
        rows, cols, vals = random_stream(rng, n_rows, n_cols, n_ins, zero_frac=0.15,
                                         col_pool=int(rng.integers(1, n_cols + 1)))
        _append_chunks(rng, st, rows, cols, vals)
        assert len(st) == n_ins
        ptr, c, v = stage(n_rows, rows, cols, vals)
        dirty = None
        if old is not None and cycle == 2:
            dirty = (rng.random(n_rows) < 0.3).astype(np.uint8)
            dirty[np.bincount(rows, minlength=n_rows) > 0] = 1
        nnz = A.lock_staged(st, dirty=dirty)
        exp = oracle.lock(ptr, c, v, old=old, dirty=dirty)
        assert nnz == exp[1].size
        assert csr_equal(A.download(), exp), f"seed {seed} cycle {cycle}"
        assert len(st) == 0
        old = exp


def test_staged_lock_keeps_insertion_order_within_rows(dev):
    """Overwrites of one (row, col) spread across appends: the last write wins;
    a zero written last deletes."""
    A = spira_amd.DeviceCSR(dev, 3, 5)
    st = spira_amd.Stage(dev, 3, 5)
    st.append([1, 0, 1], [2, 4, 2], [1.0, 7.0, 2.0])
    st.append(torch.tensor([1, 2], dtype=torch.int32).cuda(), torch.tensor([2, 0], dtype=torch.int32).cuda(),
              torch.tensor([3.0, 5.0], dtype=torch.float64).cuda(), device=True)
    st.append([0], [4], [-0.0])
    A.lock_staged(st)
    off, col, val = A.download()
    assert off.tolist() == [0, 0, 1, 2]
    assert col.tolist() == [2, 0] and val.tolist() == [3.0, 5.0]


def test_staged_lock_empty_and_errors(dev):
    A = spira_amd.DeviceCSR(dev, 4, 4)
    st = spira_amd.Stage(dev, 4, 4)
    assert A.lock_staged(st) == 0  # first lock with nothing staged: empty CSR
    assert A.download()[0].tolist() == [0, 0, 0, 0, 0]
    st.append([0, 1], [1, 2], [1.0, 2.0])
    with pytest.raises(spira_amd.OutOfRange):
        st.append(torch.tensor([0, 4], dtype=torch.int32).cuda(), torch.tensor([0, 0], dtype=torch.int32).cuda(),
                  torch.tensor([1.0, 1.0], dtype=torch.float64).cuda(), device=True)  # row 4 of 4
    with pytest.raises(spira_amd.OutOfRange):
        st.append(torch.tensor([0], dtype=torch.int32).cuda(), torch.tensor([9], dtype=torch.int32).cuda(),
                  torch.tensor([1.0], dtype=torch.float64).cuda(), device=True)
    assert len(st) == 2  # rejected appends left nothing behind
    assert A.lock_staged(st) == 2
    other = spira_amd.Stage(dev, 5, 4)
    with pytest.raises(spira_amd.InvalidArgument):
        A.lock_staged(other)


def test_staged_lock_large_relock_matches_host_staging(dev):
    """The re-lock workload (1M rows, 16 nnz/row, 2 inserts/row, 25%
    overwrite): staging on the device equals staging on the host."""
    from spira_amd import workloads as W
    n = 1 << 20
    (o3, c3, v3), (r3, ic3, iv3) = W.relock_workload(n)
    A = spira_amd.DeviceCSR(dev, n, n)
    B = spira_amd.DeviceCSR(dev, n, n)
    A.upload(o3, c3, v3)
    B.upload(o3, c3, v3)
    st = spira_amd.Stage(dev, n, n)
    st.append(torch.from_numpy(r3.astype(np.int32)).cuda(), torch.from_numpy(ic3.astype(np.int32)).cuda(),
              torch.from_numpy(iv3).cuda(), device=True)
    nA = A.lock_staged(st)
    nB = B.lock(*spira_amd.stage_from_coo(n, r3, ic3, iv3))
    assert nA == nB
    assert csr_equal(A.download(), B.download())
