"""Seeded synthetic matrices for the benchmark configurations.

Each generator returns numpy arrays in the reference's SoA CSR convention
(u64 offsets, u32 columns sorted and distinct within a row, f64 values) or an
open-mode insert stream (row, col, val in insertion order).  Seeds follow the
reference benchmark: base seed 42 (bench/spira_bench.cpp:40).
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 42


def _sorted_distinct_rows(rng, lens, n_cols, lo=None, hi=None):
    """Per row i, lens[i] distinct sorted columns, uniform in [lo_i, hi_i)
    (default [0, n_cols)).  Rows are generated as a padded 2-D block, sorted
    along rows, and rows with a repeated column are redrawn."""
    n = lens.size
    width = int(lens.max()) if n else 0
    out = np.empty((n, max(width, 1)), np.uint32)
    lo = np.zeros(n, np.int64) if lo is None else lo.astype(np.int64)
    hi = np.full(n, n_cols, np.int64) if hi is None else hi.astype(np.int64)
    todo = np.arange(n)
    pos = np.arange(max(width, 1))[None, :]
    span = (hi[todo] - lo[todo])[:, None]
    draw = lo[todo][:, None] + (rng.random((todo.size, max(width, 1))) * span).astype(np.int64)
    draw = np.minimum(draw, hi[todo][:, None] - 1).astype(np.uint32)
    valid = pos < lens[todo][:, None]
    draw[~valid] = np.iinfo(np.uint32).max
    draw.sort(axis=1)
    dup = ((draw[:, 1:] == draw[:, :-1]) & valid[:, 1:]).any(axis=1)
    out[:] = draw
    # rows that drew a repeated column: sample without replacement instead
    for i in np.nonzero(dup)[0]:
        L = int(lens[i])
        row = np.sort(rng.choice(int(hi[i] - lo[i]), size=L, replace=False) + lo[i]).astype(np.uint32)
        out[i, :L] = row
        out[i, L:] = np.iinfo(np.uint32).max
    return out


def _pack(lens, padded):
    """Flatten a padded (n, w) column block with row lengths lens into CSR."""
    n = lens.size
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    mask = np.arange(padded.shape[1])[None, :] < lens[:, None]
    return off, padded[mask]


def random_uniform_rows(n_rows, n_cols, lo, hi, seed=BASE_SEED, chunk=1 << 20):
    """n_rows rows, nnz/row uniform in [lo, hi], distinct uniform columns,
    values U(-1, 1).  Returns (offsets, cols, vals)."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, size=n_rows).astype(np.int64)
    off = np.zeros(n_rows + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    cols = np.empty(int(off[-1]), np.uint32)
    for a in range(0, n_rows, chunk):
        b = min(a + chunk, n_rows)
        blk = _sorted_distinct_rows(rng, lens[a:b], n_cols)
        _, c = _pack(lens[a:b], blk)
        cols[int(off[a]):int(off[b])] = c
    vals = rng.uniform(-1.0, 1.0, size=cols.size)
    return off, cols, vals


def powerlaw_rows(n_rows, s=1.8, clip=50000, local_frac=0.7, band=4096, seed=BASE_SEED ^ 0xDEADBEEF,
                  values="normal"):
    """Power-law rows: lengths Zipf(s) clipped to [1, clip] (numpy's zipf, so
    the clipped mean for s = 1.8 is about 26, not 12); each column is, with
    probability local_frac, uniform in [r - band, r + band] (clipped to the
    matrix) and otherwise uniform over all columns; columns distinct within a
    row; values N(0, 1) (or U(-1, 1) with values="uniform").  Square matrix."""
    rng = np.random.default_rng(seed)
    n = n_rows
    lens = np.minimum(rng.zipf(s, size=n), min(clip, n)).astype(np.int64)
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    cols = np.empty(int(off[-1]), np.uint32)
    rows_all = np.arange(n, dtype=np.int64)
    # rows are handled in width buckets so each bucket is one padded 2-D draw
    edges = [0, 8, 64, 512, 4096, clip + 1]
    for lo_w, hi_w in zip(edges[:-1], edges[1:]):
        sel = rows_all[(lens > lo_w) & (lens <= hi_w)]
        if sel.size == 0:
            continue
        width = int(lens[sel].max())
        for a in range(0, sel.size, max(1, (1 << 24) // width)):
            rs = sel[a:a + max(1, (1 << 24) // width)]
            L = lens[rs]
            loc = rng.random((rs.size, width)) < local_frac
            lo = np.maximum(rs - band, 0)[:, None]
            hi = np.minimum(rs + band + 1, n)[:, None]
            local = lo + (rng.random((rs.size, width)) * (hi - lo)).astype(np.int64)
            glob = (rng.random((rs.size, width)) * n).astype(np.int64)
            draw = np.where(loc, local, glob).astype(np.uint32)
            pos = np.arange(width)[None, :]
            valid = pos < L[:, None]
            draw[~valid] = np.iinfo(np.uint32).max
            draw.sort(axis=1)
            dup = ((draw[:, 1:] == draw[:, :-1]) & valid[:, 1:]).any(axis=1)
            for i in np.nonzero(dup)[0]:
                want = int(L[i])
                have = np.unique(draw[i, :want])
                while have.size < want:  # top up with fresh uniform columns
                    extra = rng.integers(0, n, size=want - have.size, dtype=np.int64).astype(np.uint32)
                    have = np.unique(np.concatenate([have, extra]))
                draw[i, :want] = have[:want]
            mask = pos < L[:, None]
            starts = off[rs].astype(np.int64)
            # rows in rs are increasing, so the masked flatten lands contiguously per row
            flat = draw[mask]
            idx = np.repeat(starts, L) + (np.arange(flat.size) - np.repeat(np.cumsum(L) - L, L))
            cols[idx] = flat
    if values == "normal":
        vals = rng.standard_normal(cols.size)
    else:
        vals = rng.uniform(-1.0, 1.0, size=cols.size)
    return off, cols, vals


def powerlaw_rows_device(n_rows, s=1.8, clip=50000, local_frac=0.7, band=4096, seed=BASE_SEED ^ 0xDEADBEEF,
                         device="cuda"):
    """The power-law matrix of `powerlaw_rows` drawn with torch on `device`
    (seconds instead of minutes at 8M rows): same row-length law (numpy's
    Zipf(s) clipped to [1, clip]) and column law (local_frac uniform in
    [r - band, r + band] clipped to the matrix, the rest uniform), values
    N(0, 1).  A row's repeated column draws are merged instead of redrawn,
    so rows can come out slightly shorter.  Returns (offsets i64, cols i32,
    vals f64) tensors on `device`."""
    import torch
    rng = np.random.default_rng(seed)
    n = n_rows
    lens = np.minimum(rng.zipf(s, size=n), min(clip, n)).astype(np.int64)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed) & 0x7fffffffffffffff)
    rowid = torch.repeat_interleave(torch.arange(n, device=device, dtype=torch.int64),
                                    torch.from_numpy(lens).to(device))
    m = rowid.numel()
    lo = (rowid - band).clamp_(min=0)
    hi = (rowid + band + 1).clamp_(max=n)
    u = torch.rand(m, generator=g, device=device, dtype=torch.float64)
    local = lo + (u * (hi - lo).double()).long()
    del lo, hi
    u = torch.rand(m, generator=g, device=device, dtype=torch.float64)
    glob = (u * n).long()
    del u
    pick = torch.rand(m, generator=g, device=device) < local_frac
    cols = torch.where(pick, local, glob).clamp_(max=n - 1)
    del local, glob, pick
    key, _ = torch.sort(rowid * n + cols)
    del rowid, cols
    key = torch.unique_consecutive(key)
    r = key // n
    c = (key % n).to(torch.int32)
    counts = torch.bincount(r, minlength=n)
    del r, key
    off = torch.zeros(n + 1, dtype=torch.int64, device=device)
    off[1:] = torch.cumsum(counts, 0)
    vals = torch.randn(c.numel(), generator=g, device=device, dtype=torch.float64)
    return off, c, vals


BATCH_ZIPF_CLIP = 1024  # Zipf(1.6) row lengths are clipped to min(n_cols, 1024)


def batch_matrix(seed, n_rows=None, min_rows=256, max_rows=65536, device="cpu", zipf_clip=BATCH_ZIPF_CLIP):
    """One matrix of the batched workload: square, n_rows log-uniform in
    [min_rows, max_rows] (unless given), nnz/row from a 50/50 mix of uniform
    4..64 and Zipf(1.6) clipped to min(n_cols, zipf_clip), uniform columns (a
    row's repeated draws are merged, so rows can come out slightly shorter),
    values N(0, 1) rounded to FP32 (stored as FP64).  Generated with torch on
    `device`; returns (offsets u64, cols i32, vals f64) tensors there."""
    import torch
    rng = np.random.default_rng(seed)
    if n_rows is None:
        n_rows = int(np.exp(rng.uniform(np.log(min_rows), np.log(max_rows + 1))))
        n_rows = max(min_rows, min(max_rows, n_rows))
    n = n_rows
    uni = rng.integers(4, 65, size=n)
    zpf = np.minimum(rng.zipf(1.6, size=n), min(n, zipf_clip))
    lens = np.where(rng.random(n) < 0.5, uni, zpf).astype(np.int64)
    lens = np.minimum(lens, n)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed) & 0x7fffffffffffffff)
    lt = torch.from_numpy(lens).to(device)
    rowid = torch.repeat_interleave(torch.arange(n, device=device, dtype=torch.int64), lt)
    cols = (torch.rand(rowid.numel(), generator=g, device=device, dtype=torch.float64) * n).long().clamp_(max=n - 1)
    key, _ = torch.sort(rowid * n + cols)
    key = torch.unique_consecutive(key)
    r = key // n
    c = (key % n).to(torch.int32)
    counts = torch.bincount(r, minlength=n)
    off = torch.zeros(n + 1, dtype=torch.int64, device=device)
    off[1:] = torch.cumsum(counts, 0)
    vals = torch.randn(c.numel(), generator=g, device=device, dtype=torch.float32).to(torch.float64)
    return off, c, vals


def laplacian_2d(nx, ny=None):
    """5-point Laplacian on an nx x ny grid, Dirichlet boundary: diagonal 4,
    neighbours -1, row-major grid numbering."""
    ny = nx if ny is None else ny
    n = nx * ny
    i = np.arange(n, dtype=np.int64)
    gx, gy = i % nx, i // nx
    nb = [
        (i - nx, gy > 0),
        (i - 1, gx > 0),
        (i, np.ones(n, bool)),
        (i + 1, gx < nx - 1),
        (i + nx, gy < ny - 1),
    ]
    lens = sum(m.astype(np.int64) for _, m in nb)
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    cols = np.empty(int(off[-1]), np.uint32)
    vals = np.empty(int(off[-1]), np.float64)
    cursor = off[:-1].astype(np.int64).copy()
    for k, (c, m) in enumerate(nb):  # neighbours in ascending column order
        idx = cursor[m]
        cols[idx] = c[m]
        vals[idx] = 4.0 if k == 2 else -1.0
        cursor[m] += 1
    return off, cols, vals


def insert_stream_uniform(n_rows, n_cols, per_row, seed=BASE_SEED):
    """Open-mode insert stream: per_row distinct uniform columns per row,
    rows inserted in row-major order, values U(-1, 1).  (rows, cols, vals)."""
    rng = np.random.default_rng(seed)
    lens = np.full(n_rows, per_row, np.int64)
    blk = _sorted_distinct_rows(rng, lens, n_cols)
    # insertion order within a row is the draw order, not sorted: shuffle
    perm = np.argsort(rng.random(blk.shape), axis=1)
    blk = np.take_along_axis(blk, perm, axis=1)
    rows = np.repeat(np.arange(n_rows, dtype=np.int64), per_row)
    cols = blk[:, :per_row].reshape(-1)
    vals = rng.uniform(-1.0, 1.0, size=rows.size)
    return rows, cols, vals


def relock_workload(n, nnz_per_row=16, inserts_per_row=2, overwrite_frac=0.25, seed=BASE_SEED):
    """Re-lock merge: a locked n x n matrix with nnz_per_row distinct random
    columns per row, then inserts_per_row inserts per row in row-major order of
    which exactly round(overwrite_frac * total) hit an existing column of their
    row; the others hit columns new to the row and distinct from each other.
    Returns ((off, cols, vals), (rows, cols, vals) insert stream)."""
    rng = np.random.default_rng(seed)
    lens = np.full(n, nnz_per_row, np.int64)
    blk = _sorted_distinct_rows(rng, np.full(n, nnz_per_row + inserts_per_row, np.int64), n)
    # the first nnz_per_row columns (random subset) form the locked row, the
    # rest are guaranteed-new columns for the inserts
    perm = np.argsort(rng.random(blk.shape), axis=1)
    blk = np.take_along_axis(blk, perm, axis=1)
    base = np.sort(blk[:, :nnz_per_row], axis=1)
    fresh = blk[:, nnz_per_row:nnz_per_row + inserts_per_row]
    off, cols = _pack(lens, base)
    vals = rng.uniform(-1.0, 1.0, size=cols.size)
    total = n * inserts_per_row
    n_over = int(round(overwrite_frac * total))
    over = np.zeros(total, bool)
    over[rng.permutation(total)[:n_over]] = True
    over = over.reshape(n, inserts_per_row)
    pick = rng.integers(0, nnz_per_row, size=(n, inserts_per_row))
    ins_cols = np.where(over, np.take_along_axis(base, pick, axis=1), fresh).astype(np.uint32)
    ins_rows = np.repeat(np.arange(n, dtype=np.int64), inserts_per_row)
    ins_vals = rng.uniform(-1.0, 1.0, size=total)
    return (off, cols, vals), (ins_rows, ins_cols.reshape(-1), ins_vals)


def transpose_bytes(n_rows, n_cols, nnz, s_col=4, s_val=8):
    """Algorithmic HBM bytes of one transpose: A's offsets, columns and values
    read once, A^T's written once."""
    return 8 * (n_rows + 1) + nnz * (s_col + s_val) + 8 * (n_cols + 1) + nnz * (s_col + s_val)


def dense_x(n, seed=BASE_SEED ^ 0xC0FFEE):
    """x ~ U(-1, 1) with the reference's x sub-seed (bench/spira_bench.cpp:268)."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, size=n)


def spmv_bytes(n_rows, n_cols, nnz, s_val=8, s_col=4, s_ptr=4, s_x=8, s_y=8, k=1):
    """Algorithmic bytes per SpMV call, the reference's traffic model
    (BENCHMARKS.md:173-187, bench/spira_bench.cpp:302-307) at device widths:
    nnz*(s_val+s_col) + (n_rows+1)*s_ptr + k*(n_cols*s_x + n_rows*s_y)."""
    return nnz * (s_val + s_col) + (n_rows + 1) * s_ptr + k * (n_cols * s_x + n_rows * s_y)


def lock_bytes(n_rows, nnz_old, n_staged, nnz_new, s_col=4, s_val=8):
    """Algorithmic bytes of one lock(): read old CSR + read staged + write new."""
    return (nnz_old * (s_col + s_val) + (n_rows + 1) * 8 + n_staged * (4 + 8) + (n_rows + 1) * 8
            + nnz_new * (s_col + s_val) + (n_rows + 1) * 8)


# ---- seeded corpus of the single-matrix SpMV kernel selector ----------------
SPMV_CORPUS_FAMILIES = ("uniform", "banded", "stencil", "powerlaw", "blockdiag", "local_mix")
SPMV_CORPUS_MAX_NNZ = 120_000_000


def _csr_from_entries(rowid, cols, n_rows, n_cols, g, device):
    """Sorted, duplicate-free CSR from (row, column) draws (torch tensors)."""
    import torch
    key, _ = torch.sort(rowid * n_cols + cols)
    del rowid, cols
    key = torch.unique_consecutive(key)
    r = key // n_cols
    c = (key % n_cols).to(torch.int32)
    counts = torch.bincount(r, minlength=n_rows)
    del r, key
    off = torch.zeros(n_rows + 1, dtype=torch.int64, device=device)
    off[1:] = torch.cumsum(counts, 0)
    vals = torch.rand(c.numel(), generator=g, device=device, dtype=torch.float64) * 2.0 - 1.0
    return off, c, vals


def spmv_corpus_matrix(seed, device="cuda", max_rows=1 << 23):
    """One matrix of the kernel selector's corpus (family = seed % 6): uniform
    random rows; banded random rows; 2-D 5/9-point and 3-D 7-point stencils;
    power-law rows (Zipf lengths, a local/global column mix); block-diagonal
    blocks with uniform/Zipf rows (the batched workload's matrices as one CSR);
    mostly-local rows with a uniform fraction.  Sizes are log-uniform up to
    max_rows rows and SPMV_CORPUS_MAX_NNZ entries.  Returns
    (off, cols, vals, n_cols, family) with torch tensors on `device`."""
    import torch
    rng = np.random.default_rng(seed)
    fam = SPMV_CORPUS_FAMILIES[seed % len(SPMV_CORPUS_FAMILIES)]
    g = torch.Generator(device=device)
    g.manual_seed(int(seed) & 0x7fffffffffffffff)
    logn = lambda lo, hi: int(np.exp(rng.uniform(np.log(lo), np.log(hi))))

    def row_ids(lens):
        lt = torch.from_numpy(lens.astype(np.int64)).to(device)
        return torch.repeat_interleave(torch.arange(lens.size, device=device, dtype=torch.int64), lt)

    def urand(m):
        return torch.rand(m, generator=g, device=device, dtype=torch.float64)

    if fam == "stencil":
        kind = rng.integers(0, 3)
        if kind < 2:  # 2-D, 5 or 9 points
            nx = logn(64, int(np.sqrt(max_rows)))
            ny = max(1, min(logn(64, int(np.sqrt(max_rows))), max_rows // nx))
            n = nx * ny
            i = torch.arange(n, device=device, dtype=torch.int64)
            gx, gy = i % nx, i // nx
            offs = [(0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)] + ([(-1, -1), (1, -1), (-1, 1), (1, 1)] if kind else [])
            rs, cs = [], []
            for dx, dy in offs:
                m = (gx + dx >= 0) & (gx + dx < nx) & (gy + dy >= 0) & (gy + dy < ny)
                rs.append(i[m])
                cs.append(i[m] + dx + dy * nx)
        else:  # 3-D 7 points
            side = logn(16, int(round(max_rows ** (1 / 3))))
            n = side ** 3
            i = torch.arange(n, device=device, dtype=torch.int64)
            gx, gy, gz = i % side, (i // side) % side, i // (side * side)
            rs, cs = [], []
            for dx, dy, dz in ((0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                m = ((gx + dx >= 0) & (gx + dx < side) & (gy + dy >= 0) & (gy + dy < side) & (gz + dz >= 0)
                     & (gz + dz < side))
                rs.append(i[m])
                cs.append(i[m] + dx + dy * side + dz * side * side)
        off, c, v = _csr_from_entries(torch.cat(rs), torch.cat(cs), n, n, g, device)
        return off, c, v, n, fam

    n = logn(4096, max_rows)
    if fam == "uniform":
        n_cols = max(1, int(n * 2.0 ** rng.uniform(-1.0, 1.0)))
        mean = 2.0 ** rng.uniform(0.0, 6.0)
        lens = rng.integers(int(mean // 2), int(mean * 1.5) + 1, size=n)
        lens = np.minimum(lens, min(n_cols, max(1, SPMV_CORPUS_MAX_NNZ // n)))
        rid = row_ids(lens)
        cols = (urand(rid.numel()) * n_cols).long().clamp_(max=n_cols - 1)
    elif fam == "banded":
        n_cols = n
        band = logn(1, max(2, n // 4))
        mean = 2.0 ** rng.uniform(1.0, 5.5)
        lens = np.minimum(rng.integers(1, int(2 * mean) + 1, size=n), 2 * band + 1)
        lens = np.minimum(lens, max(1, SPMV_CORPUS_MAX_NNZ // n))
        rid = row_ids(lens)
        cols = (rid + ((urand(rid.numel()) * (2 * band + 1)).long() - band)).clamp_(0, n_cols - 1)
    elif fam in ("powerlaw", "local_mix"):
        n_cols = n
        band = logn(64, max(128, n // 8))
        if fam == "powerlaw":
            s = rng.uniform(1.5, 2.5)
            lens = np.minimum(rng.zipf(s, size=n), min(n, 50000))
            local = rng.uniform(0.0, 1.0)
        else:
            mean = 2.0 ** rng.uniform(1.0, 5.5)
            lens = rng.integers(1, int(2 * mean) + 1, size=n)
            local = rng.uniform(0.5, 1.0)
        if lens.sum() > SPMV_CORPUS_MAX_NNZ:
            lens = np.maximum(1, (lens * (SPMV_CORPUS_MAX_NNZ / lens.sum())).astype(np.int64))
        rid = row_ids(lens)
        lo = (rid - band).clamp_(min=0)
        hi = (rid + band + 1).clamp_(max=n_cols)
        loc = lo + (urand(rid.numel()) * (hi - lo).double()).long()
        glob = (urand(rid.numel()) * n_cols).long()
        cols = torch.where(urand(rid.numel()) < local, loc, glob).clamp_(max=n_cols - 1)
        del lo, hi, loc, glob
    else:  # blockdiag
        sizes = []
        total = 0
        while total < n:
            b = logn(256, 65536)
            sizes.append(min(b, n - total))
            total += sizes[-1]
        n_cols = n
        starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        bstart = np.repeat(starts, sizes)
        bsize = np.repeat(sizes, sizes)
        uni = rng.integers(4, 65, size=n)
        zpf = np.minimum(rng.zipf(1.6, size=n), np.minimum(bsize, 1024))
        lens = np.minimum(np.where(rng.random(n) < 0.5, uni, zpf), bsize)
        if lens.sum() > SPMV_CORPUS_MAX_NNZ:
            lens = np.maximum(1, (lens * (SPMV_CORPUS_MAX_NNZ / lens.sum())).astype(np.int64))
        rid = row_ids(lens)
        bs = torch.from_numpy(bstart).to(device)[rid]
        bz = torch.from_numpy(bsize.astype(np.int64)).to(device)[rid]
        cols = bs + (urand(rid.numel()) * bz.double()).long().clamp_(max=bz - 1)
        del bs, bz
    off, c, v = _csr_from_entries(rid, cols, n, n_cols, g, device)
    return off, c, v, n_cols, fam
