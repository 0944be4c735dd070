"""The device scan (spira_scan_u64: tile sums, then tiles scanned with their
prefix) against numpy's cumulative sum, out of place and in place, at tile
boundaries, on both sides of the direct-prefix limit (1,024 tiles), back to
back and with words using all 64 bits."""
import numpy as np
import pytest

import spira_amd

TILE = 4096


@pytest.fixture(scope="module")
def dev():
    d = spira_amd.Device(0)
    yield d
    d.close()


def _expected(a):
    e = np.zeros(len(a) + 1, np.int64)
    e[1:] = np.cumsum(a)
    return e


SIZES = [0, 1, 2, 63, 64, 4095, TILE, TILE + 1, 3 * TILE - 7, 64 * TILE, 64 * TILE + 1, 1000 * TILE + 123,
         (1 << 22) + 5, 3_000_001]


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
def test_scan_matches_cumsum(dev, n):
    import torch
    rng = np.random.default_rng(n + 11)
    a = rng.integers(0, 1 << 20, size=n, dtype=np.int64)
    exp = _expected(a)
    src = torch.from_numpy(a).cuda()
    got = dev.scan_u64(src).cpu().numpy()
    np.testing.assert_array_equal(got, exp)
    # in place, in an (n + 1)-word buffer
    buf = torch.empty(n + 1, dtype=torch.int64, device="cuda")
    buf[:n] = src
    dev.scan_u64(buf, n=n, dst=buf)
    np.testing.assert_array_equal(buf.cpu().numpy(), exp)


@pytest.mark.gpu
def test_scan_back_to_back_and_growing(dev):
    """Many calls in a row, sizes up and down (the sums workspace grows), zeros
    and words near 2^43."""
    import torch
    rng = np.random.default_rng(5)
    for i, n in enumerate([10, 200_000, 5, 9_000_000, 4096 * 7, 0, 123_457] * 3):
        a = rng.integers(0, 1 << 16, size=n, dtype=np.int64)
        if n and i % 2:
            a[rng.integers(0, n)] = (1 << 43)
        got = dev.scan_u64(torch.from_numpy(a).cuda()).cpu().numpy()
        np.testing.assert_array_equal(got, _expected(a), err_msg=f"call {i}, n={n}")


@pytest.mark.gpu
def test_scan_full_64_bit_words(dev):
    """Words using all 64 bits (the lock's packed match << 32 | nz flags), sums
    wrapping modulo 2^64 like numpy's."""
    import torch
    rng = np.random.default_rng(9)
    n = 70 * TILE + 3
    a = (rng.integers(0, 1 << 31, size=n, dtype=np.int64) << 32) | rng.integers(0, 1 << 31, size=n, dtype=np.int64)
    got = dev.scan_u64(torch.from_numpy(a).cuda()).cpu().numpy()
    with np.errstate(over="ignore"):
        exp = _expected(a)
    np.testing.assert_array_equal(got, exp)
