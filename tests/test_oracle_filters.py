"""CPU checks of the oracle's row filters (half lists, diameter-shifted cutoffs).

- shifted_pairs (cell-list candidates + FP64 test) equals an all-pairs numpy brute force of
  the same definition over the 27 images, for random shifts of both signs;
- with zero shifts it equals the unshifted exact cell list, and with a uniform dyadic shift c
  the unshifted list at radius R + 2c (the FP64 sums are exact there);
- half_mask keeps exactly one of each mirrored pair (i, j, k) / (j, i, 26 - k) and drops the
  zero-image self pair.
"""
import numpy as np
import pytest


def mixture(n, seed):
    from neighbor_amd.workloads import colloid_mixture
    return colloid_mixture(n, seed=seed)


def rows_of(ref, idx=None):
    i = np.repeat(np.arange(ref["counts"].size), ref["counts"].astype(np.int64))
    out = np.stack([i, ref["j"].astype(np.int64), ref["k"].astype(np.int64)], 1)
    return out[np.lexsort(out.T[::-1])]


def brute(orc, xyz, types, shift, radius, L):
    img = orc.images27([L, L, L], np.float64)
    n = xyz.shape[0]
    out = []
    for k in range(27):
        d2 = orc.pair_distance2(xyz[:, None, :], xyz[None, :, :], img[k][None, None, :])
        i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        r = orc.shifted_radius(radius, types, shift, i, j)
        m = (r >= 0) & (d2 <= r * r)
        ii, jj = np.nonzero(m)
        out.append(np.stack([ii, jj, np.full(ii.size, k)], 1))
    out = np.concatenate(out).astype(np.int64)
    return out[np.lexsort(out.T[::-1])]


def test_shifted_pairs_equal_bruteforce(orc):
    xyz, types, L = mixture(1500, 4)
    rng = np.random.default_rng(1)
    shift = rng.uniform(-0.3, 0.6, xyz.shape[0]).astype(np.float32)
    radius = np.array([[1.4, 5.9], [5.9, 10.4]], np.float32)
    got = orc.shifted_pairs(xyz, types, shift, radius, [L, L, L])
    assert got.shape[0] > 1000
    assert np.array_equal(got, brute(orc, xyz, types, shift, radius, L))


def test_shift_invariants(orc):
    xyz, types, L = mixture(3000, 5)
    zero = np.zeros(xyz.shape[0], np.float32)
    radius = np.full((2, 2), 2.5, np.float32)
    q = np.concatenate([xyz, np.full((xyz.shape[0], 1), 2.5, np.float32)], 1)
    plain = rows_of(orc.exact_celllist(xyz, q, [L, L, L], 2.5))
    assert np.array_equal(orc.shifted_pairs(xyz, types, zero, radius, [L, L, L]), plain)
    c = np.full(xyz.shape[0], 0.25, np.float32)
    q[:, 3] = 3.0
    wider = rows_of(orc.exact_celllist(xyz, q, [L, L, L], 3.0))
    assert np.array_equal(orc.shifted_pairs(xyz, types, c, radius, [L, L, L]), wider)


def test_half_mask_keeps_one_of_each_mirror(orc):
    xyz, types, L = mixture(2000, 6)
    q = np.concatenate([xyz, np.full((xyz.shape[0], 1), 3.0, np.float32)], 1)
    full = rows_of(orc.exact_celllist(xyz, q, [L, L, L], 3.0))
    i, j, k = full.T
    # the full list is closed under mirroring on these inputs
    mirror = np.stack([j, i, 26 - k], 1)
    assert np.array_equal(mirror[np.lexsort(mirror.T[::-1])], full)
    keep = orc.half_mask(i, j, k)
    self0 = (i == j) & (k == 13)
    assert not np.any(keep & self0)
    h = full[keep]
    hm = np.stack([h[:, 1], h[:, 0], 26 - h[:, 2]], 1)
    both = np.concatenate([h, hm])
    both = both[np.lexsort(both.T[::-1])]
    rest = full[~self0]
    assert np.array_equal(both, rest)
