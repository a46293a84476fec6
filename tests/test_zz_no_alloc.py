"""The allocation-free scan (nb_exclusive_scan_u32_u64_noalloc) against the scans that take a
workspace (nb_exclusive_scan_u32_u64, nb_exclusive_scan_permuted_u32_u64), through the C ABI.

Bar: bit-identical offsets.  Two calls in a row, the second on a smaller input, so that the
static tile sums left by the first call cannot leak into the second; numpy's cumulative sum
checks both as well.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    import torch
    assert torch.cuda.is_available(), "gpu tests need a GPU"
    from neighbor_amd import _capi
    return _capi.lib()


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _with_workspace(L, counts, perm):
    import torch
    n = counts.shape[0]
    offs = torch.full((n + 1,), -1, dtype=torch.int64, device="cuda")
    ws = torch.empty(max(L.nb_scan_scratch_bytes(n), 8), dtype=torch.uint8, device="cuda")
    if perm is None:
        rc = L.nb_exclusive_scan_u32_u64(_p(counts), n, _p(offs), _p(ws), _stream())
    else:
        rc = L.nb_exclusive_scan_permuted_u32_u64(_p(counts), _p(perm), n, _p(offs), _p(ws), _stream())
    assert rc == 0, L.nb_last_error_string()
    return offs


def _noalloc(L, counts, perm):
    import torch
    n = counts.shape[0]
    offs = torch.full((n + 1,), -1, dtype=torch.int64, device="cuda")
    rc = L.nb_exclusive_scan_u32_u64_noalloc(_p(counts) if n else None, _p(perm), n, _p(offs), _stream())
    assert rc == 0, L.nb_last_error_string()
    return offs


def _numpy_scan(c, perm):
    c = c.astype(np.uint64)
    order = np.arange(c.shape[0]) if perm is None else perm
    run = np.concatenate([[0], np.cumsum(c[order], dtype=np.uint64)])
    out = np.empty(c.shape[0] + 1, dtype=np.uint64)
    out[order] = run[:-1]
    out[-1] = run[-1]
    return out


def test_noalloc_scan_matches_workspace_scan_twice(L):
    import torch
    rng = np.random.default_rng(7)
    # first call: 3,000,001 counts (ragged last tile) in a permuted order, counts up to 2^31 so the
    # 64-bit total exceeds 2^32; second call: 5,003 counts in index order
    cases = [(3_000_001, True, 1 << 31), (5_003, False, 200)]
    for n, permuted, hi in cases:
        log(f"[no-alloc scan] n={n} permuted={permuted}")
        c = rng.integers(0, hi, size=n, dtype=np.uint64).astype(np.uint32)
        counts = torch.from_numpy(c.view(np.int32)).cuda()
        perm_np = rng.permutation(n).astype(np.uint32) if permuted else None
        perm = torch.from_numpy(perm_np.view(np.int32)).cuda() if permuted else None
        got = _noalloc(L, counts, perm)
        ref = _with_workspace(L, counts, perm)
        torch.cuda.synchronize()
        g = got.cpu().numpy().view(np.uint64)
        r = ref.cpu().numpy().view(np.uint64)
        assert np.array_equal(g, r)
        assert np.array_equal(g, _numpy_scan(c, perm_np))


def test_noalloc_scan_empty_and_limits(L):
    import torch
    offs = _noalloc(L, torch.empty(0, dtype=torch.int32, device="cuda"), None)
    torch.cuda.synchronize()
    assert offs.cpu().tolist() == [0]
    # sizes past the static array are refused before any device work
    dummy = ctypes.c_void_p(0x1000)
    rc = L.nb_exclusive_scan_u32_u64_noalloc(dummy, None, (1 << 28) + 1, dummy, None)
    assert rc == -1 and b"2^28" in L.nb_last_error_string()
    assert L.nb_exclusive_scan_u32_u64_noalloc(None, None, 5, dummy, None) == -1
