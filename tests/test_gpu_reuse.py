"""GPU tests of tree reuse across molecular-dynamics steps.

- refit: the hierarchy (parents, children, ropes, sorted order) is untouched, every leaf box
  is its primitive's new box (the same box a fresh build gives that primitive), every internal
  box is exactly the union of its children's; exact pairs after a refit equal the oracle's FP64
  cell list of the moved points (points and spheres);
- the displacement test equals its FP64 restatement in numpy, across periodic boundaries;
- the step driver keeps the list while no particle moved more than skin/2 (every pair within
  the cutoff is then still in it), and otherwise refits (no particle crossed the box, fewer
  than max_refits refits in a row) or rebuilds into the exact list at cutoff + skin of the
  current positions.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "gpu tests need a GPU"
    import neighbor_amd
    neighbor_amd.lib()
    return neighbor_amd


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def moved(pts, L, amp, seed, wrap=True):
    """Random steps of up to amp per axis; wrapped into [0, L), or (wrap=False) reflected so
    that no point crosses the box."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(-amp, amp, pts.shape).astype(np.float32)
    if not wrap:
        out = pts + d
        bad = (out < 0) | (out >= np.float32(L))
        out[bad] = (pts - d)[bad]
        return out.astype(np.float32)
    out = np.mod(pts + d, np.float32(L)).astype(np.float32)
    out[out >= np.float32(L)] = 0.0
    return out


@pytest.mark.parametrize("radius", [None, 0.5])
def test_refit_boxes_and_pairs(nb, orc, radius):
    from neighbor_amd.workloads import box_length, images27, spheres_from_points, uniform_points
    log(f"refit radius={radius}")
    N = 20000
    L = box_length(N, 0.85)
    p0 = uniform_points(N, L, seed=3)
    p1 = moved(p0, L, 0.8, 4)
    lb = nb.LBVH()
    lb.build(dev(p0), [0, 0, 0], [L] * 3, radius=radius)
    before = {k: host(v) for k, v in lb.arrays().items()}
    d1 = dev(p1)
    lb.refit(d1, radius=radius)
    assert lb.error_word() == 0
    after = {k: host(v) for k, v in lb.arrays().items()}
    for k in ("parent", "left", "right", "primitive", "rope", "codes"):
        assert np.array_equal(before[k], after[k]), k
    prim = after["primitive"].astype(np.int64)
    # leaf boxes: what a fresh build over the moved points gives each primitive
    fresh = nb.LBVH()
    fresh.build(d1, [0, 0, 0], [L] * 3, radius=radius)
    fa = {k: host(v) for k, v in fresh.arrays().items()}
    lo_of, hi_of = np.empty((N, 3), np.float32), np.empty((N, 3), np.float32)
    lo_of[fa["primitive"].astype(np.int64)] = fa["lo"][N - 1:]
    hi_of[fa["primitive"].astype(np.int64)] = fa["hi"][N - 1:]
    assert np.array_equal(after["lo"][N - 1:], lo_of[prim]) and np.array_equal(after["hi"][N - 1:], hi_of[prim])
    if radius is None:
        assert np.array_equal(after["lo"][N - 1:], p1[prim])
    l, r = after["left"].astype(np.int64), after["right"].astype(np.int64)
    assert np.array_equal(after["lo"][:N - 1], np.minimum(after["lo"][l], after["lo"][r]))
    assert np.array_equal(after["hi"][:N - 1], np.maximum(after["hi"][l], after["hi"][r]))
    # exact pairs on the refit tree
    R = 2.4
    tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
    tr.set_bucket_size(16)
    tr.setup(lb)
    sph = spheres_from_points(p1, R)
    counts = tr.exact_count(lb, dev(sph), d1, images=dev(images27(L, dtype=np.float64)),
                            order=lb.device_pointers()["primitive"])
    ref = orc.exact_celllist(p1, sph, [L, L, L], R, pairs=False)
    assert np.array_equal(host(counts).astype(np.int64), ref["counts"].astype(np.int64))


def test_max_displacement(nb):
    import torch
    from neighbor_amd.api import max_displacement2
    log("max displacement")
    rng = np.random.default_rng(8)
    L = 30.0
    ref = rng.uniform(0, L, (50001, 3)).astype(np.float32)
    cur = moved(ref, L, 0.3, 9)
    cur[17] = np.mod(ref[17] + np.float32(1.25), np.float32(L))     # the largest move, maybe across the box edge
    wrapped = torch.zeros((), dtype=torch.int32, device="cuda")
    for box in ([L, L, L], None):
        got = float(max_displacement2(dev(cur), dev(ref), box=box, wrapped=wrapped).item())
        d = cur.astype(np.float64) - ref.astype(np.float64)
        n = np.zeros_like(d)
        if box is not None:
            n = np.rint(d / L)
            d -= L * n
        d2 = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) * (1.0 + 2.0 ** -40)
        assert got == d2.max()
        assert int(wrapped.item()) == int(np.any(n != 0, axis=1).sum())
    assert int(np.any(np.rint((cur.astype(np.float64) - ref) / L) != 0, axis=1).sum()) > 0
    empty = torch.empty((0, 3), dtype=torch.float32, device="cuda")
    assert float(max_displacement2(empty, empty).item()) == 0.0


def pairs_ij(res, n):
    counts = res["counts"].cpu().numpy().astype(np.int64)
    start = res["start"].cpu().numpy().astype(np.int64)
    neigh = res["neigh"].cpu().numpy().astype(np.int64)
    within = np.arange(counts.sum()) - np.repeat(np.cumsum(counts) - counts, counts)
    j = neigh[np.repeat(start, counts) + within]
    return set(zip(np.repeat(np.arange(n), counts).tolist(), j.tolist()))


def test_update_keeps_refits_rebuilds(nb, orc):
    import torch
    from neighbor_amd.typed import TypedNeighborList
    from neighbor_amd.workloads import box_length, uniform_points
    log("step driver: keep / refit / build")
    N = 12000
    L = box_length(N, 0.85)
    R, skin = 2.0, 0.4
    pts = uniform_points(N, L, seed=12)
    types = np.zeros(N, np.int64)
    nl = TypedNeighborList(1)
    res = nl.update(dev(pts), L, [[R]], skin, types=types, max_refits=2)
    assert res["action"] == "built"
    actions = []
    cur = pts
    for step in range(8):
        amp = 0.05 if step % 2 == 0 else 0.3      # max displacement ~0.087 (kept) or ~0.52 (recomputed)
        cur = moved(cur, L, amp, 100 + step, wrap=(step == 7))
        res = nl.update(dev(cur), L, [[R]], skin, max_refits=2)
        actions.append(res["action"])
        q = np.concatenate([cur, np.full((N, 1), R, np.float32)], 1)
        within = orc.exact_celllist(cur, q, [L, L, L], R)
        need = set(zip(np.repeat(np.arange(N), within["counts"].astype(np.int64)).tolist(),
                       within["j"].astype(np.int64).tolist()))
        have = pairs_ij(res, N)
        assert need <= have, (step, res["action"])
        if res["action"] != "kept":
            q[:, 3] = R + skin
            full = orc.exact_celllist(cur, q, [L, L, L], R + skin)
            assert res["total"] == full["total"]
    log(f"  actions {actions}")
    # kept, refit, kept, refit, kept, built (max_refits), kept, built (points crossed the box)
    assert actions == ["kept", "refit", "kept", "refit", "kept", "built", "kept", "built"]


def test_update_rebuilds_when_parameters_change(nb, orc):
    """A kept list is valid only for the cutoffs, skin, half flag and shifts it was computed with:
    changing any of them recomputes (a larger R would otherwise miss pairs)."""
    import torch
    from neighbor_amd.typed import TypedNeighborList
    from neighbor_amd.workloads import box_length, uniform_points
    log("step driver: parameter changes force a recompute")
    N = 8000
    L = box_length(N, 0.85)
    pts = uniform_points(N, L, seed=21)
    d_pts = dev(pts)
    nl = TypedNeighborList(1)
    assert nl.update(d_pts, L, [[2.0]], 0.4, types=np.zeros(N, np.int64))["action"] == "built"
    assert nl.update(d_pts, L, [[2.0]], 0.4)["action"] == "kept"
    res = nl.update(d_pts, L, [[2.5]], 0.4)                        # larger cutoff
    assert res["action"] == "built"
    q = np.concatenate([pts, np.full((N, 1), 2.9, np.float32)], 1)
    assert res["total"] == orc.exact_celllist(pts, q, [L, L, L], 2.9)["total"]
    assert nl.update(d_pts, L, [[2.5]], 0.4)["action"] == "kept"
    assert nl.update(d_pts, L, [[2.5]], 0.6)["action"] == "built"    # larger skin
    assert nl.update(d_pts, L, [[2.5]], 0.6, half=True)["action"] == "built"
    shift = torch.zeros(N, dtype=torch.float32, device="cuda")
    assert nl.update(d_pts, L, [[2.5]], 0.6, half=True, shift=shift)["action"] == "built"
    assert nl.update(d_pts, L, [[2.5]], 0.6, half=True, shift=shift)["action"] == "kept"
    shift.fill_(0.25)                                                # changed in place
    assert nl.update(d_pts, L, [[2.5]], 0.6, half=True, shift=shift)["action"] == "built"


@pytest.mark.parametrize("N", [1, 2, 3, 64, 2049])
def test_refit_small_trees(nb, N):
    """Refit of tiny trees (the one-leaf tree, chunk edges) equals the union rule, and a
    count after it equals brute force."""
    from neighbor_amd.workloads import images27, spheres_from_points
    log(f"refit N={N}")
    rng = np.random.default_rng(N)
    L = 10.0
    p0 = rng.uniform(0, L, (N, 3)).astype(np.float32)
    p1 = moved(p0, L, 0.7, N + 1, wrap=False)
    lb = nb.LBVH()
    lb.build(dev(p0), [0, 0, 0], [L] * 3)
    d1 = dev(p1)
    lb.refit(d1)
    assert lb.error_word() == 0
    a = {k: host(v) for k, v in lb.arrays().items()}
    prim = a["primitive"].astype(np.int64)
    assert np.array_equal(a["lo"][N - 1:], p1[prim]) and np.array_equal(a["hi"][N - 1:], p1[prim])
    if N > 1:
        l, r = a["left"].astype(np.int64), a["right"].astype(np.int64)
        assert np.array_equal(a["lo"][:N - 1], np.minimum(a["lo"][l], a["lo"][r]))
        assert np.array_equal(a["hi"][:N - 1], np.maximum(a["hi"][l], a["hi"][r]))
    tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
    tr.setup(lb)
    img = images27(L, dtype=np.float64)
    sph = spheres_from_points(p1, 2.0)
    counts = host(tr.exact_count(lb, dev(sph), d1, images=dev(img))).astype(np.int64)
    d2 = ((p1[:, None, None, :].astype(np.float64) + img[None, :, None, :]) - p1[None, None, :, :].astype(np.float64))
    d2 = (d2[..., 0] * d2[..., 0] + d2[..., 1] * d2[..., 1]) + d2[..., 2] * d2[..., 2]
    exp = (d2 <= np.float64(np.float32(2.0)) ** 2).sum(axis=(1, 2))
    assert np.array_equal(counts, exp)


def test_refit_rejects_other_sizes(nb):
    from neighbor_amd import NeighborError
    lb = nb.LBVH()
    lb.build(dev(np.zeros((10, 3), np.float32)), [0, 0, 0], [1, 1, 1])
    with pytest.raises(NeighborError):
        lb.refit(dev(np.zeros((11, 3), np.float32)))


def test_large_build_structure(nb):
    """A build above the large-tree chunk threshold (1024-leaf chunks): sorted codes, leaf boxes
    equal the points, every internal box is exactly the union of its children's, every node
    but the root has the parent that lists it as a child."""
    import torch
    log("large build structure N=6.5M")
    N = 6_500_000
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    L = 200.0
    d = (torch.rand((N, 3), generator=g, device="cuda") * L).contiguous()
    lb = nb.LBVH()
    lb.build(d, [0, 0, 0], [L] * 3)
    assert lb.error_word() == 0
    a = lb.arrays()
    codes = a["codes"].to(torch.int64) & 0xFFFFFFFF
    assert bool((codes[1:] >= codes[:-1]).all())
    prim = a["primitive"].long()
    assert torch.equal(a["lo"][N - 1:], d[prim]) and torch.equal(a["hi"][N - 1:], d[prim])
    l, r = a["left"].long(), a["right"].long()
    assert torch.equal(a["lo"][:N - 1], torch.minimum(a["lo"][l], a["lo"][r]))
    assert torch.equal(a["hi"][:N - 1], torch.maximum(a["hi"][l], a["hi"][r]))
    par = a["parent"].long()
    idx = torch.arange(N - 1, device="cuda")
    assert torch.equal(par[l], idx) and torch.equal(par[r], idx)
    assert int(par[0]) == 0x7FFFFFFF        # the root has no parent
