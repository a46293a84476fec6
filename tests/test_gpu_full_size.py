"""The measured workloads at their full sizes, on the GPU, sampled against the oracle's FP64 cell
list (the pattern of the reference's test/lbvh_test.cu:391-503: a large random system checked
against a host computation, here exactly instead of as a superset):

- colloid mixture, N = 2,000,000 (1% diameter 10, 99% diameter 1, volume fraction 0.3), one LBVH
  per type, the tuned per-type-pair traversals: 2000 sampled particles' rows per type pair;
- 4096 independent boxes (N per box log-uniform in [256, 65536], sum 4.95e7 points, 3.9e9 pairs,
  past 2^31) through the autotuned graph plan: every slot written, the fill's counts equal the
  count pass, 64-bit row offsets monotone in the forest's order and ending at the total, and 64
  sampled boxes (the largest among them) equal to the cell list;
- clustered, N = 4,000,000 (1000 Gaussian clusters of sigma 5 + 20% uniform), learned single
  pass: rows identical to count-then-fill and 2000 sampled rows equal to the cell list.
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


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def sorted_rows(q, k, j):
    a = np.stack([np.asarray(q, np.int64), np.asarray(k, np.int64), np.asarray(j, np.int64)], 1)
    return a[np.lexsort(a.T[::-1])]


def test_colloid_full_size_sampled(nb, orc):
    import torch
    from neighbor_amd.typed import TypedNeighborList
    from neighbor_amd.workloads import colloid_mixture
    N = 2_000_000
    d = [1.0, 10.0]
    R = [[0.5 * (d[a] + d[b]) + 0.4 for b in range(2)] for a in range(2)]
    log("colloid N=2,000,000: generate")
    pts, types, L = colloid_mixture(N, seed=7)
    nl = TypedNeighborList(2)
    nl.build(torch.from_numpy(pts).cuda(), L, types=types)
    log(f"  tune: {nl.tune(R, reps=1)}")
    res = nl.compute(R)
    counts = host(res["counts"]).astype(np.int64)
    start = host(res["start"]).astype(np.int64)
    per = {k: host(v).astype(np.int64) for k, v in res["per_pair"].items()}
    assert res["total"] == counts.sum()
    idx = [np.flatnonzero(types == b) for b in range(2)]
    for a in range(2):
        assert np.array_equal(counts[idx[a]], per[(a, 0)] + per[(a, 1)])
    # rows tile [0, total) without overlap
    o = np.argsort(start, kind="stable")
    assert start[o[0]] == 0 and np.array_equal(start[o][1:], (start[o] + counts[o])[:-1])
    neigh = res["neigh"]
    image = res["image"]
    rng = np.random.default_rng(17)
    # 1000 small and 1000 large particles (every large-large and large-small block is sampled)
    pick = np.concatenate([rng.choice(idx[0], 1000, replace=False), rng.choice(idx[1], 1000, replace=False)])
    for a in range(2):
        pa = pick[types[pick] == a]
        pos_in_type = np.searchsorted(idx[a], pa)
        exp = []
        for b in range(2):
            q = np.concatenate([pts[pa], np.full((pa.size, 1), R[a][b], np.float32)], 1)
            ref = orc.exact_celllist(pts[idx[b]], q, [L, L, L], R[a][b])
            assert np.array_equal(per[(a, b)][pos_in_type], ref["counts"].astype(np.int64)), (a, b)
            qi = np.repeat(pa, ref["counts"])
            exp.append(np.stack([qi, np.full(qi.size, b), ref["k"].astype(np.int64), idx[b][ref["j"]]], 1))
        exp = np.concatenate(exp)
        exp = exp[np.lexsort(exp.T[::-1])]
        got = []
        for i, p in enumerate(pa):
            s, c = int(start[p]), int(counts[p])
            nj = neigh[s:s + c].cpu().numpy().astype(np.int64)
            nk = image[s:s + c].cpu().numpy().astype(np.int64)
            n0 = int(per[(a, 0)][pos_in_type[i]])
            tb = (np.arange(c) >= n0).astype(np.int64)
            assert np.all(types[nj] == tb)            # tree 0's block, then tree 1's
            got.append(np.stack([np.full(c, p), tb, nk, nj], 1))
        got = np.concatenate(got)
        got = got[np.lexsort(got.T[::-1])]
        assert np.array_equal(got, exp), a
    log(f"  colloid: {res['total']} pairs, 2000 sampled rows equal the cell list")


def test_batched_full_size_sampled(nb, orc):
    import torch
    from neighbor_amd.batch import BatchPlan, LBVHBatch
    from neighbor_amd.workloads import batched_boxes, spheres_from_points
    r = 2.8
    log("batched 4096 boxes: generate")
    pts, offs, lo, hi = batched_boxes(nbox=4096, nmin=256, nmax=65536, density=0.85, seed=42)
    N = pts.shape[0]
    sph = spheres_from_points(pts, r)
    d_pts, d_sph = torch.from_numpy(pts).cuda(), torch.from_numpy(sph).cuda()
    d_offs = torch.from_numpy(offs.astype(np.int32)).cuda()
    d_lo, d_hi = torch.from_numpy(lo).cuda(), torch.from_numpy(hi).cuda()
    bt = LBVHBatch()
    choice = bt.tune(d_pts, d_offs, d_lo, d_hi, d_sph)
    log(f"  N={N}, tuned {choice}")
    counts0 = bt.count(d_sph).clone()
    total = int(counts0.to(torch.int64).sum().item())
    assert total > 2 ** 31
    plan = BatchPlan(bt, d_pts, d_offs, d_lo, d_hi, d_sph, capacity=total)
    plan.neigh.fill_(-1)
    plan.image.fill_(255)
    plan.launch()
    torch.cuda.synchronize()
    assert bt.error_word() == 0
    log(f"  graph step done, {total} pairs")
    # every slot written (in chunks: 3.9e9 entries)
    chunk = 1 << 30
    for s in range(0, total, chunk):
        e = min(total, s + chunk)
        assert int(plan.neigh[s:e].min().item()) >= 0 and int(plan.neigh[s:e].max().item()) < N
        assert int(plan.image[s:e].max().item()) < 27
    assert torch.equal(plan.nneigh[:N], counts0)
    # row offsets: 64-bit, monotone in the forest's order, ending at the total
    prim = bt.arrays()["primitive"].long()
    o_sorted = plan.offsets[:N][prim]
    c_sorted = plan.nneigh[:N].long()[prim]
    assert int(o_sorted[0].item()) == 0 and int(plan.offsets[N].item()) == total
    assert torch.equal(o_sorted[1:], (o_sorted + c_sorted)[:-1])
    # the fill's own counts equal the count pass (an eager fill of the same rows)
    fillc = torch.empty(N, dtype=torch.int32, device="cuda")
    bt.rows(d_sph, plan.offsets, plan.nneigh, plan.neigh, plan.image, nneigh=fillc)
    assert torch.equal(fillc, counts0)
    # 64 sampled boxes, the largest among them, against the FP64 cell list
    sizes = np.diff(offs)
    rng = np.random.default_rng(23)
    boxes = np.unique(np.concatenate([[int(np.argmax(sizes))], rng.choice(4096, 63, replace=False)]))
    c_all = host(plan.nneigh[:N]).astype(np.int64)
    o_all = host(plan.offsets[:N]).astype(np.int64)
    for b in boxes:
        s, e = int(offs[b]), int(offs[b + 1])
        L = float(hi[b, 0])
        ref = orc.exact_celllist(pts[s:e], sph[s:e], [L, L, L], r)
        assert np.array_equal(c_all[s:e], ref["counts"].astype(np.int64)), b
        lo_pos = int(o_all[s:e].min())
        hi_pos = int((o_all[s:e] + c_all[s:e]).max())
        nj = host(plan.neigh[lo_pos:hi_pos]).astype(np.int64)
        nk = host(plan.image[lo_pos:hi_pos]).astype(np.int64)
        within = np.arange(c_all[s:e].sum()) - np.repeat(np.cumsum(c_all[s:e]) - c_all[s:e], c_all[s:e])
        pos = np.repeat(o_all[s:e] - lo_pos, c_all[s:e]) + within
        q = np.repeat(np.arange(e - s), c_all[s:e])
        got = sorted_rows(q, nk[pos], nj[pos] - s)
        exp = sorted_rows(np.repeat(np.arange(e - s), ref["counts"]), ref["k"], ref["j"])
        assert np.array_equal(got, exp), b
    log(f"  {boxes.size} sampled boxes (largest {int(sizes.max())} points) equal the cell list")
    del plan
    torch.cuda.empty_cache()


def test_learned_clustered_full_size(nb, orc):
    import torch
    from neighbor_amd.api import exact_rows
    from neighbor_amd.learned import CapacityPredictor, LearnedFill, learned_exact_pairs
    from neighbor_amd.workloads import clustered_points, images27, spheres_from_points
    N, r = 4_000_000, 2.8

    def frame(seed):
        pts, L = clustered_points(N, 0.85, n_clusters=1000, sigma=5.0, background=0.2, seed=seed)
        d_pts = torch.from_numpy(pts).cuda()
        d_sph = torch.from_numpy(spheres_from_points(pts, r)).cuda()
        lb = nb.LBVH()
        lb.build(d_pts, [0, 0, 0], [L, L, L])
        tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
        tr.set_bucket_size(32)
        tr.setup(lb)
        return pts, L, d_pts, d_sph, lb, tr

    log("clustered N=4,000,000: train on seed 42")
    pts, L, d_pts, d_sph, lb, tr = frame(42)
    d_img = torch.from_numpy(images27(L, dtype=np.float64)).cuda()
    order = lb.device_pointers()["primitive"]
    counts = torch.empty(N, dtype=torch.int32, device="cuda")
    tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order, nneigh=counts)
    pred = CapacityPredictor(seed=1)
    pred.features(lb, [0, 0, 0], [L, L, L], r)
    pred.train(lb, counts, steps=300, batch=16384, lr=3e-3, seed=0)
    margin = pred.margin(lb, 0.995)
    log("  apply to seed 43")
    pts, L, d_pts, d_sph, lb, tr = frame(43)
    order = lb.device_pointers()["primitive"]
    tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order, nneigh=counts)
    offs = nb.exclusive_scan(counts, perm=order)
    total = int(offs[-1].item())
    neigh = torch.full((total,), -1, dtype=torch.int32, device="cuda")
    img = torch.full((total,), 255, dtype=torch.uint8, device="cuda")
    fillc = exact_rows(tr, lb, d_sph, d_pts, offs, counts, neigh, img, images=d_img, order=order,
                       nneigh=torch.empty(N, dtype=torch.int32, device="cuda"))
    assert torch.equal(fillc, counts)
    pred.features(lb, [0, 0, 0], [L, L, L], r)
    res = learned_exact_pairs(tr, lb, pred, d_sph, d_pts, d_img, margin, pad=2, order=order,
                              buffers=LearnedFill(N, total * 2))
    log(f"  {total} pairs, {res.n_overflow} rows overflowed")
    assert torch.equal(res.count, counts)
    # rows identical to count-then-fill (same pairs in the same order), compared on the device:
    # pair w of query q sits at offs[q] + w there and at start[q] + w here
    cl = counts.long()
    qs = torch.repeat_interleave(torch.arange(N, device="cuda"), cl)
    w = torch.arange(total, device="cuda", dtype=torch.int64) - torch.repeat_interleave(torch.cumsum(cl, 0) - cl, cl)
    ref_pos = offs[:N].index_select(0, qs) + w
    new_pos = res.start[:N].index_select(0, qs) + w
    del qs, w
    assert torch.equal(neigh.index_select(0, ref_pos), res.neigh.index_select(0, new_pos))
    assert torch.equal(img.index_select(0, ref_pos), res.image.index_select(0, new_pos))
    del ref_pos, new_pos
    # 2000 sampled rows against the FP64 cell list
    pick = np.sort(np.random.default_rng(29).choice(N, 2000, replace=False))
    sph = np.concatenate([pts[pick], np.full((pick.size, 1), r, np.float32)], 1)
    ref = orc.exact_celllist(pts, sph, [L, L, L], r)
    c = host(counts).astype(np.int64)
    assert np.array_equal(c[pick], ref["counts"].astype(np.int64))
    st = host(res.start[:N]).astype(np.int64)
    got = []
    for i, q in enumerate(pick):
        nj = res.neigh[st[q]:st[q] + c[q]].cpu().numpy()
        nk = res.image[st[q]:st[q] + c[q]].cpu().numpy()
        got.append(np.stack([np.full(c[q], i), nk.astype(np.int64), nj.astype(np.int64)], 1))
    got = np.concatenate(got)
    got = got[np.lexsort(got.T[::-1])]
    exp = sorted_rows(np.repeat(np.arange(pick.size), ref["counts"]), ref["k"], ref["j"])
    assert np.array_equal(got, exp)
