"""GPU tests of batched mode (many independent periodic boxes, one forest, one launch).

- every box's subtree has the shape of the tree built for that box alone (same sorted
  order and the same set of internal-node leaf ranges);
- exact pair rows of every query equal those of the single-tree path run box by box
  (itself pinned to the oracle), for both traversal variants and several bucket sizes,
  and agree with the oracle's FP64 cell list for some boxes directly;
- the graph plan replays the whole step with the same results on every launch;
- the autotuner's choice gives the same results.
Tiny boxes (1, 2, 3 points, boxes smaller than the query sphere) are included.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 2.8


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


SIZES = [1, 2, 3, 17, 63, 64, 65, 300, 1000, 2500, 4096, 129, 7, 777]


@pytest.fixture(scope="module")
def problem(nb):
    from neighbor_amd.workloads import batched_boxes, spheres_from_points
    pts, offs, lo, hi = batched_boxes(sizes=SIZES, density=0.85, seed=5)
    sph = spheres_from_points(pts, R)
    return dict(pts=pts, offs=offs, lo=lo, hi=hi, sph=sph, d_pts=dev(pts), d_offs=dev(offs.astype(np.int32)),
                d_lo=dev(lo), d_hi=dev(hi), d_sph=dev(sph))


@pytest.fixture(scope="module")
def per_box(nb, problem):
    """Single-tree path box by box: sorted order, internal ranges, counts and rows."""
    from neighbor_amd.workloads import images27
    out = []
    for b in range(len(SIZES)):
        s, e = int(problem["offs"][b]), int(problem["offs"][b + 1])
        L = float(problem["hi"][b, 0])
        pts = problem["pts"][s:e]
        lb = nb.LBVH()
        lb.build(dev(pts), [0, 0, 0], [L, L, L])
        arr = {k: host(v) for k, v in lb.arrays().items()}
        tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
        tr.set_bucket_size(8)
        order = lb.device_pointers()["primitive"]
        d_sph, d_pts, d_img = dev(problem["sph"][s:e]), dev(pts), dev(images27(L, dtype=np.float64))
        counts = tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order)
        offs = nb.exclusive_scan(counts)
        nj, nk = tr.exact_pairs(lb, d_sph, d_pts, offs, images=d_img, order=order)
        out.append(dict(primitive=arr["primitive"].view(np.uint32), left=arr.get("left"), right=arr.get("right"),
                        counts=host(counts).astype(np.int64), offs=host(offs).astype(np.int64),
                        nj=host(nj).astype(np.int64), nk=host(nk)))
    return out


def internal_ranges(left, right, root, n_int, leaf0):
    """Set of (first, last) sorted-leaf ranges of the internal nodes under root (leaf node ids
    are >= n_int, sorted position = id - leaf0)."""
    if root >= n_int:
        return set()
    out = {}
    stack = [(int(root), False)]
    while stack:
        x, done = stack.pop()
        if x >= n_int:
            continue
        l, r = int(left[x]), int(right[x])
        if done:
            f = (l - leaf0) if l >= n_int else out[l][0]
            t = (r - leaf0) if r >= n_int else out[r][1]
            out[x] = (f, t)
        else:
            stack.extend([(x, True), (l, False), (r, False)])
    return set(out.values())


def test_forest_subtrees_match_single_trees(nb, problem, per_box):
    from neighbor_amd.batch import LBVHBatch
    log("batch forest structure")
    bt = LBVHBatch()
    bt.build(problem["d_pts"], problem["d_offs"], problem["d_lo"], problem["d_hi"])
    assert bt.error_word() == 0
    a = {k: host(v) for k, v in bt.arrays().items()}
    N = problem["pts"].shape[0]
    keys = a["keys"].view(np.uint64)
    assert np.all(np.diff(keys.astype(np.float64)) >= 0)
    prim = a["primitive"].view(np.uint32).astype(np.int64)
    for b, ref in enumerate(per_box):
        s, e = int(problem["offs"][b]), int(problem["offs"][b + 1])
        assert np.all((keys[s:e] >> np.uint64(30)) == b)
        assert np.array_equal(prim[s:e] - s, ref["primitive"].astype(np.int64)), b
        root = int(a["roots"][b])
        n = e - s
        if n == 1:
            assert root == N - 1 + s
            continue
        got = internal_ranges(a["left"], a["right"], root, N - 1, N - 1)
        got = {(f - s, l - s) for f, l in got}
        exp = internal_ranges(ref["left"], ref["right"], 0, n - 1, n - 1)
        assert got == exp, b
        assert len(got) == n - 1


def check_rows(problem, per_box, counts, offs, nj, nk, rows_by_count=False):
    """offs: CSR offsets in point order, or (rows_by_count) row starts in any order with the row
    of point q ending at offs[q] + counts[q]."""
    counts = counts.astype(np.int64)
    offs = offs.astype(np.int64)
    nj = nj.astype(np.int64)
    for b, ref in enumerate(per_box):
        s, e = int(problem["offs"][b]), int(problem["offs"][b + 1])
        assert np.array_equal(counts[s:e], ref["counts"]), b
        for q in range(e - s):
            a = offs[s + q]
            z = a + counts[s + q] if rows_by_count else offs[s + q + 1]
            ra, rz = ref["offs"][q], ref["offs"][q + 1]
            assert np.array_equal(nj[a:z] - s, ref["nj"][ra:rz]), (b, q)
            assert np.array_equal(nk[a:z], ref["nk"][ra:rz]), (b, q)


@pytest.mark.parametrize("variant,bucket", [(0, 1), (0, 8), (0, 32), (0, 64), (1, 1), (1, 4)])
def test_batch_rows_equal_single_tree_path(nb, problem, per_box, variant, bucket):
    from neighbor_amd.batch import LBVHBatch
    log(f"batch rows variant={variant} bucket={bucket}")
    bt = LBVHBatch()
    bt.build(problem["d_pts"], problem["d_offs"], problem["d_lo"], problem["d_hi"])
    bt.set_params(block=256, bucket=bucket, variant=variant)
    bt.setup()
    counts = bt.count(problem["d_sph"])
    offs = nb.exclusive_scan(counts)
    nj, nk = bt.pairs(problem["d_sph"], offs)
    check_rows(problem, per_box, host(counts), host(offs), host(nj), host(nk))


def test_batch_rows_in_forest_order(nb, problem, per_box):
    """Rows placed by the permuted scan (forest Morton order) hold the same pairs."""
    import torch
    from neighbor_amd.batch import LBVHBatch
    log("batch rows in forest order")
    bt = LBVHBatch()
    bt.build(problem["d_pts"], problem["d_offs"], problem["d_lo"], problem["d_hi"])
    bt.set_params(bucket=32)
    bt.setup()
    counts = bt.count(problem["d_sph"])
    start = nb.exclusive_scan(counts, perm=bt.primitive_address())
    total = int(start[-1].item())
    neigh = torch.empty(total, dtype=torch.int32, device="cuda")
    image = torch.empty(total, dtype=torch.uint8, device="cuda")
    got = bt.rows(problem["d_sph"], start, counts, neigh, image)
    check_rows(problem, per_box, host(got), host(start), host(neigh), host(image), rows_by_count=True)
    # the rows tile [0, total) in the forest's sorted order
    prim = host(bt.arrays()["primitive"]).view(np.uint32)
    st, c = host(start)[:-1].astype(np.int64), host(counts).astype(np.int64)
    assert np.array_equal(st[prim][1:], (st[prim] + c[prim])[:-1])


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("z,pad", [(4.0, 8), (0.0, 0), (-30.0, 0)])
def test_batch_single_pass_rows(nb, problem, per_box, variant, z, pad):
    """The single pass without a count (rows sized from each box's density, z = -30 makes every
    capacity 1 so nearly every row is moved and walked again inside the fill): the rows hold
    exactly the single-tree path's pairs in its order; the capacities follow the formula; rows
    that kept their place start at the permuted scan of the capacities, moved rows past them."""
    import math
    from neighbor_amd.batch import BatchRows, LBVHBatch
    log(f"batch single pass variant={variant} z={z} pad={pad}")
    bt = LBVHBatch()
    bt.build(problem["d_pts"], problem["d_offs"], problem["d_lo"], problem["d_hi"])
    bt.set_params(bucket=16, variant=variant)
    bt.setup()
    N = problem["pts"].shape[0]
    total = sum(int(r["offs"][-1]) for r in per_box)
    buf = BatchRows(N, 2 * (total + 64 * N))
    bt.single_pass(problem["d_sph"], buf, z=z, pad=pad)
    used, moved = (int(v) for v in host(buf.spill))
    assert used <= buf.neigh.numel()
    check_rows(problem, per_box, host(buf.count), host(buf.start)[:N], host(buf.neigh), host(buf.image),
               rows_by_count=True)
    cap = host(bt.capacity(problem["d_sph"], z=z, pad=pad)).astype(np.int64)
    offs = problem["offs"]
    for b in range(len(SIZES)):
        s, e = int(offs[b]), int(offs[b + 1])
        L = problem["hi"][b].astype(np.float64)
        lam = (e - s) / (L[0] * L[1] * L[2]) * 4.0 / 3.0 * math.pi * float(np.float32(R)) ** 3
        exp = max(1, math.ceil(lam + z * math.sqrt(lam)) + pad)
        assert np.all(np.abs(cap[s:e] - exp) <= 1), (b, cap[s], exp)     # rounding of the FP64 formula
    counts = host(buf.count).astype(np.int64)
    kept = counts <= cap
    scan = host(nb.exclusive_scan(bt.capacity(problem["d_sph"], z=z, pad=pad), perm=bt.primitive_address()))
    start = host(buf.start).astype(np.int64)
    assert np.array_equal(start[:N][kept], scan[:N][kept])
    assert np.all(start[:N][~kept] >= scan[N])
    assert moved == int(np.sum(~kept))
    if z <= -30:
        assert moved == int(np.sum(counts > 1))


def test_batch_against_fp64_cell_list(nb, orc, problem):
    from neighbor_amd.batch import LBVHBatch
    log("batch vs oracle cell list")
    bt = LBVHBatch()
    bt.build(problem["d_pts"], problem["d_offs"], problem["d_lo"], problem["d_hi"])
    bt.set_params(bucket=16)
    bt.setup()
    counts = host(bt.count(problem["d_sph"])).astype(np.int64)
    offs = nb.exclusive_scan(dev(counts.astype(np.int32)))
    nj, nk = bt.pairs(problem["d_sph"], offs)
    nj, nk, o = host(nj).astype(np.int64), host(nk), host(offs).astype(np.int64)
    for b in (7, 8, 9, 10):     # boxes of 1000 .. 4096 points (side > 2r)
        s, e = int(problem["offs"][b]), int(problem["offs"][b + 1])
        L = float(problem["hi"][b, 0])
        ref = orc.exact_celllist(problem["pts"][s:e], problem["sph"][s:e], [L, L, L], R)
        assert np.array_equal(counts[s:e], ref["counts"].astype(np.int64)), b
        q_of = np.repeat(np.arange(e - s), ref["counts"])
        got_j, got_k = nj[o[s]:o[e]] - s, nk[o[s]:o[e]]
        g = np.lexsort((got_j, got_k, q_of))
        x = np.lexsort((ref["j"], ref["k"], q_of))
        assert np.array_equal(got_j[g], ref["j"][x].astype(np.int64)) and np.array_equal(got_k[g], ref["k"][x]), b


def test_graph_plan_replays(nb, problem, per_box):
    import torch
    from neighbor_amd.batch import BatchPlan, LBVHBatch
    log("batch graph plan")
    bt = LBVHBatch()
    bt.set_params(block=256, bucket=16, variant=0)
    total = sum(int(r["offs"][-1]) for r in per_box)
    plan = BatchPlan(bt, problem["d_pts"], problem["d_offs"], problem["d_lo"], problem["d_hi"], problem["d_sph"],
                     capacity=total)
    for rep in range(3):
        plan.nneigh.fill_(7)
        plan.neigh.fill_(-1)
        plan.launch()
        torch.cuda.synchronize()
        # eager work on the same data between replays
        _ = torch.cumsum(problem["d_sph"][:, 0], 0)
        check_rows(problem, per_box, host(plan.nneigh), host(plan.offsets), host(plan.neigh)[:total],
                   host(plan.image)[:total], rows_by_count=True)
        assert int(host(plan.offsets)[-1]) == total
        log(f"  replay {rep} ok")


def test_autotuner_choice_is_exact(nb, problem, per_box):
    from neighbor_amd.batch import LBVHBatch
    log("batch autotuner")
    bt = LBVHBatch()
    choice = bt.tune(problem["d_pts"], problem["d_offs"], problem["d_lo"], problem["d_hi"], problem["d_sph"])
    log(f"  choice {choice}")
    assert choice["block"] in (128, 256, 512, 1024)
    assert choice["variant"] in (0, 1)
    assert choice["bucket"] in ((8, 16, 32, 64) if choice["variant"] == 0 else (1, 4, 8))
    assert bt.get_params()["bucket"] == choice["bucket"]
    counts = bt.count(problem["d_sph"])
    offs = nb.exclusive_scan(counts)
    nj, nk = bt.pairs(problem["d_sph"], offs)
    check_rows(problem, per_box, host(counts), host(offs), host(nj), host(nk))


def test_graph_plan_refuses_a_reallocated_batch(nb, problem):
    """A plan's graph holds the batch's buffer addresses: after a rebuild with more points the
    launch is refused (NB_ERR_INVALID, nothing launched); a rebuild that fits the buffers keeps
    the plan usable, and a replay rebuilds the plan's own problem."""
    import torch
    import neighbor_amd._capi as C
    from neighbor_amd.batch import BatchPlan, LBVHBatch
    from neighbor_amd.workloads import batched_boxes, spheres_from_points
    log("batch graph plan after reallocation")
    bt = LBVHBatch()
    bt.set_params(block=256, bucket=16, variant=0)
    bt.build(problem["d_pts"], problem["d_offs"], problem["d_lo"], problem["d_hi"])
    bt.setup()
    total = int(bt.count(problem["d_sph"]).to(torch.int64).sum().item())
    plan = BatchPlan(bt, problem["d_pts"], problem["d_offs"], problem["d_lo"], problem["d_hi"], problem["d_sph"],
                     capacity=total)
    plan.launch()
    torch.cuda.synchronize()
    bt.check()
    # smaller problem: the buffers fit, the plan stays valid
    small = 5
    bt.build(problem["d_pts"][:int(problem["offs"][small])].contiguous(), problem["d_offs"][:small + 1].contiguous(),
             problem["d_lo"][:small].contiguous(), problem["d_hi"][:small].contiguous())
    plan.launch()
    torch.cuda.synchronize()
    assert int(plan.offsets[-1].item()) == total
    # larger problem: reallocated buffers, the plan refuses to launch
    pts, offs, lo, hi = batched_boxes(sizes=[5000, 7000, 9000], density=0.85, seed=9)
    big = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (pts, offs.astype(np.int32), lo, hi)]
    bt.build(*big)
    with pytest.raises(C.NeighborError) as e:
        plan.launch()
    assert e.value.code == C.NB_ERR_INVALID and "reallocated" in str(e.value)
