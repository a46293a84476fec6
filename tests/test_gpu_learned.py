"""GPU tests of the learned capacity (through the C ABI).

- the device MLP gradient (project backward kernel) against the FP64 oracle (oracle/mlp.py);
- training on counts of the exact count pass lowers the loss;
- the single-pass fill into predicted rows gives, per query, exactly the rows of the
  count-then-fill path (same pairs in the same emission order), with and without forced
  overflow of the predicted capacities (the fallback path);
- the gain: single pass vs count-then-fill on a clustered workload.
"""
import sys
import time

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


@pytest.mark.parametrize("batch", [1, 63, 1000, 4133])
def test_gradient_matches_fp64_oracle(nb, batch):
    from neighbor_amd.learned import CapacityPredictor, capacity_gradient
    from oracle import mlp
    log(f"mlp gradient batch={batch}")
    rng = np.random.default_rng(batch)
    p = CapacityPredictor(seed=batch).parameters()
    p = (p + rng.normal(0.0, 0.05, p.shape)).astype(np.float32)     # non-zero biases: mixed ReLU states
    X = rng.normal(0.0, 1.0, (batch, 8)).astype(np.float32)
    t = rng.uniform(3.0, 8.0, batch).astype(np.float32)
    g = host(capacity_gradient(dev(p), dev(X), dev(t)))
    loss, gref = mlp.loss_and_grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64))
    assert abs(float(g[-1]) - loss) <= 1e-4 * max(1.0, loss)
    scale = np.abs(gref).max()
    err = np.abs(g[:-1].astype(np.float64) - gref)
    assert np.all(err <= 2e-3 * np.abs(gref) + 1e-4 * scale), (err.max(), int(err.argmax()))


@pytest.mark.parametrize("n", [1, 64, 1000, 4133])
def test_matrix_core_forward_matches_fp64(nb, n):
    """The inference forward pass (layers 1 and 2 as 32 x 32 matrix-core products over 32-sample
    blocks, FP16 operands): equal to an FP64 forward of the same FP16-rounded operands to FP32
    accumulation error (a wrong operand layout or K permutation would be far off), and within
    the FP16 rounding of the plain FP64 forward."""
    from neighbor_amd.learned import CapacityPredictor, capacity_forward
    from oracle import mlp
    log(f"matrix-core forward n={n}")
    rng = np.random.default_rng(n)
    p = CapacityPredictor(seed=n).parameters()
    p = (p + rng.normal(0.0, 0.05, p.shape)).astype(np.float32)
    X = rng.normal(0.0, 1.5, (n, 8)).astype(np.float32)
    y = host(capacity_forward(dev(p), dev(X))).astype(np.float64)
    w1, b1, w2, b2, w3, b3 = mlp.unpack(p)
    f16 = lambda a: np.asarray(a, np.float32).astype(np.float16).astype(np.float64)  # noqa: E731
    h1 = f16(np.maximum(f16(X) @ f16(w1).T + b1, 0.0))
    h2 = np.maximum(h1 @ f16(w2).T + b2, 0.0)
    y16 = h2 @ w3 + b3
    scale = np.abs(h2) @ np.abs(w3) + abs(b3) + 1.0
    assert np.all(np.abs(y - y16) <= 1e-5 * scale), float(np.max(np.abs(y - y16) / scale))
    y64, _ = mlp.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.all(np.abs(y - y64) <= 2e-2 * scale), float(np.max(np.abs(y - y64) / scale))


@pytest.fixture(scope="module")
def clustered(nb):
    """120,000 clustered points (36 Gaussian clusters of sigma 5 + 20% uniform, rho 0.85) and
    the exact count-then-fill rows of their self-queries (r = 2.8, 27 images)."""
    from neighbor_amd.workloads import clustered_points, images27, spheres_from_points
    N, r = 120_000, 2.8
    pts, L = clustered_points(N, 0.85, n_clusters=36, sigma=5.0, background=0.2, seed=11)
    sph = spheres_from_points(pts, r)
    d = dict(N=N, r=r, L=L, pts=pts, d_pts=dev(pts), d_sph=dev(sph), d_img=dev(images27(L, dtype=np.float64)))
    lb = nb.LBVH()
    lb.build(d["d_pts"], [0, 0, 0], [L, L, L])
    tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
    tr.set_bucket_size(32)
    tr.setup(lb)
    order = lb.device_pointers()["primitive"]
    counts = tr.exact_count(lb, d["d_sph"], d["d_pts"], images=d["d_img"], order=order)
    offs = nb.exclusive_scan(counts)
    nj, nk = tr.exact_pairs(lb, d["d_sph"], d["d_pts"], offs, images=d["d_img"], order=order)
    d.update(lb=lb, tr=tr, order=order, counts=counts, c=host(counts).astype(np.int64), offs=host(offs),
             nj=host(nj), nk=host(nk))
    return d


def _rows_equal(res, ref):
    c = host(res.count).astype(np.int64)
    assert np.array_equal(c, ref["c"])
    start = host(res.start).astype(np.int64)
    total = int(ref["offs"][-1])
    pos = np.repeat(start - ref["offs"][:-1].astype(np.int64), ref["c"]) + np.arange(total)
    assert np.array_equal(host(res.neigh)[pos], ref["nj"])
    assert np.array_equal(host(res.image)[pos], ref["nk"])


@pytest.fixture(scope="module")
def trained(nb, clustered):
    from neighbor_amd.learned import CapacityPredictor
    d = clustered
    pred = CapacityPredictor(seed=3)
    pred.features(d["lb"], [0, 0, 0], [d["L"]] * 3, d["r"])
    f = host(pred.features_tensor())
    assert f.shape == (d["N"], 8) and np.all(np.isfinite(f))
    assert np.all((f[:, 4:7] >= 0) & (f[:, 4:7] <= 1))
    loss0 = pred.train(d["lb"], d["counts"], steps=1, batch=8192, lr=0.0)
    loss = pred.train(d["lb"], d["counts"], steps=400, batch=8192, lr=3e-3, seed=1)
    margin = pred.margin(d["lb"], 0.995)
    log(f"capacity training loss {loss0:.4f} -> {loss:.4f}, margin {margin:.3f}")
    return pred, loss0, loss, margin


def test_training_lowers_loss(trained):
    _, loss0, loss, margin = trained
    assert loss < 0.5 * loss0
    assert 0.0 <= margin < 2.0


@pytest.mark.parametrize("in_kernel", [True, False])
@pytest.mark.parametrize("margin_shift,pad", [(0.0, 2), (-1.0, 0), (-30.0, 0)])
def test_single_pass_rows_equal_count_then_fill(nb, clustered, trained, margin_shift, pad, in_kernel):
    """margin -30 makes every capacity 1: nearly all rows take the exact fallback.  in_kernel:
    overflowed rows are moved and walked again inside the fill launch; otherwise listed,
    relocated and filled by a second traversal."""
    from neighbor_amd.learned import learned_exact_pairs
    d = clustered
    pred, _, _, margin = trained
    log(f"learned single pass margin+{margin_shift} pad={pad} in_kernel={in_kernel}")
    res = learned_exact_pairs(d["tr"], d["lb"], pred, d["d_sph"], d["d_pts"], d["d_img"], margin + margin_shift,
                              pad=pad, order=d["order"], in_kernel=in_kernel)
    _rows_equal(res, d)
    if margin_shift <= -30:
        assert res.n_overflow == int(np.sum(d["c"] > 1))
    log(f"  overflowed rows {res.n_overflow}, capacity {res.capacity_total} for {int(d['offs'][-1])} pairs")


@pytest.mark.parametrize("margin_shift,pad", [(0.0, 2), (-1.0, 0), (-30.0, 0)])
def test_wave_placed_rows_equal_count_then_fill(nb, clustered, trained, margin_shift, pad):
    """Rows placed by the wave walk (predict_rows: capacity sums per group of 256 sorted
    positions and one scan over them; exact_rows_placed: each lane's start from its group's
    base and a prefix over the group) hold exactly the count-then-fill rows.  The bases equal
    the traversal-order scan of the same capacities at every 256th sorted position and at the
    end; every row that kept its predicted place starts where that scan puts it; moved rows
    lie past all predicted rows."""
    from neighbor_amd.learned import learned_exact_pairs
    d = clustered
    pred, _, _, margin = trained
    N, L, m = d["N"], d["L"], margin + margin_shift
    log(f"wave-placed rows margin+{margin_shift} pad={pad}")
    res = learned_exact_pairs(d["tr"], d["lb"], pred, d["d_sph"], d["d_pts"], d["d_img"], m, pad=pad,
                              placed=([0, 0, 0], [L] * 3, d["r"]))
    _rows_equal(res, d)
    cap, base = pred.predict_rows(d["lb"], [0, 0, 0], [L] * 3, d["r"], m, pad)
    scan = host(nb.exclusive_scan(cap, perm=d["order"])).astype(np.int64)
    prim = host(d["lb"].arrays()["primitive"]).astype(np.int64)
    b = host(base).astype(np.int64)
    G = (N + 255) // 256
    assert b.shape == (G + 1,)
    assert np.array_equal(b[:G], scan[prim[::256]]) and b[G] == scan[N]
    c = host(cap).astype(np.int64)
    start = host(res.start).astype(np.int64)
    kept = d["c"] <= c
    assert np.array_equal(start[kept], scan[:N][kept])
    assert np.all(start[~kept] >= scan[N])
    assert res.n_overflow == int(np.sum(~kept))


def test_prediction_beside_the_bubble(nb, clustered, trained):
    """Capacities and row bases predicted on a second stream that waits for the build's
    sorted-data event (nb_lbvh_wait_sorted), while the bubble passes still run on the first,
    equal those predicted after the whole build, for a new build of the same tree object."""
    import torch
    d = clustered
    pred, _, _, margin = trained
    L = d["L"]
    log("prediction on a second stream beside the bubble")
    lb = nb.LBVH()
    lb.enable_sorted_event(True)
    lb.build(d["d_pts"], [0, 0, 0], [L] * 3)
    torch.cuda.synchronize()
    cap0, base0 = pred.predict_rows(lb, [0, 0, 0], [L] * 3, d["r"], margin, 4)
    c0, b0 = host(cap0), host(base0)
    side = torch.cuda.Stream()
    cap1 = torch.empty_like(cap0)
    base1 = torch.empty_like(base0)
    for rep in range(3):
        cap1.fill_(0)
        base1.fill_(0)
        torch.cuda.synchronize()
        lb.build(d["d_pts"], [0, 0, 0], [L] * 3)         # main stream
        lb.wait_sorted(stream=side)
        with torch.cuda.stream(side):
            pred.predict_rows(lb, [0, 0, 0], [L] * 3, d["r"], margin, 4, cap1, base1, stream=side)
        torch.cuda.synchronize()
        assert np.array_equal(host(cap1), c0) and np.array_equal(host(base1), b0), rep
    lb.check()
    with pytest.raises(RuntimeError):
        nb.LBVH().wait_sorted()


@pytest.mark.parametrize("tail", [(0, 2), (10 ** 6, 4), (1, 2)])
@pytest.mark.parametrize("n", [1, 37, 64, 256, 257, 1000, 5000])
def test_wave_placed_rows_small(nb, trained, n, tail):
    """Small and ragged query counts (partial waves and groups, exact multiples of 64 and 256):
    rows equal the oracle's FP64 cell list per query, with the plain launch map and with
    groups split into blocks of 64 or 128 queries (tail: groups per XCD, split); refused on a
    per-lane traversal variant."""
    import oracle as orc
    from neighbor_amd.learned import LearnedFill, exact_rows_placed
    from neighbor_amd.workloads import box_length, images27, spheres_from_points, uniform_points
    pred = trained[0]
    L = box_length(max(n, 8), 0.85)
    pts = uniform_points(n, L, seed=n)
    r = min(2.8, 0.45 * L)
    sph = spheres_from_points(pts, r)
    log(f"wave-placed rows n={n} tail split {tail}")
    lb = nb.LBVH()
    lb.build(dev(pts), [0, 0, 0], [L] * 3)
    tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
    tr.set_bucket_size(16)
    tr.set_tail_split(*tail)
    tr.setup(lb)
    ref = orc.exact_celllist(pts, sph, [L, L, L], r)
    buf = LearnedFill(n)
    pred.predict_rows(lb, [0, 0, 0], [L] * 3, r, 0.0, 0, buf.cap, buf.base, buf.spill)
    buf.grow(int(host(buf.base)[-1]) + ref["total"] + 64, keep=False)   # every row may move
    img = dev(images27(L, dtype=np.float64))
    exact_rows_placed(tr, lb, dev(sph), dev(pts), buf.base, buf.start, buf.cap, buf.neigh, buf.image, buf.count,
                    spill=buf.spill, images=img)
    cnt = host(buf.count).astype(np.int64)
    assert np.array_equal(cnt, ref["counts"].astype(np.int64))
    used, moved = (int(v) for v in host(buf.spill))
    assert used <= buf.neigh.numel()
    start, nj, nk = host(buf.start).astype(np.int64), host(buf.neigh), host(buf.image)
    for q in range(n):
        a, e = int(ref["offsets"][q]), int(ref["offsets"][q + 1])
        got = sorted(zip(nk[start[q]:start[q] + cnt[q]].tolist(), nj[start[q]:start[q] + cnt[q]].tolist()))
        assert got == sorted(zip(ref["k"][a:e].tolist(), ref["j"][a:e].tolist())), q
    lane = nb.LBVHTraverser(variant=nb.NB_VARIANT_QUANTIZED)
    lane.setup(lb)
    with pytest.raises(RuntimeError):
        exact_rows_placed(lane, lb, dev(sph), dev(pts), buf.base, buf.start, buf.cap, buf.neigh, buf.image, buf.count,
                        spill=buf.spill, images=img)


@pytest.mark.parametrize("groups,split", [(0, 2), (1, 2), (16, 4), (10 ** 6, 2)])
def test_tail_split_walks_equal_default_map(nb, clustered, trained, groups, split):
    """Every launch map of the exact walks (groups = 0: one block per group of 256 queries;
    otherwise each XCD's last `groups` groups run as 2 or 4 smaller blocks) gives the counts,
    the count-then-fill pairs and images, and the placed rows of the default map (32 groups,
    split 2) with which the fixture's rows were made: the same rows, and the same starts for
    every row that kept its place."""
    from neighbor_amd.learned import learned_exact_pairs
    d = clustered
    pred, _, _, margin = trained
    L, tr, lb = d["L"], d["tr"], d["lb"]
    log(f"tail split groups={groups} split={split}")
    for bad in (0, 3, 8):
        with pytest.raises(RuntimeError):
            tr.set_tail_split(groups, bad)
    try:
        tr.set_tail_split(groups, split)
        cnt = host(tr.exact_count(lb, d["d_sph"], d["d_pts"], images=d["d_img"], order=d["order"])).astype(np.int64)
        assert np.array_equal(cnt, d["c"])
        nj, nk = tr.exact_pairs(lb, d["d_sph"], d["d_pts"], nb.exclusive_scan(d["counts"]), images=d["d_img"],
                                order=d["order"])
        assert np.array_equal(host(nj), d["nj"]) and np.array_equal(host(nk), d["nk"])
        res = learned_exact_pairs(tr, lb, pred, d["d_sph"], d["d_pts"], d["d_img"], margin, pad=2,
                                  placed=([0, 0, 0], [L] * 3, d["r"]))
        _rows_equal(res, d)
        start = host(res.start).astype(np.int64)
    finally:
        tr.set_tail_split(32, 2)
    res0 = learned_exact_pairs(tr, lb, pred, d["d_sph"], d["d_pts"], d["d_img"], margin, pad=2,
                               placed=([0, 0, 0], [L] * 3, d["r"]))
    start0 = host(res0.start).astype(np.int64)
    cap, _ = pred.predict_rows(lb, [0, 0, 0], [L] * 3, d["r"], margin, 2)
    kept = d["c"] <= host(cap).astype(np.int64)
    assert kept.sum() > 0.5 * d["N"]
    assert np.array_equal(start[kept], start0[kept])
    assert res.n_overflow == res0.n_overflow == int(np.sum(~kept))


def test_learned_gain_over_count_then_fill(nb, clustered, trained):
    """Train on one configuration, apply to an independent one of the same workload; the single
    pass must give the exact rows and beat count -> scan -> fill."""
    import torch
    from neighbor_amd.learned import LearnedFill, learned_exact_pairs
    from neighbor_amd.workloads import clustered_points, spheres_from_points
    d = clustered
    pred, _, _, margin = trained
    N, r, L = d["N"], d["r"], d["L"]
    pts, _ = clustered_points(N, 0.85, n_clusters=36, sigma=5.0, background=0.2, seed=12)
    d_pts, d_sph = dev(pts), dev(spheres_from_points(pts, r))
    lb = nb.LBVH()
    lb.build(d_pts, [0, 0, 0], [L, L, L])
    tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
    tr.set_bucket_size(32)
    tr.setup(lb)
    order = lb.device_pointers()["primitive"]
    pred.features(lb, [0, 0, 0], [L] * 3, r)
    counts = torch.empty(N, dtype=torch.int32, device="cuda")
    offs = torch.empty(N + 1, dtype=torch.int64, device="cuda")

    def two_pass():
        tr.exact_count(lb, d_sph, d_pts, images=d["d_img"], order=order, nneigh=counts)
        nb.exclusive_scan(counts, offs)
        total = int(offs[-1].item())
        return tr.exact_pairs(lb, d_sph, d_pts, offs, images=d["d_img"], order=order, total=total)

    nj, nk = two_pass()
    ref = dict(c=host(counts).astype(np.int64), offs=host(offs), nj=host(nj), nk=host(nk))
    buf = LearnedFill(N, int(ref["offs"][-1]) * 2)

    def one_pass():
        return learned_exact_pairs(tr, lb, pred, d_sph, d_pts, d["d_img"], margin, pad=2, order=order, buffers=buf)

    res = one_pass()
    _rows_equal(res, ref)

    def timed(f, reps=7):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            f()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    timed(two_pass, 2)
    timed(one_pass, 2)
    t2, t1 = timed(two_pass), timed(one_pass)
    total = int(ref["offs"][-1])
    log(f"count-then-fill {t2 * 1e3:.3f} ms, learned single pass {t1 * 1e3:.3f} ms (gain {t2 / t1:.2f}x); "
        f"overflow rows {res.n_overflow} of {N}, capacity {res.capacity_total / total:.3f}x pairs")
    assert res.n_overflow <= 0.02 * N
    assert t1 < 0.9 * t2


def test_direct_prediction_equals_features_then_predict(nb, clustered, trained):
    """predict_direct (features computed on the fly, weights by scalar loads, one kernel) gives
    the capacities of features() + predict() bit for bit."""
    d = clustered
    pred, _, _, margin = trained
    log("direct capacity prediction")
    pred.features(d["lb"], [0, 0, 0], [d["L"]] * 3, d["r"])
    a = host(pred.predict(d["lb"], margin, 2))
    b = host(pred.predict_direct(d["lb"], [0, 0, 0], [d["L"]] * 3, d["r"], margin, 2))
    assert np.array_equal(a, b)
    assert a.min() >= 1


def test_feature_cell_counts(nb, clustered):
    """The own-cell occupancies in the features equal a numpy histogram of the sorted Morton
    codes at the two feature levels (the run-length cell histogram)."""
    from neighbor_amd.learned import CapacityPredictor
    d = clustered
    pred = CapacityPredictor(seed=5)
    pred.features(d["lb"], [0, 0, 0], [d["L"]] * 3, d["r"])
    f = host(pred.features_tensor()).astype(np.float64)
    codes = host(d["lb"].arrays()["codes"]).view(np.uint32).astype(np.int64)
    N = codes.size
    L = np.float32(d["L"])
    mA = int(np.floor(np.log2(max(float(L / (np.float32(2.0) * np.float32(d["r"]))), 1.0))))
    mA = max(1, min(mA, 6))
    for col, m in ((0, mA), (2, mA + 1)):
        cell = codes >> (30 - 3 * m)
        counts = np.bincount(cell, minlength=1 << (3 * m))
        expect = np.float32(N) / np.float32(1 << (3 * m))
        own = np.rint(np.exp2(f[:, col] + np.log2(1.0 + float(expect))) - 1.0).astype(np.int64)
        assert np.array_equal(own, counts[cell]), m
