"""GPU parity of the HIP path (through the C ABI) against the oracle.

Bar: integer and index outputs bit-identical (codes, primitive order, parent/left/
right, ropes, int4 records, counts, neighbor lists in order, (neighbor, image)
pairs); float node boxes equal by value (the sign of a zero bound may depend on
merge order, see oracle.c orc_bubble); root box and bin sizes bit-identical.
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

S = 0x7FFFFFFF


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


def gpu_build(nb, pts, lo, hi, radius=None, block=256):
    lb = nb.LBVH()
    lb.build(dev(pts), lo, hi, radius=radius, block=block)
    arr = {k: host(v) for k, v in lb.arrays().items()}
    return lb, arr


def check_build(orc, arr, pts, lo, hi, radius=None):
    N = pts.shape[0]
    t = orc.build(pts, lo, hi, sphere_r=-1.0 if radius is None else radius)
    assert np.array_equal(arr["codes"].view(np.uint32), t["codes"])
    assert np.array_equal(arr["primitive"].view(np.uint32), t["primitive"])
    assert np.array_equal(arr["parent"], t["parent"])
    if N > 1:
        assert np.array_equal(arr["left"], t["left"]) and np.array_equal(arr["right"], t["right"])
    assert np.array_equal(arr["lo"], t["lo"]) and np.array_equal(arr["hi"], t["hi"])
    return t


# ------------------------------------------------------------------ reference known answers
def load_cases(golden_dir):
    with open(os.path.join(golden_dir, "reference_cases.json")) as f:
        return json.load(f)


def test_reference_basic_case_gpu(nb, golden_dir):
    import torch
    c = load_cases(golden_dir)["basic"]
    log("case basic")
    pts = np.array(c["points"], np.float32)
    for stream in (torch.cuda.Stream(), torch.cuda.default_stream()):
        with torch.cuda.stream(stream):
            lb, arr = gpu_build(nb, pts, c["lo"], c["hi"])
        assert lb.getN() == 3 and lb.getRoot() == 0
        assert list(arr["parent"]) == c["parent"]
        assert list(arr["left"]) == c["left"] and list(arr["right"]) == c["right"]
        assert list(arr["primitive"]) == c["primitive"]
    tr = nb.LBVHTraverser()
    sph = dev(np.array(c["spheres"], np.float32))
    hits = host(tr.count(lb, sph))
    assert list(hits) == c["hits"]
    d = tr.data(lb)
    rec = host(d["records"])
    bins_x = [[(int(np.uint32(r[0])) >> 20) & 0x3FF, (int(np.uint32(r[1])) >> 20) & 0x3FF] for r in rec]
    assert bins_x == c["records_x_bins_hip_path"]
    assert list(rec[:, 2]) == c["records_z"] and list(rec[:, 3]) == c["records_w"]
    lst, nn = tr.neighbor_list(lb, sph, c["max_neigh"])
    lst, nn = host(lst), host(nn)
    assert list(nn) == c["hits"]
    for q, exp in enumerate(c["list_rows"]):
        assert list(lst[q, :len(exp)]) == exp
    m = dev(np.array(c["map"] + [0] * 4, np.int32))
    tr2 = nb.LBVHTraverser()
    lst, _ = tr2.neighbor_list(lb, sph, c["max_neigh"], transform_map=m)
    lst = host(lst)
    for q, exp in enumerate(c["list_rows_mapped"]):
        assert list(lst[q, :len(exp)]) == exp


def test_reference_periodic_and_single_gpu(nb, golden_dir):
    cases = load_cases(golden_dir)
    c = cases["periodic"]
    log("case periodic")
    lb, _ = gpu_build(nb, np.array(c["points"], np.float32), c["lo"], c["hi"])
    tr = nb.LBVHTraverser()
    sph = dev(np.array(c["spheres"], np.float32))
    assert list(host(tr.count(lb, sph))) == c["hits_self"]
    from neighbor_amd.workloads import images27
    assert list(host(tr.count(lb, sph, images=dev(images27(c["box"]))))) == c["hits_27"]
    c = cases["single"]
    log("case single")
    lb, arr = gpu_build(nb, np.array(c["points"], np.float32), c["lo"], c["hi"])
    assert lb.getN() == 1 and list(arr["parent"]) == c["parent"] and list(arr["primitive"]) == c["primitive"]
    assert arr["lo"][0, 0] == np.float32(c["leaf_x"][0])
    tr = nb.LBVHTraverser()
    assert list(host(tr.count(lb, dev(np.array(c["spheres"], np.float32))))) == c["hits"]
    rec = host(tr.data(lb)["records"])
    assert list(rec[:, 2]) == c["records_z"] and list(rec[:, 3]) == c["records_w"]


def test_reference_validate_gpu(nb, orc, golden_dir):
    c = load_cases(golden_dir)["validate"]
    log("case validate")
    pts = np.load(os.path.join(golden_dir, "validate_points.npy"))
    L = np.array(c["box"], np.float32)
    lo, hi = -np.float32(0.5) * L, np.float32(0.5) * L
    lb, arr = gpu_build(nb, pts, lo, hi)
    t = check_build(orc, arr, pts, lo, hi)
    from neighbor_amd.workloads import images27
    sph = np.concatenate([pts, np.full((len(pts), 1), c["rcut"], np.float32)], axis=1)
    tr = nb.LBVHTraverser()
    hits = host(tr.count(lb, dev(sph), images=dev(images27(L))))
    ref = orc.traverse(len(pts), orc.compress(t), sph, images27(L))["counts"]
    assert np.array_equal(hits.view(np.uint32), ref)


# ------------------------------------------------------------------ build parity
@pytest.mark.parametrize("N,grid,radius,seed", [
    (2, 0, None, 0), (3, 0, None, 1), (65, 0, None, 2), (1000, 0, None, 3), (1000, 8, None, 4),
    (4097, 0, 0.3, 5), (6144, 0, None, 9), (6145, 0, None, 10), (10000, 0, None, 6), (100000, 40, None, 7),
    (262144, 0, None, 8)])
def test_build_parity(nb, orc, N, grid, radius, seed):
    log(f"build parity N={N} grid={grid} radius={radius} seed={seed}")
    rng = np.random.default_rng(seed)
    if grid:
        pts = (rng.integers(0, grid, size=(N, 3)).astype(np.float32) * np.float32(0.5))   # duplicate codes
    else:
        pts = rng.uniform(-7.0, 21.0, size=(N, 3)).astype(np.float32)
    lo, hi = [-7.0, -7.0, -7.0], [21.0, 21.0, 21.0]
    lb, arr = gpu_build(nb, pts, lo, hi, radius=radius)
    assert lb.error_word() == 0
    t = check_build(orc, arr, pts, lo, hi, radius)
    # ropes, compressed records, root box and bins
    tr = nb.LBVHTraverser()
    tr.setup(lb)
    d = tr.data(lb)
    rec, rlo, rhi, bins = orc.compress(t)
    assert np.array_equal(host(d["records"]), rec)
    assert np.array_equal(arr["rope"], rec[:, 3])
    assert np.array_equal(host(d["root_lo"]), rlo) and np.array_equal(host(d["root_hi"]), rhi)
    assert host(d["bins"]).tobytes() == bins.tobytes()


def test_build_parity_points_outside_box(nb, orc):
    """Points beyond [lo, hi] fall into the edge bins (fractionToBin clamps, LBVH.cuh:105-108);
    their leaf boxes keep the true coordinates, so the root box exceeds [lo, hi]."""
    log("build parity, points outside the box")
    rng = np.random.default_rng(13)
    pts = rng.uniform(-12.0, 26.0, size=(5000, 3)).astype(np.float32)
    lo, hi = [-7.0, -7.0, -7.0], [21.0, 21.0, 21.0]
    lb, arr = gpu_build(nb, pts, lo, hi)
    assert lb.error_word() == 0
    t = check_build(orc, arr, pts, lo, hi)
    codes = arr["codes"].view(np.uint32)
    assert codes.min() == 0 and codes.max() == (1 << 30) - 1      # clamped corners present
    tr = nb.LBVHTraverser()
    tr.setup(lb)
    d = tr.data(lb)
    rec, rlo, rhi, bins = orc.compress(t)
    assert np.array_equal(host(d["records"]), rec)
    assert np.array_equal(host(d["root_lo"]), rlo) and np.array_equal(host(d["root_hi"]), rhi)
    assert np.all(host(d["root_lo"]) < np.float32(-7.0)) and np.all(host(d["root_hi"]) > np.float32(21.0))
    sph = np.concatenate([pts[:500], np.full((500, 1), 2.5, np.float32)], axis=1)
    hits = host(tr.count(lb, dev(sph)))
    ref = orc.traverse(len(pts), (rec, rlo, rhi, bins), sph, np.zeros((1, 3), np.float32))["counts"]
    assert np.array_equal(hits.view(np.uint32), ref)


def test_build_is_deterministic_and_reusable(nb):
    log("build determinism")
    rng = np.random.default_rng(12)
    pts = rng.uniform(0, 50, size=(200000, 3)).astype(np.float32)
    lb = nb.LBVH()
    first = None
    for rep in range(3):
        lb.build(dev(pts), [0, 0, 0], [50, 50, 50])
        arr = {k: host(v) for k, v in lb.arrays().items()}
        if first is None:
            first = arr
        else:
            for k in arr:
                assert arr[k].tobytes() == first[k].tobytes(), k
    # smaller N on the same object (grow-only buffers)
    lb.build(dev(pts[:1000]), [0, 0, 0], [50, 50, 50])
    assert lb.getN() == 1000


# ------------------------------------------------------------------ traversal parity
def small_uniform(N, density, r, seed):
    from neighbor_amd.workloads import box_length, uniform_points, spheres_from_points, images27
    L = box_length(N, density)
    pts = uniform_points(N, L, seed)
    return pts, spheres_from_points(pts, r), L, images27(L)


@pytest.mark.parametrize("variant", [0, 1, 4])
def test_traverse_parity_config_s(nb, orc, variant):
    """Config S (N=10^4, rho=1, r=2.8, 27 images): counts, lists and pairs vs the oracle."""
    log(f"traverse parity S variant={variant}")
    pts, sph, L, images = small_uniform(10000, 1.0, 2.8, 42)
    lb, arr = gpu_build(nb, pts, [0, 0, 0], [L, L, L])
    t = orc.build(pts, [0, 0, 0], [L, L, L])
    comp = orc.compress(t)
    ref = orc.traverse(10000, comp, sph, images, max_neigh=96, triplets=True)
    tr = nb.LBVHTraverser(variant=variant)
    d_sph, d_img = dev(sph), dev(images)
    counts = host(tr.count(lb, d_sph, images=d_img))
    assert np.array_equal(counts.view(np.uint32), ref["counts"])
    lst, nn = tr.neighbor_list(lb, d_sph, 96, images=d_img)
    lst, nn = host(lst), host(nn)
    assert np.array_equal(nn.view(np.uint32), ref["counts"])
    mask = np.arange(96)[None, :] < np.minimum(ref["counts"], 96)[:, None]
    assert np.array_equal(lst.view(np.uint32)[mask], ref["list"].reshape(-1, 96)[mask])
    offs = nb.exclusive_scan(dev(counts))
    assert np.array_equal(host(offs), ref["offsets"].astype(np.int64))
    nj, nk = tr.pairs(lb, d_sph, offs, images=d_img)
    assert np.array_equal(host(nj).view(np.uint32), ref["prim"])
    assert np.array_equal(host(nk), ref["image"])


@pytest.mark.parametrize("variant", [0, 4])
def test_traverse_query_order_and_transform(nb, orc, variant):
    log(f"traverse order + transform variant={variant}")
    pts, sph, L, images = small_uniform(5000, 0.85, 2.8, 7)
    lb, arr = gpu_build(nb, pts, [0, 0, 0], [L, L, L])
    t = orc.build(pts, [0, 0, 0], [L, L, L])
    rng = np.random.default_rng(3)
    m = rng.permutation(5000).astype(np.uint32)
    ref = orc.traverse(5000, orc.compress(t, m), sph, images, max_neigh=128)
    tr = nb.LBVHTraverser(variant=variant)
    order = dev(arr["primitive"])       # Morton order of the self-queries
    lst, nn = tr.neighbor_list(lb, dev(sph), 128, images=dev(images), transform_map=dev(m.view(np.int32)), order=order)
    assert np.array_equal(host(nn).view(np.uint32), ref["counts"])
    mask = np.arange(128)[None, :] < np.minimum(ref["counts"], 128)[:, None]
    assert np.array_equal(host(lst).view(np.uint32)[mask], ref["list"].reshape(-1, 128)[mask])


def test_traverse_full_size_sampled(nb, orc):
    """Config P (N=10^6, rho=0.85): full build parity; traversal of 3000 sampled queries vs the oracle."""
    log("full-size P build + sampled traversal")
    pts, sph, L, images = small_uniform(1_000_000, 0.85, 2.8, 42)
    lb, arr = gpu_build(nb, pts, [0, 0, 0], [L, L, L])
    assert lb.error_word() == 0
    t = check_build(orc, arr, pts, [0, 0, 0], [L, L, L])
    comp = orc.compress(t)
    tr = nb.LBVHTraverser()
    counts = host(tr.count(lb, dev(sph), images=dev(images)))
    rng = np.random.default_rng(5)
    pick = rng.choice(1_000_000, 3000, replace=False)
    ref = orc.traverse(1_000_000, comp, sph[pick], images)
    assert np.array_equal(counts.view(np.uint32)[pick], ref["counts"])
    # size-independent property: the total hit count equals the sum over the count pass of the pair fill
    offs = nb.exclusive_scan(dev(counts))
    assert int(host(offs)[-1]) == int(counts.astype(np.int64).sum())


def test_build_parity_large_chunks(nb, orc):
    """Above 6M leaves the bubble runs in 1024-leaf chunks (two workgroups per CU); N = 6500001
    leaves the last sort tile and the last chunk ragged.  Whole tree and compression bit-exact
    against the oracle, counts of 2000 sampled queries over 27 images too."""
    N = 6_500_001
    log(f"large build N={N}")
    pts, sph, L, images = small_uniform(N, 0.85, 2.8, 23)
    lb, arr = gpu_build(nb, pts, [0, 0, 0], [L, L, L])
    assert lb.error_word() == 0
    t = check_build(orc, arr, pts, [0, 0, 0], [L, L, L])
    comp = orc.compress(t)
    tr = nb.LBVHTraverser()
    tr.setup(lb)
    d = tr.data(lb)
    assert np.array_equal(host(d["records"]), comp[0])
    pick = np.random.default_rng(6).choice(N, 2000, replace=False)
    counts = host(tr.count(lb, dev(sph[pick]), images=dev(images)))
    ref = orc.traverse(N, comp, sph[pick], images)
    assert np.array_equal(counts.view(np.uint32), ref["counts"])


# ------------------------------------------------------------------ exact FP64 mode
@pytest.mark.parametrize("variant,bucket", [(0, 8), (2, 8), (3, 1), (3, 8), (3, 32), (4, 32)])
def test_exact_pairs_equal_fp64_bruteforce(nb, orc, variant, bucket):
    log(f"exact pairs small variant={variant} bucket={bucket}")
    pts, sph, L, _ = small_uniform(3000, 0.85, 2.8, 17)
    from neighbor_amd.workloads import images27
    img64 = images27(L, dtype=np.float64)
    lb, _ = gpu_build(nb, pts, [0, 0, 0], [L, L, L])
    tr = nb.LBVHTraverser(variant=variant)
    tr.set_bucket_size(bucket)
    d_pts, d_sph, d_img = dev(pts), dev(sph), dev(img64)
    counts = tr.exact_count(lb, d_sph, d_pts, images=d_img)
    offs = nb.exclusive_scan(counts)
    nj, nk = tr.exact_pairs(lb, d_sph, d_pts, offs, images=d_img)
    ref = orc.exact_bruteforce(pts, sph, img64)
    c = host(counts).view(np.uint32)
    assert np.array_equal(c, ref["counts"])
    # per query, sort (image, j) to compare sets (emission order is image-major then leaf order)
    nj, nk, o = host(nj).view(np.uint32), host(nk), host(offs)
    for q in range(3000):
        a = sorted(zip(nk[o[q]:o[q + 1]].tolist(), nj[o[q]:o[q + 1]].tolist()))
        b = sorted(zip(ref["k"][o[q]:o[q + 1]].tolist(), ref["j"][o[q]:o[q + 1]].tolist()))
        assert a == b, q


@pytest.mark.parametrize("variant", [2, 3, 4])
def test_exact_pairs_config_s_celllist(nb, orc, variant):
    log(f"exact pairs config S vs cell list variant={variant}")
    pts, sph, L, _ = small_uniform(10000, 1.0, 2.8, 42)
    from neighbor_amd.workloads import images27
    img64 = images27(L, dtype=np.float64)
    lb, arr = gpu_build(nb, pts, [0, 0, 0], [L, L, L])
    tr = nb.LBVHTraverser(variant=variant)
    order = dev(arr["primitive"])
    counts = tr.exact_count(lb, dev(sph), dev(pts), images=dev(img64), order=order)
    ref = orc.exact_celllist(pts, sph, [L, L, L], 2.8)
    c = host(counts).view(np.uint32)
    assert np.array_equal(c, ref["counts"])
    offs = nb.exclusive_scan(counts)
    nj, nk = tr.exact_pairs(lb, dev(sph), dev(pts), offs, images=dev(img64), order=order)
    nj, nk, o = host(nj).view(np.uint32), host(nk), host(offs)
    assert int(o[-1]) == ref["total"]
    # same set per query: compare the (query, image, neighbor) triples globally after sorting
    q_of = np.repeat(np.arange(10000), c)
    got = np.lexsort((nj, nk, q_of))
    exp = np.lexsort((ref["j"], ref["k"], q_of))
    assert np.array_equal(nj[got], ref["j"][exp]) and np.array_equal(nk[got], ref["k"][exp])


def test_exact_headline_step_full_size_sampled(nb, orc):
    """The measured step at full size (N=10^6, rho=0.85, r=2.8, 27 images, wave walk with 48
    leaves per bucket, queries in Morton order): exact count, scan in traversal order, exact
    (neighbor, image) rows.  Every row is written and its count agrees with the count pass;
    2000 sampled queries' rows equal the FP64 cell list of the oracle as sets."""
    import torch
    from neighbor_amd.api import exact_rows
    from neighbor_amd.workloads import images27
    N = 1_000_000
    log("full-size exact step (wave, bucket 48), sampled vs the FP64 cell list")
    pts, sph, L, _ = small_uniform(N, 0.85, 2.8, 42)
    lb = nb.LBVH()
    d_pts, d_sph = dev(pts), dev(sph)
    d_img = dev(images27(L, dtype=np.float64))
    lb.build(d_pts, [0, 0, 0], [L, L, L])
    tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
    tr.set_bucket_size(48)
    tr.setup(lb)
    prim = lb.device_pointers()["primitive"]
    counts = torch.empty(N, dtype=torch.int32, device="cuda")
    tr.exact_count(lb, d_sph, d_pts, images=d_img, order=prim, nneigh=counts)
    offs = nb.exclusive_scan(counts, perm=prim)
    total = int(offs[-1].item())
    neigh = torch.full((total,), -1, dtype=torch.int32, device="cuda")
    img = torch.full((total,), 255, dtype=torch.uint8, device="cuda")
    fill_counts = exact_rows(tr, lb, d_sph, d_pts, offs, counts, neigh, img, images=d_img, order=prim,
                             nneigh=torch.empty(N, dtype=torch.int32, device="cuda"))
    c, o = host(counts).view(np.uint32), host(offs)
    assert np.array_equal(host(fill_counts).view(np.uint32), c)
    assert total == int(c.astype(np.int64).sum())
    nj, nk = host(neigh), host(img)
    assert nj.min() >= 0 and nj.max() < N and nk.max() < 27          # every slot written
    pick = np.sort(np.random.default_rng(7).choice(N, 2000, replace=False))
    ref = orc.exact_celllist(pts, sph[pick], [L, L, L], 2.8)
    assert np.array_equal(c[pick], ref["counts"])
    ro = np.concatenate([[0], np.cumsum(ref["counts"].astype(np.int64))])
    for i, q in enumerate(pick):
        got = sorted(zip(nk[o[q]:o[q] + c[q]].tolist(), nj[o[q]:o[q] + c[q]].tolist()))
        exp = sorted(zip(ref["k"][ro[i]:ro[i + 1]].tolist(), ref["j"][ro[i]:ro[i + 1]].tolist()))
        assert got == exp, q


def jittered_lattice(n=16, seed=3):
    """Points of a cubic lattice of spacing 1 (centers at k + 0.5), each coordinate moved by
    -2..2 ulps: neighbor distances sit at 1, 2, ... within a few ulps, inside the float band
    where only the FP64 test decides."""
    k = np.arange(n, dtype=np.float32) + np.float32(0.5)
    pts = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    steps = np.random.default_rng(seed).integers(-2, 3, size=pts.shape)
    for s in (-2, -1, 1, 2):
        m = steps == s
        for _ in range(abs(s)):
            pts[m] = np.nextafter(pts[m], np.float32(np.inf) if s > 0 else np.float32(-np.inf))
    return pts, float(n)


@pytest.mark.parametrize("variant,bucket", [(0, 8), (2, 8), (3, 1), (3, 48), (4, 32)])
@pytest.mark.parametrize("r", [1.0, 2.0])
def test_exact_band_refinement_lattice(nb, orc, variant, bucket, r):
    """Pairs at distance r within a few ulps: the float pretest leaves them in its band and the
    FP64 test decides (count pass, pair list, and the row fill of the measured step)."""
    import torch
    from neighbor_amd.api import exact_rows
    from neighbor_amd.workloads import images27
    log(f"exact band lattice variant={variant} bucket={bucket} r={r}")
    pts, L = jittered_lattice()
    N = len(pts)
    sph = np.concatenate([pts, np.full((N, 1), r, np.float32)], axis=1)
    ref = orc.exact_celllist(pts, sph, [L, L, L], r)
    assert ref["counts"].min() < ref["counts"].max()           # the jitter splits the shells
    lb, arr = gpu_build(nb, pts, [0, 0, 0], [L, L, L])
    tr = nb.LBVHTraverser(variant=variant)
    tr.set_bucket_size(bucket)
    d_pts, d_sph, d_img = dev(pts), dev(sph), dev(images27(L, dtype=np.float64))
    q_of = np.repeat(np.arange(N), ref["counts"])
    exp = np.lexsort((ref["j"], ref["k"], q_of))
    # count + pair list in query order
    counts = tr.exact_count(lb, d_sph, d_pts, images=d_img)
    c = host(counts).view(np.uint32)
    assert np.array_equal(c, ref["counts"])
    offs = nb.exclusive_scan(counts)
    nj, nk = tr.exact_pairs(lb, d_sph, d_pts, offs, images=d_img)
    nj, nk = host(nj).view(np.uint32), host(nk)
    got = np.lexsort((nj, nk, q_of))
    assert np.array_equal(nj[got], ref["j"][exp]) and np.array_equal(nk[got], ref["k"][exp])
    # rows in traversal order through the row fill
    prim = lb.device_pointers()["primitive"]
    counts = tr.exact_count(lb, d_sph, d_pts, images=d_img, order=prim)
    offs = nb.exclusive_scan(counts, perm=prim)
    total = int(offs[-1].item())
    neigh = torch.empty(total, dtype=torch.int32, device="cuda")
    img = torch.empty(total, dtype=torch.uint8, device="cuda")
    fc = exact_rows(tr, lb, d_sph, d_pts, offs, counts, neigh, img, images=d_img, order=prim)
    assert np.array_equal(host(fc).view(np.uint32), ref["counts"])
    o, nj, nk = host(offs), host(neigh).view(np.uint32), host(img)
    for q in range(0, N, 5):
        a = sorted(zip(nk[o[q]:o[q] + c[q]].tolist(), nj[o[q]:o[q] + c[q]].tolist()))
        s = ref["counts"][:q].sum(dtype=np.int64)
        b = sorted(zip(ref["k"][s:s + c[q]].tolist(), ref["j"][s:s + c[q]].tolist()))
        assert a == b, q


def test_exact_wave_transform_map(nb, orc):
    """Wave walk with MapTransformOp: emitted values are map[primitive]."""
    log("exact wave with transform map")
    pts, sph, L, _ = small_uniform(4000, 0.85, 2.8, 23)
    from neighbor_amd.workloads import images27
    img64 = images27(L, dtype=np.float64)
    lb, _ = gpu_build(nb, pts, [0, 0, 0], [L, L, L])
    m = np.random.default_rng(4).permutation(4000).astype(np.int32)
    tr = nb.LBVHTraverser(variant=3)
    d_m = dev(m)
    counts = tr.exact_count(lb, dev(sph), dev(pts), images=dev(img64), transform_map=d_m)
    offs = nb.exclusive_scan(counts)
    nj, nk = tr.exact_pairs(lb, dev(sph), dev(pts), offs, images=dev(img64), transform_map=d_m)
    ref = orc.exact_bruteforce(pts, sph, img64)
    o = host(offs)
    nj = host(nj)
    for q in range(0, 4000, 7):
        a = sorted(nj[o[q]:o[q + 1]].tolist())
        b = sorted(m[ref["j"][o[q]:o[q + 1]]].tolist())
        assert a == b, q


# ------------------------------------------------------------------ error paths
def test_error_codes(nb):
    import neighbor_amd._capi as C
    log("error codes")
    pts = np.random.default_rng(0).uniform(0, 4, size=(100, 3)).astype(np.float32)
    lb = nb.LBVH()
    with pytest.raises(C.NeighborError) as e:
        lb.build(dev(pts), [0, 0, 0], [4, 4, 4], block=100)
    assert e.value.code == C.NB_ERR_TUNING
    lb.build(dev(pts), [0, 0, 0], [4, 4, 4])
    tr = nb.LBVHTraverser()
    sph = dev(np.concatenate([pts, np.ones((100, 1), np.float32)], axis=1))
    with pytest.raises(C.NeighborError) as e:
        tr.count(lb, sph, images=dev(np.zeros((33, 3), np.float32)))
    assert e.value.code == C.NB_ERR_IMAGES
    # empty tree and empty queries are silent no-ops (LBVHTraverser.h:423-426)
    empty = nb.LBVH()
    empty.build(dev(np.zeros((0, 3), np.float32)), [0, 0, 0], [1, 1, 1])
    out = dev(np.full(100, 7, np.int32))
    tr.count(empty, sph, nneigh=out)
    assert np.all(host(out) == 7)


@pytest.mark.parametrize("N,grid,limit", [(100000, 0, 0), (300000, 0, 0), (300000, 0, 300), (300000, 24, 0),
                                           (300000, 24, 64)])
def test_span_passes_global_path(nb, orc, N, grid, limit):
    """The span passes climb lists longer than the LDS limit through global memory: with the
    limit at 0 (every span) or below the lists' sizes (some spans each way) the tree is the
    oracle's, bit for bit, one and two span levels, with and without duplicate codes; a refit
    under the same limit gives the same boxes."""
    log(f"span passes through global memory N={N} grid={grid} limit={limit}")
    rng = np.random.default_rng(N + grid + limit)
    if grid:
        pts = (rng.integers(0, grid, size=(N, 3)).astype(np.float32) * np.float32(0.5))
    else:
        pts = rng.uniform(-7.0, 21.0, size=(N, 3)).astype(np.float32)
    lo, hi = [-7.0, -7.0, -7.0], [21.0, 21.0, 21.0]
    lb = nb.LBVH()
    lb.set_span_lds_limit(limit)
    d_pts = dev(pts)
    lb.build(d_pts, lo, hi)
    assert lb.error_word() == 0
    arr = {k: host(v) for k, v in lb.arrays().items()}
    check_build(orc, arr, pts, lo, hi)
    lb.refit(d_pts)
    assert lb.error_word() == 0
    arr2 = {k: host(v) for k, v in lb.arrays().items()}
    assert np.array_equal(arr2["lo"], arr["lo"]) and np.array_equal(arr2["hi"], arr["hi"])


@pytest.mark.parametrize("shape,limit", [("identical", 1024), ("identical", 0), ("collinear", 1024), ("collinear", 0)])
def test_span_passes_degenerate_points(nb, orc, shape, limit):
    """Two span levels over degenerate inputs: 300,000 identical points (every code equal: the
    tree splits by position alone) and 300,000 points on a line (1024 distinct codes), through
    the LDS climb and through global memory; bit-exact against the oracle."""
    N = 300_000
    log(f"span passes, {shape} points, limit {limit}")
    rng = np.random.default_rng(77)
    if shape == "identical":
        pts = np.full((N, 3), 3.25, np.float32)
    else:
        pts = np.zeros((N, 3), np.float32)
        pts[:, 0] = rng.uniform(0.0, 10.0, N).astype(np.float32)
    lo, hi = [0.0, 0.0, 0.0], [10.0, 10.0, 10.0]
    lb = nb.LBVH()
    lb.set_span_lds_limit(limit)
    lb.build(dev(pts), lo, hi)
    assert lb.error_word() == 0
    arr = {k: host(v) for k, v in lb.arrays().items()}
    check_build(orc, arr, pts, lo, hi)


def test_build_check_reports_the_error_word(nb):
    """nb_lbvh_check: NB_OK after a good build; NB_ERR_INTERNAL once the error word is nonzero
    (set from the host here, as a bounded wait that gave up would set it)."""
    import ctypes
    import neighbor_amd._capi as C
    log("build check")
    pts = np.random.default_rng(1).uniform(0, 10, size=(5000, 3)).astype(np.float32)
    lb = nb.LBVH()
    lb.build(dev(pts), [0, 0, 0], [10, 10, 10])
    lb.check()
    ptr = ctypes.c_void_p()
    C.call("nb_lbvh_error_word", lb._h, ctypes.byref(ptr))
    one = dev(np.array([1], np.int32))
    C.memcpy_d2d(ptr.value, one.data_ptr(), 4, 0)
    with pytest.raises(C.NeighborError) as e:
        lb.check()
    assert e.value.code == C.NB_ERR_INTERNAL and "look-back" in str(e.value)
    lb.build(dev(pts), [0, 0, 0], [10, 10, 10])     # the next build clears it
    lb.check()


@pytest.mark.parametrize("bucket", [1, 2, 32])
def test_reference_cases_wave_quantized(nb, golden_dir, bucket):
    """The wave walk over quantized boxes reproduces the reference's known answers
    (hits, list order with truncation, transform map, periodic images, N=1)."""
    cases = load_cases(golden_dir)
    c = cases["basic"]
    log(f"reference cases, wave-quantized bucket={bucket}")
    lb, _ = gpu_build(nb, np.array(c["points"], np.float32), c["lo"], c["hi"])
    sph = dev(np.array(c["spheres"], np.float32))
    tr = nb.LBVHTraverser(variant=4)
    tr.set_bucket_size(bucket)
    assert list(host(tr.count(lb, sph))) == c["hits"]
    lst, nn = tr.neighbor_list(lb, sph, c["max_neigh"])
    lst = host(lst)
    for q, exp in enumerate(c["list_rows"]):
        assert list(lst[q, :len(exp)]) == exp
    tr2 = nb.LBVHTraverser(variant=4)
    tr2.set_bucket_size(bucket)
    lst, _ = tr2.neighbor_list(lb, sph, c["max_neigh"], transform_map=dev(np.array(c["map"] + [0] * 4, np.int32)))
    lst = host(lst)
    for q, exp in enumerate(c["list_rows_mapped"]):
        assert list(lst[q, :len(exp)]) == exp
    c = cases["periodic"]
    lb, _ = gpu_build(nb, np.array(c["points"], np.float32), c["lo"], c["hi"])
    from neighbor_amd.workloads import images27
    tr = nb.LBVHTraverser(variant=4)
    tr.set_bucket_size(bucket)
    sph = dev(np.array(c["spheres"], np.float32))
    assert list(host(tr.count(lb, sph))) == c["hits_self"]
    assert list(host(tr.count(lb, sph, images=dev(images27(c["box"]))))) == c["hits_27"]
    c = cases["single"]
    lb, _ = gpu_build(nb, np.array(c["points"], np.float32), c["lo"], c["hi"])
    tr = nb.LBVHTraverser(variant=4)
    tr.set_bucket_size(bucket)
    assert list(host(tr.count(lb, dev(np.array(c["spheres"], np.float32))))) == c["hits"]


def test_exact_rows_in_traversal_order(nb):
    """Rows placed by the scan in traversal order hold each query's CSR row, in order."""
    import torch
    from neighbor_amd.api import exact_rows
    from neighbor_amd.workloads import images27
    log("exact rows in traversal order")
    pts, sph, L, _ = small_uniform(20000, 0.85, 2.8, 23)
    lb, arr = gpu_build(nb, pts, [0, 0, 0], [L, L, L])
    tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
    tr.set_bucket_size(32)
    d_pts, d_sph, d_img = dev(pts), dev(sph), dev(images27(L, dtype=np.float64))
    order = lb.device_pointers()["primitive"]
    counts = tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order)
    offs = nb.exclusive_scan(counts)
    nj, nk = tr.exact_pairs(lb, d_sph, d_pts, offs, images=d_img, order=order)
    start = nb.exclusive_scan(counts, perm=order)
    total = int(offs[-1].item())
    assert int(start[-1].item()) == total
    neigh = torch.empty(total, dtype=torch.int32, device="cuda")
    image = torch.empty(total, dtype=torch.uint8, device="cuda")
    got = exact_rows(tr, lb, d_sph, d_pts, start, counts, neigh, image, images=d_img, order=order)
    c, o, s = host(counts).astype(np.int64), host(offs).astype(np.int64), host(start).astype(np.int64)
    assert np.array_equal(host(got).astype(np.int64), c)
    pos = np.repeat(s[:-1] - o[:-1], c) + np.arange(total)
    assert np.array_equal(host(neigh)[pos], host(nj)) and np.array_equal(host(image)[pos], host(nk))
    prim = arr["primitive"].view(np.uint32)
    assert np.array_equal(s[prim][1:], (s[prim] + c[prim])[:-1])   # rows tile [0, total) in Morton order


def test_wave_visit_log(nb):
    """Per-wave walk statistics of the wave variants: counts are unchanged with the log on; each
    wave with queries loads at least the root record; buckets tested <= records loaded; leaves
    per tested bucket within the bucket size; no lane finds more neighbors than its wave tested
    leaves; a log too small for the launch is refused; switching it off stops the writes."""
    import torch
    from neighbor_amd import NeighborError
    from neighbor_amd.workloads import images27
    log("wave visit log")
    N, bucket = 20000, 32
    pts, sph, L, _ = small_uniform(N, 0.85, 2.8, 29)
    lb, arr = gpu_build(nb, pts, [0, 0, 0], [L, L, L])
    tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
    tr.set_bucket_size(bucket)
    d_pts, d_sph, d_img = dev(pts), dev(sph), dev(images27(L, dtype=np.float64))
    order = lb.device_pointers()["primitive"]
    plain = tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order).clone()
    words = nb.LBVHTraverser.visit_log_words(N)
    assert words == 3 * 4 * ((N + 255) // 256)
    vlog = torch.full((words,), -1, dtype=torch.int32, device="cuda")
    tr.set_visit_log(vlog)
    got = tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order)
    torch.cuda.synchronize()
    assert torch.equal(got, plain)
    v = host(vlog).reshape(-1, 3).astype(np.int64)
    waves = (N + 63) // 64
    assert (v[:waves, 0] >= 1).all() and (v[:waves, 1] <= v[:waves, 0]).all()
    assert (v[:waves, 2] >= v[:waves, 1]).all() and (v[:waves, 2] <= bucket * v[:waves, 1]).all()
    assert (v[waves:] == 0).all()                       # launched waves without queries walk nothing
    prim = arr["primitive"].view(np.uint32).astype(np.int64)
    c_launch = np.zeros(waves * 64, np.int64)
    c_launch[:N] = host(plain).astype(np.int64)[prim]   # thread t serves query prim[t]
    per_wave_max = c_launch.reshape(waves, 64).max(axis=1)
    assert (per_wave_max <= v[:waves, 2]).all()
    offs = nb.exclusive_scan(plain)
    with pytest.raises(NeighborError):                  # the log records count traversals only
        tr.exact_pairs(lb, d_sph, d_pts, offs, images=d_img, order=order)
    tr.set_visit_log(vlog[: nb.LBVHTraverser.visit_log_words(64)])
    with pytest.raises(NeighborError):
        tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order)
    vlog.fill_(-1)
    tr.set_visit_log(None)
    tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order)
    torch.cuda.synchronize()
    assert (host(vlog) == -1).all()


@pytest.mark.parametrize("fmt", ["gsd", "npz"])
def test_snapshot_harness(nb, orc, tmp_path, fmt):
    """The snapshot harness (reference benchmark protocol, benchmark/lbvh_benchmark.cc:123-229) on
    two frames in HOOMD's centered box, read from a HOOMD-schema GSD file (the reference's input)
    or an .npz: one table row per frame with positive times and tuned block sizes within the
    reference's range; per particle, the raw hits of the reference-structure count traversal equal
    the oracle's restatement of that traversal (same box, images center first) and the exact hits
    equal the oracle's FP64 brute force over the same 27 images."""
    from neighbor_amd.gsd import write_hoomd_frames
    from neighbor_amd.harness import run, write_frames
    log(f"snapshot harness ({fmt})")
    rng = np.random.default_rng(11)
    N, rho, r = 4000, 0.85, 2.8
    L = float(np.float32((N / rho) ** (1.0 / 3.0)))
    pos = rng.uniform(-L / 2, L / 2, size=(2, N, 3)).astype(np.float32)
    path = str(tmp_path / f"frames.{fmt}")
    if fmt == "gsd":
        write_hoomd_frames(path, pos, [[L, L, L]] * 2)
    else:
        write_frames(path, pos, [L, L, L])
    res = run(path, 2, r, str(tmp_path / "out.txt"), calls=3, log=log, keep_hits=True)
    rows = [l.split() for l in open(tmp_path / "out.txt") if not l.startswith("#")]
    assert [int(x[0]) for x in rows] == [0, 1] and all(float(v) > 0 for x in rows for v in x[1:])
    for f, x in enumerate(res):
        assert 32 <= x["build_block"] <= 1024 and 32 <= x["traverse_reference_structure_block"] <= 1024
        sph = np.concatenate([pos[f], np.full((N, 1), r, np.float32)], 1)
        tree = orc.build(pos[f], np.float32(x["lo"]), np.float32(x["hi"]))
        raw = orc.traverse(N, orc.compress(tree), sph, x["images32"])["counts"]
        assert np.array_equal(x["raw_hits"].astype(np.uint32), raw), f
        exact = orc.exact_bruteforce(pos[f], sph, x["images64"], pairs=False, nthreads=16)["counts"]
        assert np.array_equal(x["exact_hits"].astype(np.uint32), exact), f
        assert np.all(x["raw_hits"] >= x["exact_hits"])


def test_cpp_template_api_example(nb):
    """The header-only template API used as a reference caller uses it (a C++ program built by
    the package build: LBVH::build with PointInsertOp, LBVHTraverser::traverse with
    SphereQueryOp, CountNeighborsOp / NeighborListOp and ImageListOp<float3> over 27 images)
    runs on the GPU and passes its own host brute-force check."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "neighbor_amd", "examples",
                       "periodic_neighbors.bin")
    assert os.path.exists(exe), "the example is built by __graft_entry__.build()"
    log("C++ template API example: periodic_neighbors.bin 4000 2.0")
    p = subprocess.run([exe, "4000", "2.0"], capture_output=True, text=True, timeout=180)
    assert p.returncode == 0, p.stderr
    assert p.stdout.startswith("OK N=4000")
