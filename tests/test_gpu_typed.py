"""GPU tests of the per-type neighbor lists (one LBVH per type, a radius per type pair).

A size-asymmetric mixture (1% spheres of diameter 10, 99% of diameter 1, volume fraction 0.3)
with the cutoffs (d_i + d_j)/2 + 0.4, particles in shuffled order (so that no type's global
indices coincide with its positions in its own tree): every (query type, tree type) block of
every particle's row must hold exactly the FP64 cell-list pairs of the oracle, and the row
counts their sum.  For each traversal candidate of the per-pair tuner, over points and over
spheres (SphereInsertOp leaf boxes), none of which may change a pair; and after the tuner's
own choice, across a rebuild on moved particles.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = [1.0, 10.0]
SKIN = 0.4
R = [[0.5 * (D[a] + D[b]) + SKIN for b in range(2)] for a in range(2)]


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mixture():
    import torch
    assert torch.cuda.is_available(), "gpu tests need a GPU"
    import neighbor_amd
    neighbor_amd.lib()
    from neighbor_amd.workloads import colloid_mixture
    pts, types, L = colloid_mixture(20000, seed=3)
    perm = np.random.default_rng(11).permutation(pts.shape[0])
    return np.ascontiguousarray(pts[perm]), types[perm], L


def oracle_rows(orc, pts, types, L, R=R):
    """(query global, tree type, image, neighbor global) quadruples of the oracle, sorted."""
    idx = [np.flatnonzero(types == b) for b in range(2)]
    rows = []
    for a in range(2):
        q = np.concatenate([pts[idx[a]], np.zeros((idx[a].size, 1), np.float32)], 1)
        for b in range(2):
            q[:, 3] = np.float32(R[a][b])
            ref = orc.exact_celllist(pts[idx[b]], q, [L, L, L], R[a][b])
            qi = np.repeat(idx[a], ref["counts"])
            rows.append(np.stack([qi, np.full(qi.size, b), ref["k"].astype(np.int64), idx[b][ref["j"]]], 1))
    allr = np.concatenate(rows)
    return allr[np.lexsort(allr.T[::-1])]


@pytest.fixture(scope="module")
def reference(orc, mixture):
    return oracle_rows(orc, *mixture)


def check(res, types, reference):
    counts = res["counts"].cpu().numpy().astype(np.int64)
    start = res["start"].cpu().numpy()
    neigh = res["neigh"].cpu().numpy().astype(np.int64)
    image = res["image"].cpu().numpy().astype(np.int64)
    assert res["total"] == reference.shape[0] == counts.sum()
    # each row = its tree-0 block, then its tree-1 block
    per = {k: v.cpu().numpy().astype(np.int64) for k, v in res["per_pair"].items()}
    first_tree_count = np.zeros(types.shape[0], np.int64)
    for a in range(2):
        ia = np.flatnonzero(types == a)
        assert np.array_equal(counts[ia], per[(a, 0)] + per[(a, 1)])
        first_tree_count[ia] = per[(a, 0)]
    # the rows tile [0, total) without overlap
    o = np.argsort(start, kind="stable")
    assert start[o[0]] == 0 and np.array_equal(start[o][1:], (start[o] + counts[o])[:-1])
    within = np.arange(counts.sum()) - np.repeat(np.cumsum(counts) - counts, counts)
    pos = np.repeat(start, counts) + within
    qi = np.repeat(np.arange(types.shape[0]), counts)
    tb = (within >= np.repeat(first_tree_count, counts)).astype(np.int64)
    got = np.stack([qi, tb, image[pos], neigh[pos]], 1)
    got = got[np.lexsort(got.T[::-1])]
    assert np.array_equal(got, reference)


@pytest.mark.parametrize("spheres", [False, True])
@pytest.mark.parametrize("cand", [0, 1, 2, 3])
def test_typed_rows_equal_fp64_cell_list(mixture, reference, spheres, cand):
    import torch
    from neighbor_amd.typed import CANDIDATES, TypedNeighborList
    pts, types, L = mixture
    log(f"typed lists spheres={spheres} candidate={CANDIDATES[cand]}")
    nl = TypedNeighborList(2, sphere_radii=[0.5 * D[0], 0.5 * D[1]] if spheres else None, choice=CANDIDATES[cand])
    nl.build(torch.from_numpy(pts).cuda(), L, types=types)
    res = nl.compute(R)
    torch.cuda.synchronize()
    check(res, types, reference)


def test_typed_tuned_and_rebuilt(orc, mixture, reference):
    import torch
    from neighbor_amd.typed import CANDIDATES, TypedNeighborList
    pts, types, L = mixture
    log("typed lists: tune, compute, rebuild on moved particles")
    nl = TypedNeighborList(2)
    d_pts = torch.from_numpy(pts).cuda()
    nl.build(d_pts, L, types=types)
    choice = nl.tune(R, reps=1)
    log(f"  choice {choice}")
    assert all(c in CANDIDATES for c in choice.values())
    check(nl.compute(R), types, reference)
    # moved particles (small displacements, wrapped into the box); types unchanged
    rng = np.random.default_rng(5)
    moved = np.mod(pts + rng.uniform(-0.05, 0.05, pts.shape).astype(np.float32), np.float32(L)).astype(np.float32)
    moved[moved >= np.float32(L)] = 0.0
    nl.build(torch.from_numpy(moved).cuda(), L)
    check(nl.compute(R), types, oracle_rows(orc, moved, types, L))
    # a larger skin outgrows the kept pair buffers: filled again into grown ones
    cap = nl.neigh.numel()
    R2 = [[0.5 * (D[a] + D[b]) + 1.2 for b in range(2)] for a in range(2)]
    res = nl.compute(R2)
    assert res["total"] > cap
    check(res, types, oracle_rows(orc, moved, types, L, R2))


def triples(res, n):
    """(i, j, k) of rows start[i] .. start[i] + counts[i] (gaps allowed), sorted."""
    counts = res["counts"].cpu().numpy().astype(np.int64)
    start = res["start"].cpu().numpy().astype(np.int64)
    neigh = res["neigh"].cpu().numpy().astype(np.int64)
    image = res["image"].cpu().numpy().astype(np.int64)
    within = np.arange(counts.sum()) - np.repeat(np.cumsum(counts) - counts, counts)
    pos = np.repeat(start, counts) + within
    out = np.stack([np.repeat(np.arange(n), counts), neigh[pos], image[pos]], 1)
    return out[np.lexsort(out.T[::-1])]


@pytest.mark.parametrize("half,shifted", [(True, False), (False, True), (True, True)])
def test_typed_half_and_shifted(orc, mixture, reference, half, shifted):
    import torch
    from neighbor_amd.typed import TypedNeighborList
    pts, types, L = mixture
    log(f"typed lists half={half} shifted={shifted}")
    rng = np.random.default_rng(9)
    shift = rng.uniform(-0.3, 0.5, pts.shape[0]).astype(np.float32) if shifted else None
    nl = TypedNeighborList(2)
    nl.build(torch.from_numpy(pts).cuda(), L, types=types)
    nl.tune(R, reps=1)
    res = nl.compute(R, half=half, shift=None if shift is None else torch.from_numpy(shift).cuda())
    got = triples(res, pts.shape[0])
    if shifted:
        exp = orc.shifted_pairs(pts, types, shift, np.array(R, np.float32), [L, L, L])
    else:
        exp = reference[:, [0, 3, 2]]
        exp = exp[np.lexsort(exp.T[::-1])]
    if half:
        exp = exp[orc.half_mask(exp[:, 0], exp[:, 1], exp[:, 2])]
    assert res["total"] == exp.shape[0]
    assert np.array_equal(got, exp)
    # the rows of each particle keep their tree order (type of the neighbor non-decreasing)
    counts = res["counts"].cpu().numpy().astype(np.int64)
    start = res["start"].cpu().numpy().astype(np.int64)
    nj = res["neigh"].cpu().numpy().astype(np.int64)
    for i in np.flatnonzero(counts > 1)[:2000]:
        t = types[nj[start[i]:start[i] + counts[i]]]
        assert np.all(np.diff(t) >= 0)


def test_pack_rows_pair_stream(mixture):
    import torch
    from neighbor_amd.api import pack_rows
    from neighbor_amd.typed import TypedNeighborList
    pts, types, L = mixture
    log("pack rows into CSR + pair stream")
    nl = TypedNeighborList(2)
    nl.build(torch.from_numpy(pts).cuda(), L, types=types)
    res = nl.compute(R, half=True)
    exp = triples(res, pts.shape[0])
    packed = pack_rows(res["start"], res["counts"], nl.neigh, nl.image, pair_stream=True)
    torch.cuda.synchronize()
    offs = packed["offsets"].cpu().numpy().astype(np.int64)
    counts = res["counts"].cpu().numpy().astype(np.int64)
    assert np.array_equal(np.diff(offs), counts)
    first = packed["first"].cpu().numpy().astype(np.int64)
    assert np.array_equal(first, np.repeat(np.arange(pts.shape[0]), counts))
    got = np.stack([first, packed["neigh"].cpu().numpy().astype(np.int64),
                    packed["image"].cpu().numpy().astype(np.int64)], 1)
    # rows keep their order: same sequence as the gapped rows read in particle order
    start = res["start"].cpu().numpy().astype(np.int64)
    within = np.arange(counts.sum()) - np.repeat(np.cumsum(counts) - counts, counts)
    pos = np.repeat(start, counts) + within
    nj, nk = res["neigh"].cpu().numpy().astype(np.int64), res["image"].cpu().numpy().astype(np.int64)
    assert np.array_equal(got[:, 1], nj[pos]) and np.array_equal(got[:, 2], nk[pos])
    assert np.array_equal(got[np.lexsort(got.T[::-1])], exp)


def test_typed_with_an_empty_type(orc, mixture, reference):
    """Three declared types, the third without particles: same rows as with two."""
    import torch
    from neighbor_amd.typed import TypedNeighborList
    pts, types, L = mixture
    log("typed lists with an empty type")
    R3 = [[R[a][b] if a < 2 and b < 2 else 3.0 for b in range(3)] for a in range(3)]
    nl = TypedNeighborList(3)
    nl.build(torch.from_numpy(pts).cuda(), L, types=types)
    res = nl.compute(R3)
    got = triples(res, pts.shape[0])
    exp = reference[:, [0, 3, 2]]
    assert np.array_equal(got, exp[np.lexsort(exp.T[::-1])])
    res = nl.compute(R3, half=True)
    assert res["total"] == int(orc.half_mask(exp[:, 0], exp[:, 1], exp[:, 2]).sum())


def test_filter_and_pack_empty_rows(mixture):
    import torch
    from neighbor_amd.api import filter_rows, pack_rows
    log("filter / pack of empty rows")
    z = torch.zeros(5, dtype=torch.int32, device="cuda")
    st = torch.zeros(5, dtype=torch.int64, device="cuda")
    nj = torch.zeros(1, dtype=torch.int32, device="cuda")
    im = torch.zeros(1, dtype=torch.uint8, device="cuda")
    out = filter_rows(st, z, nj, im, half=True)
    assert int(out.sum()) == 0
    e = torch.zeros(0, dtype=torch.int32, device="cuda")
    filter_rows(torch.zeros(0, dtype=torch.int64, device="cuda"), e, nj, im, half=True)
    packed = pack_rows(st, z, nj, im, pair_stream=True)
    assert packed["neigh"].numel() == 0 and int(packed["offsets"][-1]) == 0


@pytest.mark.parametrize("spheres", [False, True])
def test_postype_records_in_place(mixture, spheres):
    """16-byte (x, y, z, type) records read in place (per-type trees over the records through
    their type's index list, exact tests at stride 4): the same tree arrays as a build over the
    gathered float3 points, and the same rows, bit for bit, as the float3 path."""
    import torch
    from neighbor_amd.api import LBVH
    from neighbor_amd.typed import TypedNeighborList
    pts, types, L = mixture
    log(f"postype records in place, spheres={spheres}")
    rec = np.concatenate([pts, types[:, None].astype(np.float32)], 1).astype(np.float32)
    d_rec, d_pts = torch.from_numpy(rec).cuda(), torch.from_numpy(pts).cuda()
    idx = np.flatnonzero(types == 1).astype(np.int32)
    r = 5.0 if spheres else None
    a, b = LBVH(), LBVH()
    a.build_postype(d_rec, [0, 0, 0], [L, L, L], index=torch.from_numpy(idx).cuda(), radius=r)
    b.build(torch.from_numpy(np.ascontiguousarray(pts[idx])).cuda(), [0, 0, 0], [L, L, L], radius=r)
    ta, tb = a.arrays(), b.arrays()
    torch.cuda.synchronize()
    for k in ta:
        assert torch.equal(ta[k], tb[k]), k
    radii = [0.5 * D[0], 0.5 * D[1]] if spheres else None
    res = {}
    for name, data in (("records", d_rec), ("float3", d_pts)):
        nl = TypedNeighborList(2, sphere_radii=radii)
        nl.build(data, L, types=types)
        nl.tune(R, reps=1)
        out = nl.compute(R, half=True)
        torch.cuda.synchronize()
        res[name] = {k: (v.cpu().numpy() if torch.is_tensor(v) else v) for k, v in out.items() if k != "per_pair"}
        res[name]["choice"] = dict(nl.choice)
    g, f = res["records"], res["float3"]
    assert g["total"] == f["total"]
    assert np.array_equal(g["counts"], f["counts"])
    # rows as (particle, neighbor, image) sets (a different tuned traversal may order a row's hits
    # differently only if the choices differ; with equal choices the rows are identical)
    def rows(r):
        c = r["counts"].astype(np.int64)
        st = r["start"].astype(np.int64)
        within = np.arange(c.sum()) - np.repeat(np.cumsum(c) - c, c)
        pos = np.repeat(st, c) + within
        t = np.stack([np.repeat(np.arange(c.size), c), r["neigh"][pos].astype(np.int64),
                      r["image"][pos].astype(np.int64)], 1)
        return t
    ga, fa = rows(g), rows(f)
    if g["choice"] == f["choice"]:
        assert np.array_equal(ga, fa)
    assert np.array_equal(ga[np.lexsort(ga.T[::-1])], fa[np.lexsort(fa.T[::-1])])
