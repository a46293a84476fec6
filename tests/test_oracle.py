"""CPU tests that pin the oracle (oracle/oracle.c) before it is trusted as the checker.

Every oracle function is checked against something independent of its own code:
the reference's known answers (tests/golden/reference_cases.json), a library
routine (numpy.nextafter = C nextafterf, numpy.lexsort), or an independent
construction written here (recursive radix tree, range min/max boxes, preorder
skip pointers, flat leaf scan, FP64 brute force vs cell list).
"""
import json
import os
import operator

import numpy as np
import pytest

S = 0x7FFFFFFF
F32 = np.float32


def load_cases(golden_dir):
    with open(os.path.join(golden_dir, "reference_cases.json")) as f:
        return json.load(f)


# --------------------------------------------------------------------- directed rounding
SPECIALS = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.finfo(np.float32).max, -np.finfo(np.float32).max,
                     np.finfo(np.float32).tiny, -np.finfo(np.float32).tiny, 1.4e-45, -1.4e-45, 2.9e-39, 1e-30,
                     3.0, 0.1, 1023.0, 1e38, -1e38], dtype=np.float32)


def test_next_steps_match_library_nextafterf(orc):
    rng = np.random.default_rng(0)
    rand = rng.standard_normal(2000).astype(np.float32) * np.float32(10.0) ** rng.integers(-40, 38, 2000).astype(np.float32)
    vals = np.concatenate([SPECIALS, rand.astype(np.float32)])
    for x in vals:
        dn = np.nextafter(F32(x), F32(-np.finfo(np.float32).max))
        up = np.nextafter(F32(x), F32(np.finfo(np.float32).max))
        assert F32(orc.next_down(x)).tobytes() == dn.tobytes(), x
        assert F32(orc.next_up(x)).tobytes() == up.tobytes(), x
    assert np.isnan(orc.next_down(float("nan"))) and np.isnan(orc.next_up(float("nan")))


def test_directed_ops_are_rn_then_one_step(orc):
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(500) * 100).astype(np.float32)
    b = (rng.standard_normal(500) * 3).astype(np.float32)
    c = (rng.standard_normal(500)).astype(np.float32)
    dn = lambda v: np.nextafter(F32(v), F32(-np.finfo(np.float32).max))  # noqa: E731
    up = lambda v: np.nextafter(F32(v), F32(np.finfo(np.float32).max))  # noqa: E731
    with np.errstate(all="ignore"):
        for x, y, z in zip(a, b, c):
            assert F32(orc.approx("fadd_rd", x, y)) == dn(x + y)
            assert F32(orc.approx("fadd_ru", x, y)) == up(x + y)
            assert F32(orc.approx("fsub_rd", x, y)) == dn(x - y)
            assert F32(orc.approx("fsub_ru", x, y)) == up(x - y)
            assert F32(orc.approx("fmul_rd", x, y)) == dn(x * y)
            assert F32(orc.approx("fmul_ru", x, y)) == up(x * y)
            assert F32(orc.approx("fdiv_rd", x, y)) == dn(x / y)
            assert F32(orc.approx("fdiv_ru", x, y)) == up(x / y)
            assert F32(orc.approx("frcp_rd", y)) == dn(F32(1) / y)
            # fmaf: single rounding of the exact x*y+z, computed here through float64 + exact residual
            exact = np.float64(x) * np.float64(y) + np.float64(z)   # exact for these magnitudes (48-bit products)
            assert F32(orc.approx("fmaf_rd", x, y, z)) == dn(F32(exact)) or abs(exact) > 2.0 ** 50


def test_double2float_directed(orc):
    for x in [1.000000000001, -1.000000000001, 0.1, -0.1, 1e-50, -1e-50, 3.0, 1e39]:
        rd = F32(orc.approx("double2float_rd", x))
        ru = F32(orc.approx("double2float_ru", x))
        assert np.float64(rd) <= x <= np.float64(ru) or np.isinf(ru)
        # tight: at most one float apart, equal when x is a float (1e39 rounds up to inf)
        with np.errstate(over="ignore"):
            if np.float64(F32(x)) == np.float64(x):
                assert rd == ru == F32(x)
            else:
                assert np.nextafter(rd, F32(np.inf)) == ru


def test_approx_math_known_answers(orc, golden_dir):
    cases = load_cases(golden_dir)["approx_math"]["cases"]
    ops = {"==": operator.eq, "<": operator.lt, ">": operator.gt, "<=": operator.le, ">=": operator.ge}
    for name, args, rel, ref in cases:
        val = orc.approx(name, *args)
        target = ref if name.startswith(("fdiv", "frcp")) else float(F32(ref))
        assert ops[rel](float(F32(val)), target), (name, args, rel, ref, val)


def test_saturating_conversion(orc):
    assert orc.sat_f2u(-1.0) == 0
    assert orc.sat_f2u(float("nan")) == 0
    assert orc.sat_f2u(-0.0) == 0
    assert orc.sat_f2u(1023.9) == 1023
    assert orc.sat_f2u(5e9) == 0xFFFFFFFF


# --------------------------------------------------------------------- Morton + sort
def interleave_bruteforce(x, y, z):
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def test_morton_interleave(orc):
    rng = np.random.default_rng(2)
    for x, y, z in rng.integers(0, 1024, size=(500, 3)):
        assert orc.morton3(x, y, z) == interleave_bruteforce(int(x), int(y), int(z))
    assert orc.morton3(1023, 1023, 1023) == (1 << 30) - 1


def test_morton_codes_bins(orc):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-5, 15, size=(300, 3)).astype(np.float32)
    lo, hi = np.float32([-1, 0, 2]), np.float32([11, 12, 9])
    codes = orc.morton_codes(pts, lo, hi)
    for p, c in zip(pts, codes):
        q = []
        for a in range(3):
            f = (p[a] - lo[a]) / (hi[a] - lo[a])           # float32 ops
            q.append(int(np.clip(F32(f) * F32(1023), F32(0), F32(1023))))
        assert c == interleave_bruteforce(*q)


def test_sort_is_stable_by_code_then_index(orc):
    rng = np.random.default_rng(4)
    codes = rng.integers(0, 50, size=5000).astype(np.uint32)    # many duplicates
    sc, prim = orc.sort_codes(codes)
    order = np.lexsort((np.arange(5000), codes))
    assert np.array_equal(prim, order.astype(np.uint32))
    assert np.array_equal(sc, codes[order])


# --------------------------------------------------------------------- hierarchy, boxes, ropes
def radix_tree_recursive(codes):
    """Independent top-down radix tree on the 64-bit keys (code << 32 | position)."""
    N = len(codes)
    keys = [(int(c) << 32) | i for i, c in enumerate(codes)]
    parent = np.full(2 * N - 1, -1, np.int64)
    left = np.zeros(N - 1, np.int64)
    right = np.zeros(N - 1, np.int64)
    parent[0] = S
    stack = [(0, 0, N - 1)]          # (node index, first, last)
    while stack:
        node, first, last = stack.pop()
        hb = (keys[first] ^ keys[last]).bit_length()
        # split: last position whose key shares the bits above hb-1 with keys[first]
        split = first
        for p in range(first, last):
            if (keys[first] ^ keys[p]).bit_length() < hb:
                split = p
        lc = split + N - 1 if split == first else split
        rc = split + N if split + 1 == last else split + 1
        left[node], right[node] = lc, rc
        parent[lc] = node
        parent[rc] = node
        if lc < N - 1:
            stack.append((lc, first, split))
        if rc < N - 1:
            stack.append((rc, split + 1, last))
    return parent, left, right


def node_ranges(N, left, right):
    rng = {}

    def visit(n):
        if n >= N - 1:
            i = n - (N - 1)
            rng[n] = (i, i)
            return rng[n]
        a = visit(left[n])
        b = visit(right[n])
        rng[n] = (a[0], b[1])
        return rng[n]

    import sys
    sys.setrecursionlimit(10000)
    visit(0)
    return rng


@pytest.mark.parametrize("N,levels,seed", [(2, 4, 0), (3, 4, 1), (17, 2, 2), (257, 8, 3), (1000, 1024, 4), (1000, 3, 5)])
def test_karras_matches_recursive_radix_tree(orc, N, levels, seed):
    rng = np.random.default_rng(seed)
    codes = np.sort(rng.integers(0, levels, size=N).astype(np.uint32))
    p, l, r = orc.karras(codes)
    P, Lc, Rc = radix_tree_recursive(codes)
    assert np.array_equal(p, P) and np.array_equal(l, Lc) and np.array_equal(r, Rc)


def test_bubble_boxes_are_leaf_range_unions(orc):
    rng = np.random.default_rng(6)
    pts = rng.uniform(0, 10, size=(400, 3)).astype(np.float32)
    for radius in (-1.0, 0.3):
        t = orc.build(pts, [0, 0, 0], [10, 10, 10], sphere_r=radius)
        N = 400
        ranges = node_ranges(N, t["left"], t["right"])
        leaf_lo = t["lo"][N - 1:]
        leaf_hi = t["hi"][N - 1:]
        p = pts[t["primitive"]]
        r = F32(max(radius, 0))
        exp_lo = (p - r) if radius > 0 else p
        exp_hi = (p + r) if radius > 0 else p
        assert np.array_equal(leaf_lo, exp_lo) and np.array_equal(leaf_hi, exp_hi)
        for n in range(N - 1):
            a, b = ranges[n]
            assert np.array_equal(t["lo"][n], leaf_lo[a:b + 1].min(axis=0))
            assert np.array_equal(t["hi"][n], leaf_hi[a:b + 1].max(axis=0))


def preorder_ropes(N, left, right):
    order = []
    stack = [0]
    while stack:
        n = stack.pop()
        order.append(n)
        if n < N - 1:
            stack.append(right[n])
            stack.append(left[n])
    ranges = node_ranges(N, left, right) if N > 1 else {0: (0, 0)}
    pos = {n: i for i, n in enumerate(order)}
    rope = {}
    for n in order:
        # the first node after n's subtree in preorder
        size = 2 * (ranges[n][1] - ranges[n][0]) + 1
        j = pos[n] + size
        rope[n] = order[j] if j < len(order) else S
    return rope


def test_compress_ropes_and_bins(orc):
    rng = np.random.default_rng(7)
    N = 600
    pts = rng.uniform(-3, 7, size=(N, 3)).astype(np.float32)
    t = orc.build(pts, [-3, -3, -3], [7, 7, 7])
    rec, rlo, rhi, bins = orc.compress(t)
    rope = preorder_ropes(N, t["left"], t["right"])
    assert all(rec[n, 3] == rope[n] for n in range(2 * N - 1))
    assert np.array_equal(rec[:N - 1, 2], t["left"])
    assert np.array_equal(rec[N - 1:, 2], ~t["primitive"].astype(np.int32))
    # decoded quantized boxes contain the exact boxes, and are within two bins of them
    lo3 = rec[:, 0].astype(np.uint32)
    hi3 = rec[:, 1].astype(np.uint32)
    for a, sh in enumerate((20, 10, 0)):
        lb = ((lo3 >> sh) & 0x3FF).astype(np.float32)
        hb = ((hi3 >> sh) & 0x3FF).astype(np.float32)
        dlo = rlo[a] + lb * bins[a]
        dhi = rhi[a] - hb * bins[a]
        assert np.all(dlo <= t["lo"][:, a]) and np.all(dhi >= t["hi"][:, a])
        assert np.all(t["lo"][:, a] - dlo <= 2.01 * bins[a]) and np.all(dhi - t["hi"][:, a] <= 2.01 * bins[a])


# --------------------------------------------------------------------- reference known answers
def _bins_x(rec):
    return [[(int(np.uint32(r[0])) >> 20) & 0x3FF, (int(np.uint32(r[1])) >> 20) & 0x3FF] for r in rec]


def test_reference_basic_case(orc, golden_dir):
    c = load_cases(golden_dir)["basic"]
    pts = np.array(c["points"], np.float32)
    t = orc.build(pts, c["lo"], c["hi"])
    assert list(t["parent"]) == c["parent"]
    assert list(t["left"]) == c["left"] and list(t["right"]) == c["right"]
    assert list(t["primitive"]) == c["primitive"]
    for n, (l, h) in c["leaf_x"].items():
        assert t["lo"][int(n), 0] == F32(l) and t["hi"][int(n), 0] == F32(h)
    for n, (l, h) in c["internal_x"].items():
        assert t["lo"][int(n), 0] == F32(l) and t["hi"][int(n), 0] == F32(h)
    comp = orc.compress(t)
    assert _bins_x(comp[0]) == c["records_x_bins_hip_path"]
    assert list(comp[0][:, 2]) == c["records_z"] and list(comp[0][:, 3]) == c["records_w"]
    sph = np.array(c["spheres"], np.float32)
    tr = orc.traverse(3, comp, sph, max_neigh=c["max_neigh"])
    assert list(tr["counts"]) == c["hits"]
    rows = tr["list"].reshape(-1, c["max_neigh"])
    for q, exp in enumerate(c["list_rows"]):
        assert list(rows[q, :len(exp)]) == exp
    comp_m = orc.compress(t, np.array(c["map"], np.uint32))
    rows = orc.traverse(3, comp_m, sph, max_neigh=c["max_neigh"])["list"].reshape(-1, c["max_neigh"])
    for q, exp in enumerate(c["list_rows_mapped"]):
        assert list(rows[q, :len(exp)]) == exp


def test_reference_periodic_case(orc, golden_dir):
    c = load_cases(golden_dir)["periodic"]
    t = orc.build(np.array(c["points"], np.float32), c["lo"], c["hi"])
    comp = orc.compress(t)
    sph = np.array(c["spheres"], np.float32)
    assert list(orc.traverse(3, comp, sph)["counts"]) == c["hits_self"]
    assert list(orc.traverse(3, comp, sph, orc.images27(c["box"]))["counts"]) == c["hits_27"]


def test_reference_single_case(orc, golden_dir):
    c = load_cases(golden_dir)["single"]
    t = orc.build(np.array(c["points"], np.float32), c["lo"], c["hi"])
    assert list(t["parent"]) == c["parent"] and list(t["primitive"]) == c["primitive"]
    assert t["lo"][0, 0] == F32(c["leaf_x"][0]) and t["hi"][0, 0] == F32(c["leaf_x"][1])
    comp = orc.compress(t)
    assert list(comp[0][:, 2]) == c["records_z"] and list(comp[0][:, 3]) == c["records_w"]
    assert list(orc.traverse(1, comp, np.array(c["spheres"], np.float32))["counts"]) == c["hits"]


def test_reference_validate_superset(orc, golden_dir):
    """lbvh_validate: hits >= float minimum-image brute force (test/lbvh_test.cu:459-502)."""
    c = load_cases(golden_dir)["validate"]
    pts = np.load(os.path.join(golden_dir, "validate_points.npy"))
    L = np.array(c["box"], np.float32)
    N = pts.shape[0]
    assert N == c["N"]
    t = orc.build(pts, -F32(0.5) * L, F32(0.5) * L)
    comp = orc.compress(t)
    sph = np.concatenate([pts, np.full((N, 1), c["rcut"], np.float32)], axis=1)
    hits = orc.traverse(N, comp, sph, orc.images27(L))["counts"]
    # the reference test's brute force in float with minimum image
    hi = F32(0.5) * L
    lo = -hi
    ref = np.zeros(N, np.int64)
    for i in range(N):
        dr = pts[i:] - pts[i]
        dr = np.where(dr >= hi, dr - L, np.where(dr < lo, dr + L, dr)).astype(np.float32)
        d2 = dr[:, 0] * dr[:, 0] + dr[:, 1] * dr[:, 1] + dr[:, 2] * dr[:, 2]
        m = d2 <= F32(c["rcut"]) * F32(c["rcut"])
        ref[i] += m.sum()
        ref[i + 1:] += m[1:]
    assert np.all(hits >= ref)
    assert hits.mean() < ref.mean() * 1.5


def test_golden_oracle_vectors_unchanged(orc, golden_dir):
    g = np.load(os.path.join(golden_dir, "oracle_n1000.npz"))
    t = orc.build(g["points"], g["lo"], g["hi"])
    for k in ("codes", "primitive", "parent", "left", "right"):
        assert np.array_equal(t[k], g[k]), k
    assert np.array_equal(t["lo"], g["node_lo"]) and np.array_equal(t["hi"], g["node_hi"])
    comp = orc.compress(t)
    assert np.array_equal(comp[0], g["records"])
    tr = orc.traverse(1000, comp, g["spheres"], g["images"], max_neigh=64)
    assert np.array_equal(tr["counts"], g["counts"]) and np.array_equal(tr["list"], g["nlist"])
    assert tr["visits"] == int(g["visits"])


# --------------------------------------------------------------------- traversal vs flat scan
def flat_scan(orc, N, comp, spheres, images):
    """Hits as {leaf with decoded box overlapping} for each image whose sphere overlaps the
    unquantized root box, images ascending then leaves in sorted order."""
    rec, rlo, rhi, bins = comp
    Nn = 2 * N - 1
    first_leaf = N - 1
    out = []
    lo3 = rec[:, 0].astype(np.uint32)
    hi3 = rec[:, 1].astype(np.uint32)
    boxes = []
    for n in range(first_leaf, Nn):
        blo, bhi = [], []
        for a, sh in enumerate((20, 10, 0)):
            blo.append(orc.approx("fadd_rd", rlo[a], orc.approx("fmul_rd", float((lo3[n] >> sh) & 0x3FF), bins[a])))
            bhi.append(orc.approx("fsub_ru", rhi[a], orc.approx("fmul_rd", float((hi3[n] >> sh) & 0x3FF), bins[a])))
        boxes.append((blo, bhi))

    def overlap(o, Rsq, blo, bhi):
        d = [orc.approx("fsub_rd", float(min(max(F32(o[a]), F32(blo[a])), F32(bhi[a]))), o[a]) for a in range(3)]
        d2 = orc.approx("fmaf_rd", d[0], d[0], orc.approx("fmaf_rd", d[1], d[1], orc.approx("fmul_rd", d[2], d[2])))
        return F32(d2) <= F32(Rsq)

    for s in spheres:
        Rsq = orc.approx("fmul_ru", s[3], s[3])
        hits = []
        for k, img in enumerate(images):
            o = [F32(s[a]) + F32(img[a]) for a in range(3)]
            if not overlap(o, Rsq, rlo, rhi):
                continue
            for li, (blo, bhi) in enumerate(boxes):
                if overlap(o, Rsq, blo, bhi):
                    hits.append((k, int(~rec[first_leaf + li, 2])))
        out.append(hits)
    return out


def test_rope_walk_equals_flat_leaf_scan(orc):
    rng = np.random.default_rng(8)
    N, L = 150, 6.0
    pts = rng.uniform(0, L, size=(N, 3)).astype(np.float32)
    t = orc.build(pts, [0, 0, 0], [L, L, L])
    comp = orc.compress(t)
    sph = np.concatenate([pts[:40], np.full((40, 1), 1.3, np.float32)], axis=1)
    images = orc.images27([L, L, L])
    tr = orc.traverse(N, comp, sph, images, triplets=True)
    flat = flat_scan(orc, N, comp, sph, images)
    offs = tr["offsets"]
    for q in range(40):
        got = list(zip(tr["image"][offs[q]:offs[q + 1]].tolist(), tr["prim"][offs[q]:offs[q + 1]].tolist()))
        assert got == flat[q], q


# --------------------------------------------------------------------- exact pairs
def test_exact_bruteforce_tiny_hand_case(orc):
    pts = np.array([[0.5, 0.5, 0.5], [1.5, 0.5, 0.5], [9.6, 0.5, 0.5]], np.float32)
    q = np.array([[0.5, 0.5, 0.5, 1.0]], np.float32)
    images = orc.images27([10, 10, 10], dtype=np.float64)
    r = orc.exact_bruteforce(pts, q, images)
    got = sorted(zip(r["k"].tolist(), r["j"].tolist()))
    # self (image 13), neighbor 1 at distance exactly 1.0 (image 13), point 2 through the image
    # t = (+10, 0, 0) (index 22): (0.5 + 10) - 9.6 = 0.9
    assert got == [(13, 0), (13, 1), (22, 2)]


def test_exact_celllist_equals_bruteforce(orc):
    from neighbor_amd.workloads import uniform_points, spheres_from_points
    L = 12.0
    pts = uniform_points(1200, L, seed=9)
    q = spheres_from_points(pts, 2.8)
    images = orc.images27([L, L, L], dtype=np.float64)
    bf = orc.exact_bruteforce(pts, q, images)
    cl = orc.exact_celllist(pts, q, [L, L, L], 2.8)
    assert np.array_equal(bf["counts"], cl["counts"])
    assert np.array_equal(bf["j"], cl["j"]) and np.array_equal(bf["k"], cl["k"])
    # the raw (reference) hits are a superset in practice
    t = orc.build(pts, [0, 0, 0], [L, L, L])
    raw = orc.traverse(1200, orc.compress(t), q, orc.images27([L, L, L]))["counts"]
    assert np.all(raw >= bf["counts"])


def test_exact_references_independent_of_thread_count(orc):
    """The OpenMP brute force and cell list give the same counts and rows on 1 and 4 threads."""
    from neighbor_amd.workloads import uniform_points, spheres_from_points
    L = 10.0
    pts = uniform_points(900, L, seed=4)
    q = spheres_from_points(pts, 2.1)
    images = orc.images27([L, L, L], dtype=np.float64)
    for f in (lambda n: orc.exact_bruteforce(pts, q, images, nthreads=n),
              lambda n: orc.exact_celllist(pts, q, [L, L, L], 2.1, nthreads=n)):
        one, four = f(1), f(4)
        assert one["total"] == four["total"] > 0
        assert np.array_equal(one["counts"], four["counts"])
        assert np.array_equal(one["j"], four["j"]) and np.array_equal(one["k"], four["k"])


# --------------------------------------------------------------------- BoundingSphere
def _ulp32(x):
    x = abs(F32(x))
    return float(np.nextafter(x, F32(np.inf))) - float(x)


def _exact(x):
    from fractions import Fraction
    return Fraction(float(x))


def test_sphere_double_ctor_closed_form_and_enclosure(orc):
    """BoundingSphere(double3 o, double r) (BoundingVolumes.h:135-147) pinned without retyping it:
    the origin is the largest float <= o on every axis (numpy nextafter); the radius covers r plus
    the largest per-axis displacement o - origin (exact rational arithmetic, 80-digit square
    roots); it is tight to a few ulps; and with float-representable o and r it reduces to the
    float ctor at next_up(r).

    The reference pads by the widest per-axis rounding interval, not by the Euclidean length of
    the displacement (up to sqrt(3) times more), so its float sphere need not contain the double
    sphere when r is a few ulps of o; this test records that too.  The exact queries of this
    library do not use this ctor (ExactSphereQueryOp pads by its own bound)."""
    from decimal import Decimal, getcontext
    getcontext().prec = 80
    rng = np.random.default_rng(21)
    cases = []
    for _ in range(3000):
        scale = 10.0 ** rng.integers(-3, 6)
        o = rng.uniform(-1, 1, 3) * scale
        r = float(rng.uniform(0.0, 3.0) * 10.0 ** rng.integers(-2, 3))
        cases.append((o, r))
    cases += [(np.array([1.0, 2.0, 3.0]), 2.0), (np.array([0.1, 0.2, 0.3]), 2.8), (np.array([-1e-3, 5e5, 7.25]), 0.0),
              (np.array([1e-320, -1e-320, 0.0]), 1.0), (np.array([105.56671905517578, 0.0, -0.0]), 2.8)]
    n_short = 0
    for o, r in cases:
        origin, rsq = orc.sphere_double(o, r)
        for a in range(3):
            assert np.float64(origin[a]) <= o[a] < np.float64(np.nextafter(origin[a], F32(np.inf))), (o, a, origin)
        # per-axis cover: sqrt(Rsq) >= r + max_a (o_a - origin_a) in exact arithmetic
        lhs = Decimal(float(rsq)).sqrt()
        dmax = max(_exact(o[a]) - _exact(origin[a]) for a in range(3))
        assert lhs >= Decimal(r) + Decimal(dmax.numerator) / Decimal(dmax.denominator), (o, r, rsq)
        d2 = sum((_exact(o[a]) - _exact(origin[a])) ** 2 for a in range(3))
        dist = (Decimal(d2.numerator) / Decimal(d2.denominator)).sqrt()
        n_short += lhs < Decimal(r) + dist
        # tight: the padding is a few ulps of o and of r
        maxulp = max(_ulp32(v) for v in origin)
        bound = (r * (1 + 2.0 ** -22) + 2.5 * maxulp + 4 * 2.0 ** -149) * (1 + 2.0 ** -20)
        assert float(lhs) <= bound, (o, r, rsq, bound)
    # the Euclidean enclosure fails only rarely (small r against large |o|)
    assert 0 < n_short < len(cases) // 10
    # float-representable inputs: the pad is one denormal step, R = next_up(r)
    for o, r in [((1.0, 2.0, 3.0), 2.0), ((0.5, -7.25, 1e3), 2.8), ((0.0, 0.0, 0.0), 1.0)]:
        r32 = F32(r)
        origin, rsq = orc.sphere_double(np.array(o, np.float64), float(r32))
        f_origin, f_rsq = orc.sphere_float(np.array(o, np.float32), float(np.nextafter(r32, F32(np.inf))))
        assert origin.tobytes() == f_origin.tobytes() and F32(rsq).tobytes() == F32(f_rsq).tobytes()


def test_sphere_overlap_conservative_and_tight(orc):
    """BoundingSphere::overlap (BoundingVolumes.h:163-171): never false when the exact squared
    distance from the origin to the box is <= Rsq; false once it exceeds Rsq by a relative 2^-19."""
    rng = np.random.default_rng(22)
    n_true = n_false = 0
    for _ in range(4000):
        origin = rng.uniform(-10, 10, 3).astype(np.float32)
        rsq = F32(rng.uniform(0.01, 25.0))
        c = origin + rng.standard_normal(3).astype(np.float32) * F32(np.sqrt(rsq))
        half = np.abs(rng.standard_normal(3)).astype(np.float32) * F32(rng.choice([0.0, 0.3, 2.0]))
        lo, hi = (c - half).astype(np.float32), (c + half).astype(np.float32)
        d2 = sum((max(min(_exact(origin[a]), _exact(hi[a])), _exact(lo[a])) - _exact(origin[a])) ** 2 for a in range(3))
        got = orc.sphere_overlap(origin, rsq, lo, hi)
        if d2 <= _exact(rsq):
            assert got, (origin, rsq, lo, hi)
            n_true += 1
        elif d2 > _exact(rsq) * (1 + _exact(2.0 ** -19)):
            assert not got, (origin, rsq, lo, hi)
            n_false += 1
    assert n_true > 500 and n_false > 500
    # a box touching the sphere exactly (distance == R, exact in floats) overlaps
    assert orc.sphere_overlap([0, 0, 0], 4.0, [2.0, -1, -1], [3.0, 1, 1])
    assert not orc.sphere_overlap([0, 0, 0], 4.0, [2.0001, -1, -1], [3.0, 1, 1])
