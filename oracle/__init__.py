"""CPU oracle for the LBVH neighbor-search path -- TEST INFRASTRUCTURE ONLY.

This package wraps ``liboracle.so`` (built from ``oracle.c`` by ``oracle/Makefile``),
a plain-C restatement of the reference's HIP-path semantics (mphowardlab/neighbor,
``include/neighbor/kernels/LBVH.cuh``, ``kernels/LBVHTraverser.cuh``,
``ApproximateMath.h``, ``BoundingVolumes.h`` and the op headers).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
it, and only as the checker; the product (``neighbor_amd``) never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
SENTINEL = 0x7FFFFFFF

_lib = None


def compile_library() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            compile_library()
        L = ctypes.CDLL(_LIB_PATH)
        f, d, u32 = ctypes.c_float, ctypes.c_double, ctypes.c_uint32
        P = ctypes.c_void_p
        for name in ("next_down", "next_up", "frcp_rd", "frcp_ru"):
            fn = getattr(L, "orc_" + name)
            fn.restype, fn.argtypes = f, [f]
        for name in ("fadd_rd", "fadd_ru", "fsub_rd", "fsub_ru", "fmul_rd", "fmul_ru", "fdiv_rd", "fdiv_ru"):
            fn = getattr(L, "orc_" + name)
            fn.restype, fn.argtypes = f, [f, f]
        for name in ("fmaf_rd", "fmaf_ru"):
            fn = getattr(L, "orc_" + name)
            fn.restype, fn.argtypes = f, [f, f, f]
        for name in ("double2float_rd", "double2float_ru"):
            fn = getattr(L, "orc_" + name)
            fn.restype, fn.argtypes = f, [d]
        L.orc_sat_f2u.restype, L.orc_sat_f2u.argtypes = u32, [f]
        L.orc_morton3.restype, L.orc_morton3.argtypes = u32, [u32, u32, u32]
        L.orc_morton_codes.restype = None
        L.orc_morton_codes.argtypes = [P, u32, f, P, P, P]
        L.orc_sort_codes.restype = None
        L.orc_sort_codes.argtypes = [u32, P, P, P]
        L.orc_karras.restype = None
        L.orc_karras.argtypes = [u32, P, P, P, P]
        L.orc_bubble.restype = None
        L.orc_bubble.argtypes = [u32, P, P, P, P, P, f, P, P]
        L.orc_build.restype = None
        L.orc_build.argtypes = [P, u32, f, P, P, P, P, P, P, P, P, P]
        L.orc_compress.restype = None
        L.orc_compress.argtypes = [u32, P, P, P, P, P, P, P, P, P, P, P]
        L.orc_traverse.restype = ctypes.c_uint64
        L.orc_traverse.argtypes = [u32, P, P, P, P, P, u32, P, u32, P, P, u32, P, P, P, P]
        L.orc_exact_bruteforce.restype = ctypes.c_uint64
        L.orc_exact_bruteforce.argtypes = [P, u32, P, u32, P, u32, P, P, P, P]
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_sphere_double.restype = None
        L.orc_sphere_double.argtypes = [P, d, P, P]
        L.orc_sphere_float.restype = None
        L.orc_sphere_float.argtypes = [P, f, P, P]
        L.orc_sphere_overlap.restype = ctypes.c_int
        L.orc_sphere_overlap.argtypes = [P, f, P, P]
        L.orc_exact_celllist.restype = ctypes.c_uint64
        L.orc_exact_celllist.argtypes = [P, u32, P, u32, P, d, P, P, P, P, ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


# ---------------------------------------------------------------- scalars
def next_down(x):
    return lib().orc_next_down(float(x))


def next_up(x):
    return lib().orc_next_up(float(x))


def approx(name, *args):
    """Call one of the approx:: restatements by name, e.g. approx('fadd_rd', a, b)."""
    return getattr(lib(), "orc_" + name)(*[float(a) for a in args])


def sphere_double(o, r):
    """BoundingSphere(double3 o, double r): returns (origin float32[3], Rsq float32)."""
    o = np.ascontiguousarray(o, np.float64)
    origin = np.zeros(3, np.float32)
    rsq = np.zeros(1, np.float32)
    lib().orc_sphere_double(_p(o), float(r), _p(origin), _p(rsq))
    return origin, rsq[0]


def sphere_float(o, r):
    """BoundingSphere(float3 o, float r): returns (origin float32[3], Rsq float32)."""
    o = _f32(o)
    origin = np.zeros(3, np.float32)
    rsq = np.zeros(1, np.float32)
    lib().orc_sphere_float(_p(o), float(np.float32(r)), _p(origin), _p(rsq))
    return origin, rsq[0]


def sphere_overlap(origin, rsq, lo, hi):
    """BoundingSphere::overlap(BoundingBox(lo, hi))."""
    return bool(lib().orc_sphere_overlap(_p(_f32(origin)), float(np.float32(rsq)), _p(_f32(lo)), _p(_f32(hi))))


def sat_f2u(x):
    return lib().orc_sat_f2u(float(x))


def morton3(x, y, z):
    return lib().orc_morton3(int(x), int(y), int(z))


# ---------------------------------------------------------------- build
def morton_codes(xyz, lo, hi, sphere_r=-1.0):
    xyz = _f32(xyz, (-1, 3))
    N = xyz.shape[0]
    codes = np.zeros(N, np.uint32)
    lib().orc_morton_codes(_p(xyz), N, float(sphere_r), _p(_f32(lo)), _p(_f32(hi)), _p(codes))
    return codes


def sort_codes(codes):
    codes = np.ascontiguousarray(codes, np.uint32)
    N = codes.shape[0]
    sc = np.zeros(N, np.uint32)
    prim = np.zeros(N, np.uint32)
    lib().orc_sort_codes(N, _p(codes), _p(sc), _p(prim))
    return sc, prim


def karras(sorted_codes):
    sc = np.ascontiguousarray(sorted_codes, np.uint32)
    N = sc.shape[0]
    parent = np.zeros(max(2 * N - 1, 0), np.int32)
    left = np.zeros(max(N - 1, 0), np.int32)
    right = np.zeros(max(N - 1, 0), np.int32)
    lib().orc_karras(N, _p(sc), _p(parent), _p(left), _p(right))
    return parent, left, right


def build(xyz, lo, hi, sphere_r=-1.0):
    """Full LBVH build; returns a dict with the reference's arrays (LBVH.h:135-169)."""
    xyz = _f32(xyz, (-1, 3))
    N = xyz.shape[0]
    Nn = max(2 * N - 1, 0)
    out = dict(
        N=N,
        codes=np.zeros(N, np.uint32),
        primitive=np.zeros(N, np.uint32),
        parent=np.zeros(Nn, np.int32),
        left=np.zeros(max(N - 1, 0), np.int32),
        right=np.zeros(max(N - 1, 0), np.int32),
        lo=np.zeros((Nn, 3), np.float32),
        hi=np.zeros((Nn, 3), np.float32),
    )
    if N:
        lib().orc_build(_p(xyz), N, float(sphere_r), _p(_f32(lo)), _p(_f32(hi)), _p(out["codes"]),
                        _p(out["primitive"]), _p(out["parent"]), _p(out["left"]), _p(out["right"]),
                        _p(out["lo"]), _p(out["hi"]))
    return out


def compress(tree, transform_map=None):
    """Rope compression; returns (records[Nn,4] int32, root_lo, root_hi, bins)."""
    N = tree["N"]
    Nn = max(2 * N - 1, 0)
    rec = np.zeros((Nn, 4), np.int32)
    rlo = np.zeros(3, np.float32)
    rhi = np.zeros(3, np.float32)
    bins = np.zeros(3, np.float32)
    m = None if transform_map is None else np.ascontiguousarray(transform_map, np.uint32)
    if N:
        lib().orc_compress(N, _p(tree["parent"]), _p(tree["left"]), _p(tree["right"]), _p(tree["primitive"]),
                           _p(tree["lo"]), _p(tree["hi"]), _p(m), _p(rec), _p(rlo), _p(rhi), _p(bins))
    return rec, rlo, rhi, bins


def traverse(N, compressed, spheres, images=None, max_neigh=0, triplets=False):
    """Rope traversal.  Returns dict(counts, list|None, prim|None, image|None, visits, total)."""
    rec, rlo, rhi, bins = compressed
    spheres = _f32(spheres, (-1, 4))
    Nq = spheres.shape[0]
    if images is None:
        images = np.zeros((1, 3), np.float32)
    images = _f32(images, (-1, 3))
    Nimg = images.shape[0]
    counts = np.zeros(Nq, np.uint32)
    lst = np.zeros(Nq * max_neigh, np.uint32) if max_neigh > 0 else None
    visits = ctypes.c_uint64(0)
    out = dict(list=None, prim=None, image=None)
    L = lib()
    total = L.orc_traverse(N, _p(rec), _p(rlo), _p(rhi), _p(bins), _p(spheres), Nq, _p(images), Nimg,
                           _p(counts), _p(lst), max_neigh, None, None, None, ctypes.byref(visits))
    if triplets:
        offs = np.zeros(Nq + 1, np.uint64)
        np.cumsum(counts, out=offs[1:])
        prim = np.zeros(int(total), np.uint32)
        img = np.zeros(int(total), np.uint8)
        L.orc_traverse(N, _p(rec), _p(rlo), _p(rhi), _p(bins), _p(spheres), Nq, _p(images), Nimg,
                       _p(counts), None, 0, _p(offs), _p(prim), _p(img), None)
        out.update(prim=prim, image=img, offsets=offs)
    out.update(counts=counts, list=lst, visits=int(visits.value), total=int(total))
    return out


# ---------------------------------------------------------------- exact pairs
def exact_bruteforce(xyz, queries, images, pairs=True, nthreads=0):
    """FP64 brute force over explicit images (the reference test's CPU loop restated); nthreads > 0
    sets the OpenMP thread count."""
    lib().orc_set_threads(int(nthreads))
    xyz = _f32(xyz, (-1, 3))
    queries = _f32(queries, (-1, 4))
    images = np.ascontiguousarray(images, np.float64).reshape(-1, 3)
    Nq = queries.shape[0]
    counts = np.zeros(Nq, np.uint32)
    L = lib()
    total = L.orc_exact_bruteforce(_p(xyz), xyz.shape[0], _p(queries), Nq, _p(images), images.shape[0],
                                   _p(counts), None, None, None)
    out = dict(counts=counts, total=int(total))
    if pairs:
        offs = np.zeros(Nq + 1, np.uint64)
        np.cumsum(counts, out=offs[1:])
        j = np.zeros(int(total), np.uint32)
        k = np.zeros(int(total), np.uint8)
        L.orc_exact_bruteforce(_p(xyz), xyz.shape[0], _p(queries), Nq, _p(images), images.shape[0],
                               _p(counts), _p(offs), _p(j), _p(k))
        out.update(offsets=offs, j=j, k=k)
    return out


def exact_celllist(xyz, queries, box, rmax, pairs=True, nthreads=0):
    xyz = _f32(xyz, (-1, 3))
    queries = _f32(queries, (-1, 4))
    box = np.ascontiguousarray(box, np.float64).reshape(3)
    Nq = queries.shape[0]
    counts = np.zeros(Nq, np.uint32)
    L = lib()
    total = L.orc_exact_celllist(_p(xyz), xyz.shape[0], _p(queries), Nq, _p(box), float(rmax), _p(counts),
                                 None, None, None, int(nthreads))
    out = dict(counts=counts, total=int(total))
    if pairs:
        offs = np.zeros(Nq + 1, np.uint64)
        np.cumsum(counts, out=offs[1:])
        j = np.zeros(int(total), np.uint32)
        k = np.zeros(int(total), np.uint8)
        L.orc_exact_celllist(_p(xyz), xyz.shape[0], _p(queries), Nq, _p(box), float(rmax), _p(counts),
                             _p(offs), _p(j), _p(k), int(nthreads))
        out.update(offsets=offs, j=j, k=k)
    return out


def images27(box, dtype=np.float32):
    """27 periodic image vectors in the reference test's order (test/lbvh_test.cu:427-440):
    ix, iy, iz in -1..1 with iz fastest; the self image is index 13."""
    box = np.asarray(box, np.float64).reshape(3)
    out = []
    for ix in (-1, 0, 1):
        for iy in (-1, 0, 1):
            for iz in (-1, 0, 1):
                if dtype == np.float32:
                    # float products as the reference test forms them: L.x*((float)ix)
                    out.append([np.float32(box[0]) * np.float32(ix), np.float32(box[1]) * np.float32(iy),
                                np.float32(box[2]) * np.float32(iz)])
                else:
                    out.append([box[0] * ix, box[1] * iy, box[2] * iz])
    return np.asarray(out, dtype=dtype)


# ---------------------------------------------------------------- row filters (typed lists)
def pair_distance2(xi, xj, t):
    """d2 of the exact pair test in IEEE double: d = ((double)x_i + t) - (double)x_j per axis,
    d2 = (dx*dx + dy*dy) + dz*dz (rows of xi, xj, t broadcast together)."""
    d = (np.asarray(xi, np.float32).astype(np.float64) + np.asarray(t, np.float64)) \
        - np.asarray(xj, np.float32).astype(np.float64)
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


def half_mask(i, j, k, n_images=27):
    """Half-list rule: j > i, or j == i at an image k of the first half of the mirrored list."""
    i, j, k = (np.asarray(v, np.int64) for v in (i, j, k))
    return (j > i) | ((j == i) & (2 * k + 1 < n_images))


def shifted_radius(radius, types, shift, i, j):
    """Pair cutoff r_ij = ((double)radius[t_i][t_j] + shift_i) + shift_j in IEEE double."""
    radius = np.asarray(radius, np.float32)
    types = np.asarray(types, np.int64)
    shift = np.asarray(shift, np.float32).astype(np.float64)
    return (radius[types[i], types[j]].astype(np.float64) + shift[i]) + shift[j]


def shifted_pairs(xyz, types, shift, radius, box):
    """All (i, j, k) of particles xyz (periodic cube box, 27 images) with r_ij >= 0 and
    d2 <= r_ij^2 (pair_distance2, shifted_radius).  Candidates: the exact cell list with the
    per-particle radius max_b ((radius[t_i][b] + shift_i) + max shift of type b), rounded up
    to float, which bounds every r_ij; then the FP64 test.  Returns sorted (i, j, k) int64 rows."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    types = np.asarray(types, np.int64)
    shift32 = np.asarray(shift, np.float32)
    s = shift32.astype(np.float64)
    radius = np.asarray(radius, np.float32)
    nt = radius.shape[0]
    smax = np.array([s[types == b].max() if np.any(types == b) else -np.inf for b in range(nt)])
    rc = np.full(xyz.shape[0], 0.0)
    for b in range(nt):
        rc = np.maximum(rc, (radius[types, b].astype(np.float64) + s) + smax[b])
    r32 = rc.astype(np.float32)
    low = r32.astype(np.float64) < rc
    r32[low] = np.nextafter(r32[low], np.float32(np.inf))
    q = np.concatenate([xyz, r32[:, None]], 1)
    ref = exact_celllist(xyz, q, box, float(r32.max()))
    i = np.repeat(np.arange(xyz.shape[0]), ref["counts"].astype(np.int64))
    j = ref["j"].astype(np.int64)
    k = ref["k"].astype(np.int64)
    img = images27(box, np.float64)
    r = shifted_radius(radius, types, shift32, i, j)
    keep = (r >= 0) & (pair_distance2(xyz[i], xyz[j], img[k]) <= r * r)
    out = np.stack([i[keep], j[keep], k[keep]], 1)
    return out[np.lexsort(out.T[::-1])]
