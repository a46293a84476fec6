"""Deterministic synthetic inputs for the benchmark configurations.

Counter-based generator: value i of stream s is splitmix64(seed, s, i), so every
config is reproducible on any host from (seed, sizes) alone.  Coordinates are
24-bit fractions times the box length, formed in float32.
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n counter-based 64-bit values of stream `stream`."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0x2545F4914F6CDD1D) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.arange(offset + 1, offset + n + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def unit24(seed: int, stream: int, n: int) -> np.ndarray:
    """n float32 values in [0, 1) with 24 random bits."""
    u = (splitmix64(seed, stream, n) >> np.uint64(40)).astype(np.float32)
    return u * np.float32(2.0 ** -24)


def box_length(N: int, density: float) -> float:
    return float(np.float32((N / density) ** (1.0 / 3.0)))


def uniform_points(N: int, L: float, seed: int = 42) -> np.ndarray:
    """N float32 points uniform in the periodic cube [0, L)^3."""
    Lf = np.float32(L)
    pts = np.empty((N, 3), np.float32)
    for a in range(3):
        x = unit24(seed, a, N) * Lf
        x[x >= Lf] = np.float32(0.0)
        pts[:, a] = x
    return pts


def spheres_from_points(points: np.ndarray, r: float) -> np.ndarray:
    s = np.empty((points.shape[0], 4), np.float32)
    s[:, :3] = points
    s[:, 3] = np.float32(r)
    return s


def images27(L, dtype=np.float32) -> np.ndarray:
    """27 periodic image vectors, ix, iy, iz in -1..1 with iz fastest (self image at 13)."""
    L = np.broadcast_to(np.asarray(L, np.float64), (3,))
    out = np.empty((27, 3), dtype)
    i = 0
    for ix in (-1, 0, 1):
        for iy in (-1, 0, 1):
            for iz in (-1, 0, 1):
                if dtype == np.float32:
                    out[i] = [np.float32(L[0]) * np.float32(ix), np.float32(L[1]) * np.float32(iy),
                              np.float32(L[2]) * np.float32(iz)]
                else:
                    out[i] = [L[0] * ix, L[1] * iy, L[2] * iz]
                i += 1
    return out


def expected_neighbors(r: float, density: float) -> float:
    """Mean neighbors per particle including self, 4/3 pi r^3 rho."""
    return 4.0 / 3.0 * math.pi * r ** 3 * density


# configuration S: N=10,000 at density 1.0, r = 2.8
SMALL = dict(name="uniform N=10000 rho=1.0 r=2.8 27 images", N=10_000, density=1.0, r=2.8)
# configuration P: N=1,000,000 at density 0.85, r = 2.8
LARGE = dict(name="uniform N=1000000 rho=0.85 r=2.8 27 images", N=1_000_000, density=0.85, r=2.8)


def make_uniform(cfg: dict, seed: int = 42):
    L = box_length(cfg["N"], cfg["density"])
    pts = uniform_points(cfg["N"], L, seed)
    return pts, L


def clustered_points(N: int, density: float = 0.85, n_clusters: int = 1000, sigma: float = 5.0,
                     background: float = 0.2, seed: int = 42):
    """N float32 points in the periodic cube of mean density `density`: a fraction `background`
    uniform, the rest in `n_clusters` Gaussian clusters (std `sigma`) at uniform centers.
    Returns (points, L).  Gaussians by Box-Muller on counter-based uniforms, wrapped into [0, L)."""
    L = box_length(N, density)
    Lf = np.float32(L)
    n_bg = int(round(background * N))
    n_cl = N - n_bg
    pts = np.empty((N, 3), np.float32)
    for a in range(3):
        x = unit24(seed, a, n_bg) * Lf
        x[x >= Lf] = np.float32(0.0)
        pts[:n_bg, a] = x
    if n_cl == 0:
        return pts, L
    centers = np.stack([unit24(seed, 10 + a, n_clusters).astype(np.float64) * L for a in range(3)], axis=1)
    which = (splitmix64(seed, 20, n_cl) % np.uint64(n_clusters)).astype(np.int64)
    for a in range(3):
        u1 = ((splitmix64(seed, 30 + 2 * a, n_cl) >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
        u2 = (splitmix64(seed, 31 + 2 * a, n_cl) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
        x = np.mod(centers[which, a] + sigma * g, L).astype(np.float32)
        x[(x >= Lf) | (x < 0)] = np.float32(0.0)
        pts[n_bg:, a] = x
    return pts, L


# configuration X: N=4,000,000 clustered (1000 Gaussian clusters, sigma 5, 20% uniform), rho=0.85, r=2.8
CLUSTERED = dict(name="clustered N=4000000 rho=0.85 1000 clusters sigma=5 20% uniform r=2.8 27 images",
                 N=4_000_000, density=0.85, n_clusters=1000, sigma=5.0, background=0.2, r=2.8)


def batched_boxes(nbox: int = 4096, nmin: int = 256, nmax: int = 65536, density: float = 0.85, seed: int = 42,
                  sizes=None):
    """Independent periodic cubes: box b holds N_b points (log-uniform in [nmin, nmax] unless `sizes`
    is given) uniform in [0, L_b)^3 with L_b = (N_b / density)^(1/3).  Returns (points (sum N_b, 3)
    float32, offsets (nbox + 1) int64, box_lo (nbox, 3) float32, box_hi (nbox, 3) float32)."""
    if sizes is None:
        u = (splitmix64(seed, 100, nbox) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        lo, hi = math.log(nmin), math.log(nmax + 1)
        sizes = np.minimum(np.floor(np.exp(lo + u * (hi - lo))), nmax).astype(np.int64)
    sizes = np.asarray(sizes, np.int64)
    nbox = sizes.shape[0]
    offsets = np.zeros(nbox + 1, np.int64)
    np.cumsum(sizes, out=offsets[1:])
    points = np.empty((int(offsets[-1]), 3), np.float32)
    box_lo = np.zeros((nbox, 3), np.float32)
    box_hi = np.empty((nbox, 3), np.float32)
    for b in range(nbox):
        n = int(sizes[b])
        L = box_length(max(n, 1), density)
        box_hi[b] = L
        if n:
            points[offsets[b]:offsets[b + 1]] = uniform_points(n, L, seed=seed * 1_000_003 + b)
    return points, offsets, box_lo, box_hi


# configuration B: 4096 periodic boxes, N log-uniform in [256, 65536], rho=0.85, r=2.8
BATCHED = dict(name="batched 4096 periodic boxes N log-uniform [256, 65536] rho=0.85 r=2.8 27 images",
               nbox=4096, nmin=256, nmax=65536, density=0.85, r=2.8)


def colloid_mixture(N: int = 2_000_000, frac_large: float = 0.01, d_large: float = 10.0, d_small: float = 1.0,
                    phi: float = 0.3, seed: int = 42):
    """Size-asymmetric sphere mixture in a periodic cube at total volume fraction phi: large spheres
    placed by random sequential addition without overlap, then small spheres uniform outside every
    large sphere (small-small overlaps allowed: the small fraction is dilute).  Returns
    (points (N, 3) float32, types (N,) int32: 0 small / 1 large, L)."""
    n_l = int(round(frac_large * N))
    n_s = N - n_l
    vol = n_l * math.pi / 6.0 * d_large ** 3 + n_s * math.pi / 6.0 * d_small ** 3
    L = float(np.float32((vol / phi) ** (1.0 / 3.0)))
    rng = np.random.default_rng(seed)
    # large: random sequential addition on a grid of cells of side >= d_large; candidates come
    # in batches, checked against the placed spheres at once, then against each other in order
    nc = max(1, int(L // d_large))
    cell = L / nc
    kcap = 8    # centers at least d_large apart: at most the 8 corners of a cell of side < 2 d / sqrt(3)
    table = np.full((nc, nc, nc, kcap), -1, np.int64)
    fill = np.zeros((nc, nc, nc), np.int64)
    large = np.empty((n_l, 3), np.float64)
    placed = 0
    dl2 = d_large * d_large
    offs = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]
    while placed < n_l:
        cand = rng.uniform(0.0, L, size=(2048, 3))
        ci = (cand // cell).astype(np.int64) % nc
        bad = np.zeros(cand.shape[0], bool)
        for dx, dy, dz in offs:
            ids = table[(ci[:, 0] + dx) % nc, (ci[:, 1] + dy) % nc, (ci[:, 2] + dz) % nc]
            for s_ in range(kcap):
                k = ids[:, s_]
                has = k >= 0
                if not has.any():
                    continue
                d = cand[has] - large[k[has]]
                d -= L * np.round(d / L)
                bad[np.flatnonzero(has)[(d * d).sum(1) < dl2]] = True
        surv, scell = cand[~bad], ci[~bad]
        acc = []
        for i in range(surv.shape[0]):
            if placed + len(acc) == n_l:
                break
            if acc:
                d = surv[acc] - surv[i]
                d -= L * np.round(d / L)
                if ((d * d).sum(1) < dl2).any():
                    continue
            acc.append(i)
        for i in acc:
            cc = scell[i]
            large[placed] = surv[i]
            table[cc[0], cc[1], cc[2], fill[cc[0], cc[1], cc[2]]] = placed
            fill[cc[0], cc[1], cc[2]] += 1
            placed += 1
    # small: uniform, rejected inside any large sphere (center distance < (d_large + d_small) / 2)
    from scipy.spatial import cKDTree
    tree = cKDTree(large, boxsize=L) if n_l else None
    rmin = 0.5 * (d_large + d_small)
    small = np.empty((0, 3), np.float64)
    while small.shape[0] < n_s:
        cand = rng.uniform(0.0, L, size=(min(2 * (n_s - small.shape[0]) + 1024, 1 << 22), 3))
        cand[cand >= L] = 0.0
        if tree is not None:
            dist, _ = tree.query(cand, k=1, distance_upper_bound=rmin, workers=-1)
            cand = cand[~(dist < rmin)]
        small = np.concatenate([small, cand])[:n_s]
    pts = np.concatenate([small, large]).astype(np.float32)
    pts[pts >= np.float32(L)] = np.float32(0.0)
    types = np.concatenate([np.zeros(n_s, np.int32), np.ones(n_l, np.int32)])
    return pts, types, L


# configuration C: 2e6 spheres, 1% d=10 + 99% d=1, phi=0.3, per-pair radius (d_i + d_j)/2 + 0.4
COLLOID = dict(name="colloid N=2000000 1% d=10 99% d=1 phi=0.3 skin=0.4 27 images", N=2_000_000, frac_large=0.01,
               d_large=10.0, d_small=1.0, phi=0.3, skin=0.4)
