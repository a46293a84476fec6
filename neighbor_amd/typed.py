"""Neighbor lists of particles of several types (one LBVH per type, a radius per type pair).

The caller-side driver of the tree path the way a molecular-dynamics code uses it (HOOMD-blue's
tree neighbor list builds one LBVH per particle type and traverses each with the queries of
every type, with a cutoff per type pair such as (d_i + d_j)/2 + skin; the reference's HOOMD
ops are ``benchmark/lbvh_benchmark.cuh:25-92``): here with exact FP64 pairs and periodic
images.  Query type a against tree b uses the radius R[a][b]; the neighbor index emitted is the
particle's global index (each tree is set up and traversed with a transform map = the global
indices of its type, as the reference requires: ``LBVHTraverser.h:366-390``, the same transform
at setup and traversal).  Particle i's row holds its hits in tree 0, then tree 1, ...; rows are
laid out per query type, each type in its tree's Morton order.

Each (query type, tree type) pair has its own best traversal: few neighbors per query (small
radii, or sparse queries over a dense tree) favour one lane per query (float-rope records),
many favour the wave walk over leaf buckets.  ``tune`` times the candidates on the current
configuration and keeps the fastest per pair; every candidate gives the same pairs.
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _capi
from .api import LBVH, LBVHTraverser, exact_rows, exclusive_scan, filter_rows

#: (variant, leaves per bucket) candidates of the per-pair tuner
CANDIDATES: Tuple[Tuple[int, int], ...] = (
    (_capi.NB_VARIANT_FLOAT_ROPE_EXACT, 1),
    (_capi.NB_VARIANT_WAVE, 8),
    (_capi.NB_VARIANT_WAVE, 16),
    (_capi.NB_VARIANT_WAVE, 32),
)


class TypedNeighborList:
    def __init__(self, ntypes: int, sphere_radii: Optional[Sequence[float]] = None,
                 choice: Optional[Tuple[int, int]] = None, block: int = 256):
        """sphere_radii: if given, tree b is built over spheres of that radius (SphereInsertOp:
        leaf boxes p +- r) instead of points; the pairs are the same, the tree boxes looser.
        choice: (variant, bucket) used for every type pair until ``tune`` (default wave, 32)."""
        self.ntypes = ntypes
        self.sphere_radii = sphere_radii
        self.block = block
        self.trees = [LBVH() for _ in range(ntypes)]
        default = choice if choice is not None else (_capi.NB_VARIANT_WAVE, 32)
        self.choice: Dict[Tuple[int, int], Tuple[int, int]] = {
            (a, b): default for a in range(ntypes) for b in range(ntypes)}
        self._travs: Dict[Tuple[int, int, int], LBVHTraverser] = {}
        self._built = False
        self.index = None
        self._L = None
        self.neigh = None
        self.image = None
        self._side = []
        self._ref = None
        self._last = None
        self._last_key = None
        self._refits = 0
        self._wrapped = None

    # ---------------------------------------------------------------- configuration
    def set_types(self, types: np.ndarray, device="cuda") -> None:
        """types (N,) host ints in [0, ntypes): the per-type global index (the trees' transform
        maps) and the per-type buffers.  Call again when the types change."""
        types = np.asarray(types)
        self.N = int(types.shape[0])
        self.index, self._index_long, self.sub, self.count = [], [], [], []
        for b in range(self.ntypes):
            idx = np.flatnonzero(types == b).astype(np.int32)
            d_idx = torch.from_numpy(idx).to(device)
            self.index.append(d_idx)
            self._index_long.append(d_idx.long())
            self.sub.append(torch.empty((idx.size, 3), dtype=torch.float32, device=device))
            self.count.append(int(idx.size))
        # a type holding every particle needs neither the gather nor the transform map
        self._map = [None if self.count[b] == self.N else self.index[b] for b in range(self.ntypes)]
        # query spheres (x, y, z, R[a][b]) and per-pair counts of the type-a queries
        self._q = {(a, b): torch.empty((self.count[a], 4), dtype=torch.float32, device=device)
                   for a in range(self.ntypes) for b in range(self.ntypes)}
        self._per = {(a, b): torch.zeros(self.count[a], dtype=torch.int32, device=device)
                     for a in range(self.ntypes) for b in range(self.ntypes)}
        self._tot = [torch.empty(self.count[a], dtype=torch.int32, device=device) for a in range(self.ntypes)]
        self._fill = [torch.empty(max(self.count[a], 1), dtype=torch.int32, device=device) for a in range(self.ntypes)]
        self._starts = [torch.empty(self.count[a] + 1, dtype=torch.int64, device=device) for a in range(self.ntypes)]
        self._types = torch.from_numpy(types.astype(np.int32)).to(device)
        self._counts = torch.zeros(self.N, dtype=torch.int32, device=device)
        self._start = torch.zeros(self.N, dtype=torch.int64, device=device)
        self._R = {}
        self._built = False

    def _traverser(self, b: int, variant: int, bucket: int) -> LBVHTraverser:
        key = (b, variant, bucket)
        t = self._travs.get(key)
        if t is None:
            t = LBVHTraverser(variant=variant)
            t.set_bucket_size(bucket)
            self._travs[key] = t
            if self._built and self.count[b]:
                t.setup(self.trees[b], transform_map=self._map[b], block=self.block)
        return t

    # ---------------------------------------------------------------- build
    def build(self, points: torch.Tensor, box_L, types: Optional[np.ndarray] = None, refit: bool = False,
              stream=None) -> None:
        """points (N, 3) float32 on the device, or (N, 4) 16-byte records (x, y, z, w: HOOMD's
        postype layout, w not read), periodic cube [0, L)^3.  Records are read in place: each
        type's tree reads its particles' records through the type's index list
        (nb_lbvh_build_postype) and the exact tests read them with stride 4, so no per-type
        copy of the positions is made.  types: see set_types (required on the first call).
        refit: keep each tree's hierarchy and refit its boxes to the new positions (same types
        and box; nb_lbvh_refit_*) instead of building it."""
        if types is not None:
            self.set_types(types, device=points.device)
        if self.index is None:
            raise ValueError("set the particle types first (set_types or build(types=...))")
        if points.shape[0] != self.N:
            raise ValueError(f"{points.shape[0]} points for {self.N} typed particles")
        if points.dim() != 2 or points.shape[1] not in (3, 4):
            raise ValueError("points must be (N, 3) positions or (N, 4) postype records")
        self.points = points
        records = points.shape[1] == 4
        L = float(box_L)
        if refit and (types is not None or float(box_L) != self._L or not self._built):
            raise ValueError("refit needs a built list with the same types and box")
        if L != self._L:
            from .workloads import images27
            self.images = torch.from_numpy(images27(L, dtype=np.float64)).to(points.device)
            self._L = L
        lo, hi = [0.0, 0.0, 0.0], [L] * 3
        # traversers of the chosen (tree, variant, bucket) only; the others would hold stale records
        used = {(b, *self.choice[(a, b)]) for a in range(self.ntypes) for b in range(self.ntypes)}
        self._travs = {k: t for k, t in self._travs.items() if k in used}
        self._built = False
        for key in used:
            self._traverser(*key)
        # type b's gather, tree and traverser set-ups on stream b (the trees are independent)
        S = self._streams(stream)
        for b in range(1, self.ntypes):
            S[b].wait_stream(S[0])
        for b in range(self.ntypes):
            if self.count[b] == 0:
                continue
            with torch.cuda.stream(S[b]):
                r = None if self.sphere_radii is None else float(self.sphere_radii[b])
                if records:
                    # the type's records in place; the query spheres take their x, y, z (gathered
                    # straight into the preallocated query arrays: their 4th column is the radius,
                    # written by compute)
                    q0 = self._q[(b, 0)]
                    if self._map[b] is None:
                        q0.copy_(points)
                    else:
                        torch.index_select(points, 0, self._index_long[b], out=q0)
                    for a in range(1, self.ntypes):
                        self._q[(b, a)].copy_(q0)
                    if refit:
                        self.trees[b].refit_postype(points, index=self._map[b], radius=r, block=self.block, stream=S[b])
                    else:
                        self.trees[b].build_postype(points, lo, hi, index=self._map[b], radius=r, block=self.block,
                                                    stream=S[b])
                else:
                    if self._map[b] is None:
                        self.sub[b] = points
                    else:
                        torch.index_select(points, 0, self._index_long[b], out=self.sub[b])
                    for a in range(self.ntypes):
                        self._q[(b, a)][:, :3].copy_(self.sub[b])
                    if refit:
                        self.trees[b].refit(self.sub[b], radius=r, block=self.block, stream=S[b])
                    else:
                        self.trees[b].build(self.sub[b], lo, hi, radius=r, block=self.block, stream=S[b])
                for (tb, v, k), t in self._travs.items():
                    if tb == b:
                        t.setup(self.trees[b], transform_map=self._map[b], block=self.block, stream=S[b])
        for b in range(1, self.ntypes):
            S[0].wait_stream(S[b])
        self._built = True

    def _streams(self, stream=None):
        """[the caller's stream] + one side stream per further type (default priority)."""
        main = torch.cuda.current_stream() if stream is None else stream
        if len(self._side) != self.ntypes - 1:
            self._side = [torch.cuda.Stream(device=main.device) for _ in range(self.ntypes - 1)]
        return [main] + self._side

    def _query_radius(self, a: int, b: int, R, shift: Optional[torch.Tensor]) -> None:
        """Radius column of the (a, b) query spheres: R[a][b], or with shifts the smallest float
        >= (R[a][b] + shift_i) + max_j shift_j (FP64, never below 0), which bounds every pair
        cutoff (R[a][b] + shift_i) + shift_j of the type-b particles j."""
        col = self._q[(a, b)][:, 3]
        if shift is None:
            col.fill_(float(R[a][b]))
            return
        s_a = shift.index_select(0, self._index_long[a]).double()
        s_max = shift.index_select(0, self._index_long[b]).max().double()
        # the float32 cutoff, as the filter reads it
        r64 = torch.clamp((s_a + float(np.float32(R[a][b]))) + s_max, min=0.0)
        r32 = r64.float()
        r32 = torch.where(r32.double() < r64, torch.nextafter(r32, torch.full_like(r32, float("inf"))), r32)
        col.copy_(r32)

    def _radii(self, R) -> None:
        for a in range(self.ntypes):
            for b in range(self.ntypes):
                self._q[(a, b)][:, 3].fill_(float(R[a][b]))

    def _count(self, a: int, b: int, stream=None) -> torch.Tensor:
        variant, bucket = self.choice[(a, b)]
        t = self._traverser(b, variant, bucket)
        order = self.trees[a].device_pointers()["primitive"]
        return t.exact_count(self.trees[b], self._q[(a, b)], self.points, images=self.images,
                             transform_map=self._map[b], order=order, nneigh=self._per[(a, b)],
                             block=self.block, stream=stream)

    # ---------------------------------------------------------------- tuning
    def tune(self, R, candidates=CANDIDATES, reps: int = 3) -> Dict[Tuple[int, int], Tuple[int, int]]:
        """Time the count pass of every (query type, tree type) pair for each candidate
        (variant, bucket) on the built configuration and keep the fastest.  Synchronizes."""
        if not self._built:
            raise ValueError("build before tune")
        self._radii(R)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        timings = {}
        for a in range(self.ntypes):
            for b in range(self.ntypes):
                if self.count[a] == 0 or self.count[b] == 0:
                    continue
                best = None
                for cand in candidates:
                    self.choice[(a, b)] = cand
                    self._count(a, b)      # warm-up (sets the traverser up)
                    ts = []
                    for _ in range(reps):
                        e0.record()
                        self._count(a, b)
                        e1.record()
                        e1.synchronize()
                        ts.append(e0.elapsed_time(e1))
                    ms = float(np.median(ts))
                    timings[(a, b, cand)] = ms
                    if best is None or ms < best[0]:
                        best = (ms, cand)
                self.choice[(a, b)] = best[1]
        self.timings = timings
        return dict(self.choice)

    # ---------------------------------------------------------------- pairs
    def _total(self, S, prev) -> int:
        S[0].wait_event(prev[0])
        with torch.cuda.stream(S[0]):
            return int(prev[1].item())

    def _grow(self, S, total: int) -> None:
        cap = max(int(total * 1.1), 1)
        with torch.cuda.stream(S[0]):
            self.neigh = torch.empty(cap, dtype=torch.int32, device=self.points.device)
            self.image = torch.empty(cap, dtype=torch.uint8, device=self.points.device)

    def _fill_rows(self, S) -> None:
        nt = self.ntypes
        # the other streams wait for the pair buffers' allocation only, not for type 0's fills
        ready = S[0].record_event()
        for a in range(nt):
            if self.count[a] == 0:
                continue
            if a > 0:
                S[a].wait_event(ready)
                self.neigh.record_stream(S[a])
                self.image.record_stream(S[a])
            with torch.cuda.stream(S[a]):
                order = self.trees[a].device_pointers()["primitive"]
                sub_start = self._starts[a][:self.count[a]].clone()
                for b in range(nt):
                    if self.count[b] == 0:
                        continue
                    variant, bucket = self.choice[(a, b)]
                    exact_rows(self._traverser(b, variant, bucket), self.trees[b], self._q[(a, b)], self.points,
                               sub_start, self._per[(a, b)], self.neigh, self.image, images=self.images, order=order,
                               nneigh=self._fill[a], transform_map=self._map[b], block=self.block, stream=S[a])
                    sub_start += self._per[(a, b)]
        for a in range(1, nt):
            S[0].wait_stream(S[a])

    def compute(self, R, half: bool = False, shift: Optional[torch.Tensor] = None, stream=None) -> dict:
        """Exact pairs for the radii R[a][b].  Returns dict(counts (N,) int32, start (N,) int64,
        neigh, image, total, per_pair {(a, b): counts of the type-a queries in tree b}): particle
        i's pairs are neigh/image[start[i], start[i] + counts[i]).  Reads the pair total back to
        the host once, at the end (the pair buffers are kept across calls with 10% headroom; a
        configuration that outgrows them is filled again).

        shift (N,) float32 on the device, diameter shifting: the cutoff of pair (i, j) becomes
        (R[t_i][t_j] + shift_i) + shift_j in FP64 (shift = (d - 1) / 2 for HOOMD's shifted
        potentials).  Each type pair is then traversed with the per-query radius
        (R + shift_i) + max shift of the tree's type, rounded up to float, and the rows are
        filtered exactly (nb_rows_filter).  half: keep each pair once (j > i; a particle's own
        periodic images once per mirrored pair, never itself).  After either filter the rows
        keep their starts and leave gaps behind their (smaller) counts; per_pair holds the
        counts before filtering."""
        if not self._built:
            raise ValueError("build before compute")
        nt = self.ntypes
        S = self._streams(stream)
        for a in range(1, nt):
            S[a].wait_stream(S[0])
        # per query type a on stream a: counts against every tree, then its rows in its Morton
        # order (a permuted scan), placed after the rows of the earlier types
        prev = None     # (event, device total) of the previous type's rows
        for a in range(nt):
            if self.count[a] == 0:
                continue
            with torch.cuda.stream(S[a]):
                tot = self._tot[a]
                tot.zero_()
                for b in range(nt):
                    if self.count[b] == 0:
                        self._per[(a, b)].zero_()
                        continue
                    self._query_radius(a, b, R, shift)
                    tot += self._count(a, b, S[a])
                order = self.trees[a].device_pointers()["primitive"]
                st = exclusive_scan(tot, self._starts[a], perm=order, stream=S[a])
                if prev is not None:
                    S[a].wait_event(prev[0])
                    st += prev[1]
                prev = (S[a].record_event(), st[self.count[a]])
                self._counts.index_copy_(0, self._index_long[a], tot)
                self._start.index_copy_(0, self._index_long[a], st[:self.count[a]])
        if prev is None:
            total = 0
            if self.neigh is None:
                self._grow(S, 0)
        elif self.neigh is None:
            total = self._total(S, prev)
            self._grow(S, total)
            self._fill_rows(S)
        else:
            # rows past the kept capacity are dropped but counted: the one host read comes
            # after the fills, and only a grown configuration fills again
            self._fill_rows(S)
            total = self._total(S, prev)
            if total > self.neigh.numel():
                self._grow(S, total)
                self._fill_rows(S)
        filtered = bool(half) or shift is not None
        if filtered and total:
            with torch.cuda.stream(S[0]):
                radius = torch.tensor(R, dtype=torch.float32).to(self.points.device)
                pts3 = None
                if shift is not None:
                    # the radius filter reads float3 positions
                    pts3 = self.points if self.points.shape[1] == 3 else self.points[:, :3].contiguous()
                filter_rows(self._start, self._counts, self.neigh, self.image, counts_out=self._counts, half=half,
                            n_images=self.images.shape[0], points=pts3,
                            images=self.images, types=self._types, shift=shift, radius=radius, stream=S[0])
                kept = int(self._counts.to(torch.int64).sum().item())
        else:
            kept = total
        return dict(counts=self._counts, start=self._start, neigh=self.neigh[:total], image=self.image[:total],
                    total=kept, per_pair={k: v for k, v in self._per.items()}, filtered=filtered)

    # ---------------------------------------------------------------- MD steps
    def update(self, points: torch.Tensor, box_L, R, skin: float, half: bool = False,
               shift: Optional[torch.Tensor] = None, refit: bool = True, max_refits: int = 8,
               types: Optional[np.ndarray] = None, stream=None) -> dict:
        """One molecular-dynamics step of the list.  The list holds the exact pairs at the
        cutoffs R + skin of the configuration it was last computed from; while no particle has
        moved more than skin / 2 since then (nb_max_displacement2, minimum image; one host read),
        every pair within the cutoffs R of the current configuration is still in it, and it is
        returned as it is.  Otherwise it is recomputed: refit (the trees keep their hierarchy,
        at most max_refits times in a row, and only while no particle has crossed the box since
        the last build: a crossed leaf stays at its old place in the hierarchy and stretches its
        ancestors' boxes across the box) or rebuilt.  Returns compute()'s dict plus action
        ("kept", "refit", "built") and max_disp2.  Image indices of a kept list refer to the
        configuration it was computed from.  A call with other cutoffs, skin, half flag or shifts
        than the last computed list (a different shift tensor, or the same one changed in place)
        rebuilds."""
        from .api import max_displacement2
        L = float(box_L)
        Rs = [[float(R[a][b]) + skin for b in range(self.ntypes)] for a in range(self.ntypes)]
        # the kept list is valid only for the cutoffs, skin, half flag and shifts it was made with
        # (a shift tensor changed in place bumps its version counter)
        key = (tuple(tuple(r) for r in Rs), float(skin), bool(half),
               None if shift is None else (shift.data_ptr(), shift._version, tuple(shift.shape)))
        fresh = self._last is None or types is not None or L != self._L or self._ref is None \
            or self._ref.shape[0] != points.shape[0] or key != self._last_key
        disp2 = None
        wrapped = 0
        if not fresh:
            if self._wrapped is None:
                self._wrapped = torch.zeros((), dtype=torch.int32, device=points.device)
            # the displacement test reads float3 positions (a copy of the records' x, y, z)
            pts3 = points if points.shape[1] == 3 else points[:, :3].contiguous()
            disp2 = float(max_displacement2(pts3, self._ref, box=[L, L, L], wrapped=self._wrapped,
                                            stream=stream).item())
            if 4.0 * disp2 <= skin * skin:
                out = dict(self._last)
                out.update(action="kept", max_disp2=disp2)
                return out
            if refit and self._refits < max_refits:
                wrapped = int(self._wrapped.item())
        if fresh or not refit or self._refits >= max_refits or wrapped:
            self.build(points, L, types=types, stream=stream)
            self._refits = 0
            action = "built"
        else:
            self.build(points, L, refit=True, stream=stream)
            self._refits += 1
            action = "refit"
        res = self.compute(Rs, half=half, shift=shift, stream=stream)
        if self._ref is None or self._ref.shape[0] != points.shape[0]:
            self._ref = torch.empty((points.shape[0], 3), dtype=torch.float32, device=points.device)
        self._ref.copy_(points[:, :3])
        self._last = res
        self._last_key = key
        out = dict(res)
        out.update(action=action, max_disp2=disp2)
        return out
