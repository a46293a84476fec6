"""Batched boxes over the C ABI: one forest build, one persistent traversal, one graph.

Points of all boxes are concatenated (box b holds rows offsets[b]:offsets[b+1]); each box
has its own lower/upper corner (its periodic cell).  Queries are the points themselves,
one sphere (x, y, z, r) per point, each against its own box.  Neighbor indices and the
per-query outputs use the global (concatenated) point index.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from ._capi import NB_BATCH_LANE, NB_BATCH_WAVE, P, call, lib  # noqa: F401
from .api import _ptr, _stream, exclusive_scan

U32 = ctypes.c_uint32


def _check(t: torch.Tensor, dtype, name: str, cols: Optional[int] = None) -> torch.Tensor:
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous CUDA tensor of {dtype}")
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise ValueError(f"{name} must have shape (n, {cols})")
    return t


class BatchRows:
    """Buffers of the batched single pass: row capacity, start (N + 1) and count per point, the
    pair arrays (neighbor, image) and the spill counter."""

    def __init__(self, N: int, capacity: int, device="cuda"):
        self.cap = torch.empty(N, dtype=torch.int32, device=device)
        self.start = torch.empty(N + 1, dtype=torch.int64, device=device)
        self.count = torch.empty(N, dtype=torch.int32, device=device)
        self.neigh = torch.empty(max(capacity, 1), dtype=torch.int32, device=device)
        self.image = torch.empty(max(capacity, 1), dtype=torch.uint8, device=device)
        self.spill = torch.zeros(2, dtype=torch.int64, device=device)


class LBVHBatch:
    """Forest of per-box LBVHs (see include/neighbor_amd.h, nb_batch_*)."""

    def __init__(self):
        h = ctypes.c_void_p()
        call("nb_batch_create", ctypes.byref(h))
        self._h = h
        self._destroy = lib().nb_batch_destroy
        self._keep = ()
        self.N = 0
        self.nbox = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def build(self, points: torch.Tensor, offsets: torch.Tensor, box_lo: torch.Tensor, box_hi: torch.Tensor,
              stream=None) -> None:
        """points (N, 3) float32; offsets (nbox + 1) int32 (uint32 values); box_lo/box_hi (nbox, 3) float32."""
        _check(points, torch.float32, "points", 3)
        _check(offsets, torch.int32, "offsets")
        _check(box_lo, torch.float32, "box_lo", 3)
        _check(box_hi, torch.float32, "box_hi", 3)
        nbox = offsets.shape[0] - 1
        if box_lo.shape[0] != nbox or box_hi.shape[0] != nbox:
            raise ValueError("box_lo/box_hi need one row per box")
        self._keep = (points, offsets, box_lo, box_hi)     # the forest refers to them
        self.N, self.nbox = points.shape[0], nbox
        call("nb_batch_build", self._h, _ptr(points), self.N, _ptr(offsets), _ptr(box_lo), _ptr(box_hi), nbox,
             _stream(stream))

    def error_word(self) -> int:
        """The forest build's error word (0 after a good build; synchronizes)."""
        from .api import _read_error_word
        return _read_error_word("nb_batch_error_word", self._h)

    def check(self, stream=None) -> None:
        """Synchronize the stream and raise NeighborError if the last forest build gave up a
        bounded wait or list."""
        call("nb_batch_check", self._h, _stream(stream))

    def set_params(self, block: int = 256, bucket: int = 16, variant: int = NB_BATCH_WAVE, grid: int = 0) -> None:
        call("nb_batch_set_params", self._h, block, bucket, variant, grid)

    def get_params(self) -> dict:
        b, k, g = U32(), U32(), U32()
        v = ctypes.c_int()
        call("nb_batch_get_params", self._h, ctypes.byref(b), ctypes.byref(k), ctypes.byref(v), ctypes.byref(g))
        return {"block": b.value, "bucket": k.value, "variant": v.value, "grid": g.value}

    def setup(self, stream=None) -> None:
        call("nb_batch_setup", self._h, _stream(stream))

    def count(self, spheres: torch.Tensor, periodic: bool = True, nneigh: Optional[torch.Tensor] = None,
              stream=None) -> torch.Tensor:
        _check(spheres, torch.float32, "spheres", 4)
        if nneigh is None:
            nneigh = torch.empty(self.N, dtype=torch.int32, device=spheres.device)
        call("nb_batch_count", self._h, _ptr(spheres), int(periodic), _ptr(nneigh), _stream(stream))
        return nneigh

    def pairs(self, spheres: torch.Tensor, offsets: torch.Tensor, periodic: bool = True, total: Optional[int] = None,
              neigh: Optional[torch.Tensor] = None, image: Optional[torch.Tensor] = None, with_image: bool = True,
              stream=None):
        _check(spheres, torch.float32, "spheres", 4)
        if total is None:
            total = int(offsets[-1].item())
        if neigh is None:
            neigh = torch.empty(max(total, 1), dtype=torch.int32, device=spheres.device)
        if image is None and with_image:
            image = torch.empty(max(total, 1), dtype=torch.uint8, device=spheres.device)
        call("nb_batch_pairs", self._h, _ptr(spheres), int(periodic), _ptr(offsets), _ptr(neigh), _ptr(image), None,
             neigh.numel(), _stream(stream))
        return neigh[:total], (image[:total] if image is not None else None)

    def rows(self, spheres: torch.Tensor, start: torch.Tensor, limit: torch.Tensor, neigh: torch.Tensor,
             image: Optional[torch.Tensor] = None, nneigh: Optional[torch.Tensor] = None, periodic: bool = True,
             spill: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
        """Pairs into rows [start[q], start[q] + limit[q]) (see nb_batch_rows); returns the counts.
        spill (2 int64: {first free position past all rows, 0}): overflowed rows are moved there
        and walked again inside the launch (start/limit rewritten; spill[1] counts them)."""
        _check(spheres, torch.float32, "spheres", 4)
        if nneigh is None:
            nneigh = torch.empty(self.N, dtype=torch.int32, device=spheres.device)
        call("nb_batch_rows", self._h, _ptr(spheres), int(periodic), _ptr(start), _ptr(limit), _ptr(neigh), _ptr(image),
             _ptr(nneigh), neigh.numel(), _ptr(spill), _stream(stream))
        return nneigh

    def capacity(self, spheres: torch.Tensor, z: float = 4.0, pad: int = 8, cap: Optional[torch.Tensor] = None,
                 stream=None) -> torch.Tensor:
        """Row capacities from each box's density (nb_batch_capacity)."""
        _check(spheres, torch.float32, "spheres", 4)
        if cap is None:
            cap = torch.empty(self.N, dtype=torch.int32, device=spheres.device)
        call("nb_batch_capacity", self._h, _ptr(spheres), float(z), int(pad), _ptr(cap), _stream(stream))
        return cap

    def single_pass(self, spheres: torch.Tensor, buffers: "BatchRows", z: float = 4.0, pad: int = 8,
                    periodic: bool = True, stream=None) -> "BatchRows":
        """Exact pairs of the built forest in one fill pass, without a count pass: rows sized by
        capacity(), laid out in the forest's Morton order, overflowed rows moved and walked again
        inside the fill.  No host synchronization: afterwards buffers.spill = {used positions
        (every row complete iff at most buffers.neigh.numel()), moved rows}."""
        self.capacity(spheres, z, pad, buffers.cap, stream)
        exclusive_scan(buffers.cap, buffers.start, stream, perm=self.primitive_address())
        buffers.spill[0:1].copy_(buffers.start[self.N:self.N + 1])
        buffers.spill[1:2].zero_()
        self.rows(spheres, buffers.start, buffers.cap, buffers.neigh, buffers.image, buffers.count, periodic,
                  spill=buffers.spill, stream=stream)
        return buffers

    def primitive_address(self) -> int:
        """Device address of the forest's primitive order (sorted position -> point), N u32."""
        ptrs = [ctypes.c_void_p() for _ in range(9)]
        N, nbox, nodes = U32(), U32(), U32()
        call("nb_batch_arrays", self._h, *[ctypes.byref(p) for p in ptrs], ctypes.byref(N), ctypes.byref(nbox),
             ctypes.byref(nodes))
        return ptrs[1].value or 0

    def arrays(self) -> dict:
        """Copies of the forest arrays (synchronizes)."""
        ptrs = [ctypes.c_void_p() for _ in range(9)]
        N, nbox, nodes = U32(), U32(), U32()
        call("nb_batch_arrays", self._h, *[ctypes.byref(p) for p in ptrs], ctypes.byref(N), ctypes.byref(nbox),
             ctypes.byref(nodes))
        torch.cuda.synchronize()
        from ._capi import memcpy_d2d
        spec = [("roots", torch.int32, nbox.value), ("primitive", torch.int32, N.value), ("keys", torch.int64, N.value),
                ("parent", torch.int32, nodes.value), ("left", torch.int32, max(N.value - 1, 0)),
                ("right", torch.int32, max(N.value - 1, 0)), ("rope", torch.int32, nodes.value),
                ("lo", torch.float32, 3 * nodes.value), ("hi", torch.float32, 3 * nodes.value)]
        out = {}
        for (name, dt, n), p in zip(spec, ptrs):
            t = torch.empty(n, dtype=dt, device="cuda")
            if n and p.value:
                memcpy_d2d(t.data_ptr(), p.value, t.numel() * t.element_size(), 0)
            out[name] = t
        torch.cuda.synchronize()
        out["lo"] = out["lo"].view(-1, 3)
        out["hi"] = out["hi"].view(-1, 3)
        return out

    def tune(self, points, offsets, box_lo, box_hi, spheres, periodic: bool = True, stream=None) -> dict:
        """Autotune block size, bucket size and variant on this problem (leaves it built and set up)."""
        self._keep = (points, offsets, box_lo, box_hi)
        self.N, self.nbox = points.shape[0], offsets.shape[0] - 1
        scratch = torch.empty(max(self.N, 1), dtype=torch.int32, device=points.device)
        b, k = U32(), U32()
        v = ctypes.c_int()
        tb, tc = ctypes.c_float(), ctypes.c_float()
        call("nb_batch_tune", self._h, _ptr(points), self.N, _ptr(offsets), _ptr(box_lo), _ptr(box_hi), self.nbox,
             _ptr(spheres), int(periodic), _ptr(scratch), ctypes.byref(b), ctypes.byref(k), ctypes.byref(v),
             ctypes.byref(tb), ctypes.byref(tc), _stream(stream))
        return {"block": b.value, "bucket": k.value, "variant": v.value, "build_ms": tb.value, "count_ms": tc.value}


class BatchPlan:
    """One HIP graph replaying build + setup + count + scan + pair fill of a batch.  Point q's pairs
    are neigh/image[offsets[q] : offsets[q] + nneigh[q]] (rows in the forest's Morton order)."""

    def __init__(self, batch: LBVHBatch, points, offsets, box_lo, box_hi, spheres, capacity: int,
                 periodic: bool = True, with_image: bool = True):
        N = points.shape[0]
        dev = points.device
        self.batch = batch
        batch._keep = (points, offsets, box_lo, box_hi)
        batch.N, batch.nbox = N, offsets.shape[0] - 1
        self._inputs = (points, offsets, box_lo, box_hi, spheres)
        self.nneigh = torch.empty(max(N, 1), dtype=torch.int32, device=dev)
        self.offsets = torch.empty(N + 1, dtype=torch.int64, device=dev)
        self.scratch = torch.empty(max(int(lib().nb_scan_scratch_bytes(N)), 256), dtype=torch.uint8, device=dev)
        self.neigh = torch.empty(max(capacity, 1), dtype=torch.int32, device=dev)
        self.image = torch.empty(max(capacity, 1), dtype=torch.uint8, device=dev) if with_image else None
        h = ctypes.c_void_p()
        call("nb_batch_plan_create", ctypes.byref(h), batch.handle, _ptr(points), N, _ptr(offsets), _ptr(box_lo),
             _ptr(box_hi), batch.nbox, _ptr(spheres), int(periodic), _ptr(self.nneigh), _ptr(self.offsets),
             _ptr(self.scratch), _ptr(self.neigh), _ptr(self.image), self.neigh.numel())
        self._h = h
        self._destroy = lib().nb_batch_plan_destroy

    def launch(self, stream=None) -> None:
        call("nb_batch_plan_launch", self._h, _stream(stream))

    def __del__(self):
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self._h = None
