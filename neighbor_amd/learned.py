"""Learned neighbor capacity and the single-pass exact pair fill (over the C ABI).

Pipeline (``learned_exact_pairs``):
  1. cap[q] = capacity predicted by the MLP from Morton-cell occupancy features;
  2. rows start at the exclusive scan of cap;
  3. one exact traversal writes every query's (neighbor, image) pairs into its row and
     records the true count;
  4. rows whose count exceeds their capacity are relocated after the padded rows and
     redone by the exact traversal (the count-then-fill fallback), so the pair sets are
     exactly those of the two-pass path whatever the network predicts.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from . import _capi
from ._capi import P, call, lib
from .api import LBVH, LBVHTraverser, _exact_points, _ptr, _stream, _vec3, exclusive_scan

N_PARAMS = 1377
N_FEATURES = 8


class CapacityPredictor:
    """8-32-32-1 FP32 MLP predicting log2(1 + neighbor count) per query."""

    def __init__(self, seed: int = 1):
        h = ctypes.c_void_p()
        call("nb_capacity_create", ctypes.byref(h), int(seed))
        self._h = h
        self._destroy = lib().nb_capacity_destroy
        self.n_features_rows = 0
        self._feat_ptr = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self._h = None

    def parameters(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, np.float32)
        n = ctypes.c_uint32()
        call("nb_capacity_parameters", self._h, out.ctypes.data_as(P), None, ctypes.byref(n))
        return out

    def set_parameters(self, params) -> None:
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == N_PARAMS
        n = ctypes.c_uint32()
        call("nb_capacity_parameters", self._h, None, p.ctypes.data_as(P), ctypes.byref(n))

    def features(self, lbvh: LBVH, lo, hi, r: float, stream=None) -> int:
        """Compute features of lbvh's primitives (sorted order); returns the device address (N x 8 float)."""
        fp = ctypes.c_void_p()
        call("nb_capacity_features", self._h, lbvh.handle, _vec3(lo), _vec3(hi), float(r), ctypes.byref(fp),
             _stream(stream))
        self.n_features_rows = lbvh.getN()
        self._feat_ptr = fp.value or 0
        return self._feat_ptr

    def features_tensor(self) -> torch.Tensor:
        t = torch.empty((self.n_features_rows, N_FEATURES), dtype=torch.float32, device="cuda")
        if self._feat_ptr and t.numel():
            _capi.memcpy_d2d(t.data_ptr(), self._feat_ptr, t.numel() * 4, _stream(None).value or 0)
        return t

    def train(self, lbvh: LBVH, counts: torch.Tensor, steps: int = 300, batch: int = 16384, lr: float = 3e-3,
              seed: int = 0, stream=None) -> float:
        loss = ctypes.c_float()
        call("nb_capacity_train", self._h, lbvh.handle, _ptr(counts), int(steps), int(batch), float(lr), int(seed),
             ctypes.byref(loss), _stream(stream))
        return float(loss.value)

    def margin(self, lbvh: LBVH, quantile: float = 0.995, stream=None) -> float:
        m = ctypes.c_float()
        call("nb_capacity_margin", self._h, lbvh.handle, float(quantile), ctypes.byref(m), _stream(stream))
        return float(m.value)

    def predict(self, lbvh: LBVH, margin: float, pad: int = 2, cap: Optional[torch.Tensor] = None,
                stream=None) -> torch.Tensor:
        N = lbvh.getN()
        if cap is None:
            cap = torch.empty(N, dtype=torch.int32, device="cuda")
        call("nb_capacity_predict", self._h, lbvh.handle, float(margin), int(pad), _ptr(cap), _stream(stream))
        return cap


def _predict_direct(self, lbvh: LBVH, lo, hi, r: float, margin: float, pad: int = 2,
                    cap: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """Capacities straight from the tree: features computed on the fly, in one kernel."""
    N = lbvh.getN()
    if cap is None:
        cap = torch.empty(N, dtype=torch.int32, device="cuda")
    call("nb_capacity_predict_direct", self._h, lbvh.handle, _vec3(lo), _vec3(hi), float(r), float(margin), int(pad),
         _ptr(cap), _stream(stream))
    return cap


CapacityPredictor.predict_direct = _predict_direct


def _predict_rows(self, lbvh: LBVH, lo, hi, r: float, margin: float, pad: int = 2, cap: Optional[torch.Tensor] = None,
                  base: Optional[torch.Tensor] = None, spill: Optional[torch.Tensor] = None, stream=None):
    """Capacities (as predict_direct) and the row bases of exact_rows_placed: base[g] = the
    capacities of the sorted positions before 256 g, base[ceil(N / 256)] = all of them; spill
    (2 int64, optional) is set to {that total, 0}.  Returns (cap, base)."""
    N = lbvh.getN()
    if cap is None:
        cap = torch.empty(N, dtype=torch.int32, device="cuda")
    if base is None:
        base = torch.empty((N + 255) // 256 + 1, dtype=torch.int64, device="cuda")
    if base.numel() < (N + 255) // 256 + 1:
        raise ValueError("base must hold ceil(N / 256) + 1 values")
    call("nb_capacity_predict_rows", self._h, lbvh.handle, _vec3(lo), _vec3(hi), float(r), float(margin), int(pad),
         _ptr(cap), _ptr(base), _ptr(spill), _stream(stream))
    return cap, base


CapacityPredictor.predict_rows = _predict_rows


def exact_rows_placed(trav: LBVHTraverser, lbvh: LBVH, spheres: torch.Tensor, points: torch.Tensor, base: torch.Tensor,
                    start: torch.Tensor, limit: torch.Tensor, neigh: torch.Tensor, image: Optional[torch.Tensor],
                    nneigh: torch.Tensor, spill: Optional[torch.Tensor] = None, images: Optional[torch.Tensor] = None,
                    transform_map: Optional[torch.Tensor] = None, stream=None) -> None:
    """Exact (neighbor, image) pairs of the tree's own primitives into rows placed by the wave
    walk (nb_traverse_exact_rows_placed): the query at sorted position t starts at
    base[t // 256] plus the limits of the positions before it in its group of 256; the start
    is written to start[q].  spill as in exact_rows (predict_rows initializes it)."""
    N = spheres.shape[0]
    _exact_points(trav, points)
    nimg = 0 if images is None else images.shape[0]
    call("nb_traverse_exact_rows_placed", trav._h, lbvh.handle, _ptr(spheres), N, _ptr(points), _ptr(images), nimg,
         _ptr(transform_map), _ptr(base), _ptr(start), _ptr(limit), _ptr(neigh), _ptr(image), _ptr(nneigh),
         neigh.numel(), _ptr(spill), _stream(stream))


def capacity_forward(params: torch.Tensor, features: torch.Tensor, stream=None) -> torch.Tensor:
    """Device inference forward pass y = MLP(features) (matrix layers on the matrix cores)."""
    y = torch.empty(features.shape[0], dtype=torch.float32, device="cuda")
    call("nb_capacity_forward", _ptr(params), _ptr(features), int(features.shape[0]), _ptr(y), _stream(stream))
    return y


def capacity_gradient(params: torch.Tensor, features: torch.Tensor, targets: torch.Tensor, stream=None) -> torch.Tensor:
    """Device gradient (1377 values + the loss at index 1377) of mean((MLP(x) - t)^2)."""
    g = torch.empty(N_PARAMS + 1, dtype=torch.float32, device="cuda")
    call("nb_capacity_gradient", _ptr(params), _ptr(features), _ptr(targets), int(features.shape[0]), _ptr(g),
         _stream(stream))
    return g


class PairRows:
    """Exact pairs in rows: query q's pairs are neigh[start[q] : start[q] + count[q]] (image likewise)."""

    def __init__(self, start, count, neigh, image, n_overflow, capacity_total):
        self.start = start
        self.count = count
        self.neigh = neigh
        self.image = image
        self.n_overflow = n_overflow
        self.capacity_total = capacity_total


class LearnedFill:
    """Reusable buffers of the single-pass learned fill (the pair buffer only grows)."""

    def __init__(self, N: int, capacity: int = 0):
        self.N = N
        self.cap = torch.empty(N, dtype=torch.int32, device="cuda")
        self.start = torch.empty(N + 1, dtype=torch.int64, device="cuda")
        self.count = torch.empty(N, dtype=torch.int32, device="cuda")
        self.list = torch.empty(N, dtype=torch.int32, device="cuda")
        self.n_list = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.spill = torch.zeros(2, dtype=torch.int64, device="cuda")
        self.base = torch.empty((N + 255) // 256 + 1, dtype=torch.int64, device="cuda")
        self.neigh = torch.empty(max(capacity, 1), dtype=torch.int32, device="cuda")
        self.image = torch.empty(max(capacity, 1), dtype=torch.uint8, device="cuda")
        self.sized = capacity > 0
        self.scratch = torch.empty(max(int(lib().nb_relocate_scratch_bytes(N)), 256), dtype=torch.uint8, device="cuda")

    def grow(self, capacity: int, keep: bool = True) -> None:
        if capacity > self.neigh.numel():
            n = torch.empty(capacity, dtype=torch.int32, device="cuda")
            i = torch.empty(capacity, dtype=torch.uint8, device="cuda")
            if keep:
                n[: self.neigh.numel()].copy_(self.neigh)
                i[: self.image.numel()].copy_(self.image)
            self.neigh, self.image = n, i
        self.sized = True


def learned_exact_pairs(trav: LBVHTraverser, lbvh: LBVH, pred: CapacityPredictor, spheres: torch.Tensor,
                        points: torch.Tensor, images: Optional[torch.Tensor], margin: float, pad: int = 2,
                        order=None, buffers: Optional[LearnedFill] = None, block: int = 256, stream=None,
                        in_kernel: bool = True, sync: bool = True, placed: Optional[tuple] = None) -> PairRows:
    """Exact pairs in one fill pass over rows of predicted capacity.

    in_kernel=True: overflowed rows are moved past the predicted rows and walked again inside
    the same launch (one host synchronization, to read the used size; if the buffer was too
    small it is grown and the pass repeated).  in_kernel=False: the fill drops the overflow,
    then the overflowed rows are listed, relocated and filled by a second traversal.

    sync=False (in_kernel only, buffers already sized): no host synchronization at all; the
    caller reads buffers.spill later (spill[0] = used positions, complete iff it is at most
    buffers.neigh.numel(); spill[1] = moved rows), and n_overflow / capacity_total are None.

    placed=(lo, hi, r) (in_kernel only; the queries are the tree's own primitives): the
    capacities come from predict_rows with the tree's scene box and query radius, and the rows
    are placed by the wave walk (exact_rows_placed) in the tree's Morton order, with no scan over
    the queries; `order` is not used.  The wave traversal variants only."""
    N = spheres.shape[0]
    _exact_points(trav, points)
    if buffers is None:
        buffers = LearnedFill(N)
    s = _stream(stream)
    nimg = 0 if images is None else images.shape[0]
    if not sync and not (in_kernel and buffers.sized):
        raise ValueError("sync=False needs in_kernel=True and pre-sized buffers")
    if placed is not None:
        if not in_kernel:
            raise ValueError("placed rows need in_kernel=True")
        lo, hi, r = placed
        G = (N + 255) // 256
        for attempt in range(2):
            pred.predict_rows(lbvh, lo, hi, r, margin, pad, buffers.cap, buffers.base, buffers.spill, stream)
            if not buffers.sized:
                total_cap = int(buffers.base[G].item())
                buffers.grow(total_cap + total_cap // 16 + 1024, keep=False)
            exact_rows_placed(trav, lbvh, spheres, points, buffers.base, buffers.start, buffers.cap, buffers.neigh,
                            buffers.image, buffers.count, spill=buffers.spill, images=images, stream=stream)
            if not sync:
                return PairRows(buffers.start[:N], buffers.count, buffers.neigh, buffers.image, None, None)
            used, moved = (int(v) for v in buffers.spill.tolist())
            if used <= buffers.neigh.numel():
                return PairRows(buffers.start[:N], buffers.count, buffers.neigh, buffers.image, moved, used)
            buffers.grow(used + used // 16, keep=False)
        raise RuntimeError("learned fill did not fit after growing its buffer")
    for attempt in range(2):
        pred.predict(lbvh, margin, pad, buffers.cap, stream)
        # rows in traversal order when one is given: a wave's queries fill adjacent rows
        exclusive_scan(buffers.cap, buffers.start, stream, perm=order)
        if not buffers.sized or not in_kernel:
            total_cap = int(buffers.start[N].item())
            buffers.grow(total_cap + (0 if not in_kernel else total_cap // 16 + 1024), keep=False)
        spill = None
        if in_kernel:
            buffers.spill[0:1].copy_(buffers.start[N:N + 1])
            buffers.spill[1:2].zero_()
            spill = buffers.spill
        call("nb_traverse_exact_rows", trav._h, lbvh.handle, _ptr(spheres), N, _ptr(points), _ptr(images), nimg, None,
             _ptr(order), _ptr(buffers.start), _ptr(buffers.cap), _ptr(buffers.neigh), _ptr(buffers.image),
             _ptr(buffers.count), buffers.neigh.numel(), _ptr(spill), block, s)
        if not in_kernel:
            n_ov = overflow_fallback(trav, lbvh, spheres, points, images, buffers, total_cap, block, stream)
            return PairRows(buffers.start[:N], buffers.count, buffers.neigh, buffers.image, n_ov, total_cap)
        if not sync:
            return PairRows(buffers.start[:N], buffers.count, buffers.neigh, buffers.image, None, None)
        used, moved = (int(v) for v in buffers.spill.tolist())
        if used <= buffers.neigh.numel():
            return PairRows(buffers.start[:N], buffers.count, buffers.neigh, buffers.image, moved, used)
        buffers.grow(used + used // 16, keep=False)
    raise RuntimeError("learned fill did not fit after growing its buffer")


def overflow_fallback(trav, lbvh, spheres, points, images, buffers: LearnedFill, tail: int, block: int = 256,
                      stream=None) -> int:
    """Find rows whose count exceeded their capacity, give them rows after `tail` and redo them exactly."""
    N = spheres.shape[0]
    s = _stream(stream)
    buffers.n_list.zero_()
    call("nb_overflow_rows", _ptr(buffers.count), _ptr(buffers.cap), N, _ptr(buffers.list), _ptr(buffers.n_list), s)
    n_ov = int(buffers.n_list.item())
    if n_ov == 0:
        return 0
    need = tail + int(buffers.count[buffers.list[:n_ov].long()].long().sum().item())
    buffers.grow(need)
    call("nb_relocate_rows", _ptr(buffers.list), n_ov, _ptr(buffers.count), int(tail), _ptr(buffers.start),
         _ptr(buffers.cap), _ptr(buffers.scratch), s)
    nimg = 0 if images is None else images.shape[0]
    call("nb_traverse_exact_rows", trav._h, lbvh.handle, _ptr(spheres), n_ov, _ptr(points), _ptr(images), nimg, None,
         _ptr(buffers.list), _ptr(buffers.start), _ptr(buffers.cap), _ptr(buffers.neigh), _ptr(buffers.image),
         _ptr(buffers.count), buffers.neigh.numel(), None, block, s)
    return n_ov
