"""Python mirror of the reference's LBVH / LBVHTraverser over the C ABI.

Names and argument meaning follow mphowardlab/neighbor (``LBVH.h:47-219``,
``LBVHTraverser.h:60-355``); tensors are torch CUDA(HIP) tensors, streams are
torch streams.  Errors raise NeighborError (the reference raises
std::runtime_error for an invalid block size and for more than 32 images).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import torch

from . import _capi
from ._capi import P, call, lib


def _ptr(t):
    """Device pointer of a tensor (or a raw integer address); None stays NULL."""
    if t is None:
        return None
    if isinstance(t, int):
        return P(t)
    return P(t.data_ptr())


def _stream(stream) -> P:
    if stream is None:
        stream = torch.cuda.current_stream()
    return P(stream.cuda_stream)


def _vec3(v) -> ctypes.Array:
    a = (ctypes.c_float * 3)(*[float(x) for x in v])
    return a


def _exact_points(trav, points: torch.Tensor) -> torch.Tensor:
    """Points of the exact calls: float3 rows or 16-byte (x, y, z, w) records; sets the
    traverser's record stride to match (nb_traverser_set_point_stride)."""
    if not points.is_cuda or points.dtype != torch.float32 or not points.is_contiguous() or points.dim() != 2 \
            or points.shape[1] not in (3, 4):
        raise ValueError("points must be a contiguous float32 device tensor of shape [n, 3] or [n, 4]")
    call("nb_traverser_set_point_stride", trav._h, int(points.shape[1]))
    return points


def _require(t: torch.Tensor, dtype: torch.dtype, name: str, cols: Optional[int] = None) -> torch.Tensor:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise ValueError(f"{name} must have shape [n, {cols}]")
    return t


class LBVH:
    """Linear BVH (reference ``include/neighbor/LBVH.h:47``)."""

    def __init__(self):
        h = ctypes.c_void_p()
        call("nb_lbvh_create", ctypes.byref(h))
        self._h = h
        self._destroy = lib().nb_lbvh_destroy
        self._points = None

    def __del__(self):
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def build(self, points: torch.Tensor, lo: Sequence[float], hi: Sequence[float], radius: Optional[float] = None,
              block: int = 256, stream=None) -> None:
        """Build over float3 points (PointInsertOp) or constant-radius spheres (SphereInsertOp)."""
        _require(points, torch.float32, "points", 3)
        self._points = points
        N = points.shape[0]
        if radius is None:
            call("nb_lbvh_build_points", self._h, _ptr(points), N, _vec3(lo), _vec3(hi), block, _stream(stream))
        else:
            call("nb_lbvh_build_spheres", self._h, _ptr(points), N, float(radius), _vec3(lo), _vec3(hi), block,
                 _stream(stream))

    def build_postype(self, postype: torch.Tensor, lo: Sequence[float], hi: Sequence[float],
                      index: Optional[torch.Tensor] = None, radius: Optional[float] = None, block: int = 256,
                      stream=None) -> None:
        """Build over 16-byte records (x, y, z, w) read in place (nb_lbvh_build_postype): primitive
        i is record index[i] (int32 device tensor; None: every record), points or constant-radius
        spheres as build()."""
        _require(postype, torch.float32, "postype", 4)
        if index is not None:
            _require(index, torch.int32, "index")
        self._points = (postype, index)
        N = postype.shape[0] if index is None else index.shape[0]
        call("nb_lbvh_build_postype", self._h, _ptr(postype), _ptr(index), N, -1.0 if radius is None else float(radius),
             _vec3(lo), _vec3(hi), block, _stream(stream))

    def refit_postype(self, postype: torch.Tensor, index: Optional[torch.Tensor] = None, radius: Optional[float] = None,
                      block: int = 256, stream=None) -> None:
        """refit() over 16-byte records, as build_postype reads them."""
        _require(postype, torch.float32, "postype", 4)
        self._points = (postype, index)
        N = postype.shape[0] if index is None else index.shape[0]
        call("nb_lbvh_refit_postype", self._h, _ptr(postype), _ptr(index), N, -1.0 if radius is None else float(radius),
             block, _stream(stream))

    def refit(self, points: torch.Tensor, radius: Optional[float] = None, block: int = 256, stream=None) -> None:
        """New point (or sphere) boxes under the last build's hierarchy (nb_lbvh_refit_*); the
        same number of points.  Traversers set up on this tree need setup() again."""
        _require(points, torch.float32, "points", 3)
        self._points = points
        N = points.shape[0]
        if radius is None:
            call("nb_lbvh_refit_points", self._h, _ptr(points), N, block, _stream(stream))
        else:
            call("nb_lbvh_refit_spheres", self._h, _ptr(points), N, float(radius), block, _stream(stream))

    def info(self):
        N, ni, nn, root = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int32()
        call("nb_lbvh_info", self._h, ctypes.byref(N), ctypes.byref(ni), ctypes.byref(nn), ctypes.byref(root))
        return N.value, ni.value, nn.value, root.value

    def getN(self) -> int:
        return self.info()[0]

    def getNInternal(self) -> int:
        return self.info()[1]

    def getNNodes(self) -> int:
        return self.info()[2]

    def getRoot(self) -> int:
        return self.info()[3]

    def getTunableParameters(self):
        n = ctypes.c_uint32()
        call("nb_lbvh_tunable_parameters", self._h, None, 0, ctypes.byref(n))
        buf = (ctypes.c_uint32 * n.value)()
        call("nb_lbvh_tunable_parameters", self._h, buf, n.value, ctypes.byref(n))
        return list(buf)

    def error_word(self) -> int:
        """The build's error word (0 after a good build; synchronizes)."""
        return _read_error_word("nb_lbvh_error_word", self._h)

    def check(self, stream=None) -> None:
        """Synchronize the stream and raise NeighborError if the last build or refit gave up a
        bounded wait or list (its tree is incomplete)."""
        call("nb_lbvh_check", self._h, _stream(stream))

    def enable_sorted_event(self, on: bool = True) -> None:
        """Record an event in every later build once the sorted data are complete (nb_lbvh_enable_sorted_event)."""
        call("nb_lbvh_enable_sorted_event", self._h, int(bool(on)))

    def set_span_lds_limit(self, limit: int) -> None:
        """Span lists longer than `limit` climb through global memory in the bubble's span passes
        (nb_lbvh_set_span_lds_limit; 0: every span; a testing knob, the tree is the same)."""
        call("nb_lbvh_set_span_lds_limit", self._h, int(limit))

    def wait_sorted(self, stream=None) -> None:
        """Make `stream` wait for the sorted data of the last build (nb_lbvh_wait_sorted)."""
        call("nb_lbvh_wait_sorted", self._h, _stream(stream))

    def device_pointers(self) -> dict:
        """Raw device addresses of the tree arrays (owned by the LBVH; valid until the next build)."""
        ptrs = [ctypes.c_void_p() for _ in range(8)]
        call("nb_lbvh_arrays", self._h, *[ctypes.byref(p) for p in ptrs])
        names = ["parent", "left", "right", "primitive", "lo", "hi", "rope", "codes"]
        return {n: (p.value or 0) for n, p in zip(names, ptrs)}

    def arrays(self, stream=None) -> dict:
        """Copies of the tree arrays: parent, left, right, primitive, lo, hi, rope, codes."""
        N, ni, nn, _ = self.info()
        ptrs = [ctypes.c_void_p() for _ in range(8)]
        call("nb_lbvh_arrays", self._h, *[ctypes.byref(p) for p in ptrs])
        names = ["parent", "left", "right", "primitive", "lo", "hi", "rope", "codes"]
        shapes = {"parent": (nn,), "left": (ni,), "right": (ni,), "primitive": (N,), "lo": (nn, 3), "hi": (nn, 3),
                  "rope": (nn,), "codes": (N,)}
        dtypes = {"parent": torch.int32, "left": torch.int32, "right": torch.int32, "primitive": torch.int32,
                  "lo": torch.float32, "hi": torch.float32, "rope": torch.int32, "codes": torch.int32}
        s = _stream(stream)
        out = {}
        for name, p in zip(names, ptrs):
            t = torch.empty(shapes[name], dtype=dtypes[name], device="cuda")
            if p.value and t.numel():
                _capi.memcpy_d2d(t.data_ptr(), p.value, t.numel() * t.element_size(), s.value or 0)
            out[name] = t
        return out


class LBVHTraverser:
    """Rope traverser (reference ``include/neighbor/LBVHTraverser.h:60``)."""

    def __init__(self, variant: int = _capi.NB_VARIANT_QUANTIZED):
        h = ctypes.c_void_p()
        call("nb_traverser_create", ctypes.byref(h))
        self._h = h
        self._destroy = lib().nb_traverser_destroy
        self.set_variant(variant)

    def __del__(self):
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self._h = None

    def set_variant(self, variant: int) -> None:
        call("nb_traverser_set_variant", self._h, int(variant))
        self.variant = variant

    def set_bucket_size(self, bucket: int) -> None:
        call("nb_traverser_set_bucket_size", self._h, int(bucket))

    def set_tail_split(self, groups: int, split: int = 2) -> None:
        """Launch shape of the exact wave walks: per XCD, the last `groups` groups of 256
        queries run as `split` (1, 2 or 4) smaller blocks (default 32, 2); groups 0 = one block
        per group.  Results do not depend on it."""
        call("nb_traverser_set_tail_split", self._h, int(groups), int(split))

    @staticmethod
    def visit_log_words(n_queries: int) -> int:
        """u32 words of visit log a wave traversal of n_queries writes (3 per launched wave)."""
        return int(lib().nb_visit_log_words(int(n_queries)))

    def set_visit_log(self, log: Optional[torch.Tensor]) -> None:
        """Wave-walk statistics (nb_traverser_set_visit_log): while set, each wave-variant count
        traversal writes per launched wave (node records loaded, buckets tested, leaves in them)
        into log (int32 device tensor, visit_log_words(n) entries or more); other wave traversals
        are refused meanwhile; None switches it off."""
        self._visit_log = log
        call("nb_traverser_set_visit_log", self._h, _ptr(log), 0 if log is None else log.numel())

    def setup(self, lbvh: LBVH, transform_map: Optional[torch.Tensor] = None, block: int = 256, stream=None) -> None:
        call("nb_traverser_setup", self._h, lbvh.handle, _ptr(transform_map), block, _stream(stream))

    def reset(self) -> None:
        call("nb_traverser_reset", self._h)

    def data(self, lbvh: LBVH, stream=None) -> dict:
        """Copies of the compressed records (int4 per node) and root box / bin sizes."""
        nn = lbvh.getNNodes()
        ptrs = [ctypes.c_void_p() for _ in range(5)]
        call("nb_traverser_data", self._h, *[ctypes.byref(p) for p in ptrs])
        s = _stream(stream).value or 0
        rec = torch.empty((nn, 4), dtype=torch.int32, device="cuda")
        small = torch.empty((3, 3), dtype=torch.float32, device="cuda")
        if ptrs[0].value and nn:
            _capi.memcpy_d2d(rec.data_ptr(), ptrs[0].value, rec.numel() * 4, s)
        for i in range(3):
            if ptrs[1 + i].value:
                _capi.memcpy_d2d(small[i].data_ptr(), ptrs[1 + i].value, 12, s)
        out = dict(records=rec, root_lo=small[0], root_hi=small[1], bins=small[2], float_records=None)
        if ptrs[4].value and nn:
            fr = torch.empty((nn, 2, 4), dtype=torch.float32, device="cuda")
            _capi.memcpy_d2d(fr.data_ptr(), ptrs[4].value, fr.numel() * 4, s)
            out["float_records"] = fr
        return out

    # ----------------------------------------------------------------- raw (reference) semantics
    def count(self, lbvh: LBVH, spheres: torch.Tensor, images: Optional[torch.Tensor] = None,
              transform_map: Optional[torch.Tensor] = None, order: Optional[torch.Tensor] = None,
              nneigh: Optional[torch.Tensor] = None, block: int = 256, stream=None) -> torch.Tensor:
        """CountNeighborsOp with SphereQueryOp (and ImageListOp<float3> when images are given)."""
        _require(spheres, torch.float32, "spheres", 4)
        Nq = spheres.shape[0]
        if nneigh is None:
            nneigh = torch.empty(Nq, dtype=torch.int32, device=spheres.device)
        nimg = 0 if images is None else _require(images, torch.float32, "images", 3).shape[0]
        call("nb_traverse_count", self._h, lbvh.handle, _ptr(spheres), Nq, _ptr(images), nimg, _ptr(transform_map),
             _ptr(order), _ptr(nneigh), block, _stream(stream))
        return nneigh

    def neighbor_list(self, lbvh: LBVH, spheres: torch.Tensor, max_neigh: int, images: Optional[torch.Tensor] = None,
                      transform_map: Optional[torch.Tensor] = None, order: Optional[torch.Tensor] = None,
                      block: int = 256, stream=None):
        """NeighborListOp: returns (neigh_list [Nq, max_neigh], nneigh [Nq])."""
        _require(spheres, torch.float32, "spheres", 4)
        Nq = spheres.shape[0]
        lst = torch.zeros((Nq, max_neigh), dtype=torch.int32, device=spheres.device)
        nneigh = torch.empty(Nq, dtype=torch.int32, device=spheres.device)
        nimg = 0 if images is None else _require(images, torch.float32, "images", 3).shape[0]
        call("nb_traverse_list", self._h, lbvh.handle, _ptr(spheres), Nq, _ptr(images), nimg, _ptr(transform_map),
             _ptr(order), _ptr(lst), _ptr(nneigh), max_neigh, block, _stream(stream))
        return lst, nneigh

    def pairs(self, lbvh: LBVH, spheres: torch.Tensor, offsets: torch.Tensor, images: Optional[torch.Tensor] = None,
              transform_map: Optional[torch.Tensor] = None, order: Optional[torch.Tensor] = None,
              with_image: bool = True, block: int = 256, stream=None):
        """Raw hits as (neighbor, image) pairs into CSR rows given by offsets [Nq+1] (int64)."""
        Nq = spheres.shape[0]
        total = int(offsets[-1].item()) if Nq else 0
        neigh = torch.empty(max(total, 1), dtype=torch.int32, device=spheres.device)
        img = torch.empty(max(total, 1), dtype=torch.uint8, device=spheres.device) if with_image else None
        nimg = 0 if images is None else images.shape[0]
        call("nb_traverse_pairs", self._h, lbvh.handle, _ptr(spheres), Nq, _ptr(images), nimg, _ptr(transform_map),
             _ptr(order), _ptr(offsets), _ptr(neigh), _ptr(img), None, neigh.numel(), block, _stream(stream))
        return neigh[:total], (img[:total] if img is not None else None)

    # ----------------------------------------------------------------- exact FP64 semantics
    def exact_count(self, lbvh: LBVH, spheres: torch.Tensor, points: torch.Tensor,
                    images: Optional[torch.Tensor] = None, transform_map: Optional[torch.Tensor] = None,
                    order: Optional[torch.Tensor] = None, nneigh: Optional[torch.Tensor] = None,
                    block: int = 256, stream=None) -> torch.Tensor:
        _require(spheres, torch.float32, "spheres", 4)
        _exact_points(self, points)
        Nq = spheres.shape[0]
        if nneigh is None:
            nneigh = torch.empty(Nq, dtype=torch.int32, device=spheres.device)
        nimg = 0 if images is None else _require(images, torch.float64, "images", 3).shape[0]
        call("nb_traverse_exact_count", self._h, lbvh.handle, _ptr(spheres), Nq, _ptr(points), _ptr(images), nimg,
             _ptr(transform_map), _ptr(order), _ptr(nneigh), block, _stream(stream))
        return nneigh

    def exact_pairs(self, lbvh: LBVH, spheres: torch.Tensor, points: torch.Tensor, offsets: torch.Tensor,
                    images: Optional[torch.Tensor] = None, transform_map: Optional[torch.Tensor] = None,
                    order: Optional[torch.Tensor] = None, with_image: bool = True, total: Optional[int] = None,
                    neigh: Optional[torch.Tensor] = None, img: Optional[torch.Tensor] = None,
                    block: int = 256, stream=None):
        _exact_points(self, points)
        Nq = spheres.shape[0]
        if total is None:
            total = int(offsets[-1].item()) if Nq else 0
        if neigh is None:
            neigh = torch.empty(max(total, 1), dtype=torch.int32, device=spheres.device)
        if img is None and with_image:
            img = torch.empty(max(total, 1), dtype=torch.uint8, device=spheres.device)
        nimg = 0 if images is None else images.shape[0]
        call("nb_traverse_exact_pairs", self._h, lbvh.handle, _ptr(spheres), Nq, _ptr(points), _ptr(images), nimg,
             _ptr(transform_map), _ptr(order), _ptr(offsets), _ptr(neigh), _ptr(img), None, neigh.numel(), block,
             _stream(stream))
        return neigh[:total], (img[:total] if img is not None else None)


def _read_error_word(fn: str, handle) -> int:
    ptr = ctypes.c_void_p()
    call(fn, handle, ctypes.byref(ptr))
    if not ptr.value:
        return 0
    out = torch.empty(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    _capi.memcpy_d2d(out.data_ptr(), ptr.value, 4, 0)
    torch.cuda.synchronize()
    return int(out.item())


def exact_rows(trav: "LBVHTraverser", lbvh: LBVH, spheres: torch.Tensor, points: torch.Tensor, start: torch.Tensor,
               limit: torch.Tensor, neigh: torch.Tensor, image: Optional[torch.Tensor] = None,
               images: Optional[torch.Tensor] = None, order=None, nneigh: Optional[torch.Tensor] = None,
               spill: Optional[torch.Tensor] = None, transform_map=None, block: int = 256, stream=None) -> torch.Tensor:
    """Exact pairs into rows [start[q], start[q] + limit[q]) (nb_traverse_exact_rows); returns the counts.
    points: float3 rows or 16-byte (x, y, z, w) records."""
    _exact_points(trav, points)
    Nq = spheres.shape[0]
    if nneigh is None:
        nneigh = torch.empty(max(Nq, 1), dtype=torch.int32, device=spheres.device)
    nimg = 0 if images is None else images.shape[0]
    call("nb_traverse_exact_rows", trav._h, lbvh.handle, _ptr(spheres), Nq, _ptr(points), _ptr(images), nimg,
         _ptr(transform_map),
         _ptr(order), _ptr(start), _ptr(limit), _ptr(neigh), _ptr(image), _ptr(nneigh), neigh.numel(), _ptr(spill),
         block, _stream(stream))
    return nneigh


def max_displacement2(points: torch.Tensor, ref: torch.Tensor, box=None, out: Optional[torch.Tensor] = None,
                      wrapped: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """Device scalar (float64): max squared minimum-image displacement of points from ref
    (nb_max_displacement2); box: periodic lengths (3 floats) or None.  wrapped: optional
    device int32 scalar receiving the number of points moved across the box."""
    _require(points, torch.float32, "points", 3)
    _require(ref, torch.float32, "ref", 3)
    if ref.shape[0] != points.shape[0]:
        raise ValueError("points and ref differ in length")
    if out is None:
        out = torch.empty((), dtype=torch.float64, device=points.device)
    call("nb_max_displacement2", _ptr(points), _ptr(ref), points.shape[0], None if box is None else _vec3(box), _ptr(out),
         _ptr(wrapped), _stream(stream))
    return out


def filter_rows(start: torch.Tensor, counts: torch.Tensor, neigh: torch.Tensor, image: Optional[torch.Tensor] = None,
                counts_out: Optional[torch.Tensor] = None, query_index=None, half: bool = False, n_images: int = 27,
                points: Optional[torch.Tensor] = None, images: Optional[torch.Tensor] = None, types=None, shift=None,
                radius: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """In-place half-list and per-pair-radius filter of neighbor rows (nb_rows_filter); returns
    the kept counts.  radius: (ntypes, ntypes) float32 device tensor."""
    _require(start, torch.int64, "start")
    _require(counts, torch.int32, "counts")
    _require(neigh, torch.int32, "neigh")
    if image is not None:
        _require(image, torch.uint8, "image")
    if start.shape[0] < counts.shape[0]:
        raise ValueError("start needs one entry per row")
    if points is not None:
        _require(points, torch.float32, "points", 3)
        if images is None or radius is None:
            raise ValueError("the radius filter needs images and radius")
        _require(images, torch.float64, "images", 3)
        _require(radius, torch.float32, "radius")
        if radius.dim() != 2 or radius.shape[0] != radius.shape[1]:
            raise ValueError("radius must be (ntypes, ntypes)")
    Nq = counts.shape[0]
    if counts_out is None:
        counts_out = torch.empty_like(counts)
    ntypes = 0 if radius is None else int(radius.shape[0])
    call("nb_rows_filter", _ptr(start), _ptr(counts), Nq, _ptr(query_index), _ptr(neigh), _ptr(image), _ptr(counts_out),
         int(half), int(n_images), _ptr(points), _ptr(images), _ptr(types), _ptr(shift), _ptr(radius), ntypes,
         _stream(stream))
    return counts_out


def pack_rows(start: torch.Tensor, counts: torch.Tensor, neigh: torch.Tensor, image: Optional[torch.Tensor] = None,
              query_index=None, pair_stream: bool = False, stream=None) -> dict:
    """Gap-free CSR rows (offsets = scan of counts in row-index order) and, with pair_stream,
    the particle index of every pair (nb_rows_pack).  Returns dict(offsets, neigh, image, first)."""
    _require(start, torch.int64, "start")
    _require(counts, torch.int32, "counts")
    _require(neigh, torch.int32, "neigh")
    if image is not None:
        _require(image, torch.uint8, "image")
    if start.shape[0] < counts.shape[0]:
        raise ValueError("start needs one entry per row")
    Nq = counts.shape[0]
    offsets = exclusive_scan(counts, stream=stream)
    total = int(offsets[-1].item())
    dev = counts.device
    neigh_out = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
    image_out = torch.empty(max(total, 1), dtype=torch.uint8, device=dev) if image is not None else None
    first = torch.empty(max(total, 1), dtype=torch.int32, device=dev) if pair_stream else None
    call("nb_rows_pack", _ptr(start), _ptr(counts), Nq, _ptr(query_index), _ptr(offsets), _ptr(neigh), _ptr(image),
         _ptr(neigh_out), _ptr(image_out), _ptr(first), _stream(stream))
    return dict(offsets=offsets, neigh=neigh_out[:total], image=None if image_out is None else image_out[:total],
                first=None if first is None else first[:total])


_scan_scratch = {}


def exclusive_scan(counts: torch.Tensor, offsets: Optional[torch.Tensor] = None, stream=None, perm=None) -> torch.Tensor:
    """offsets[0..n] (int64) = exclusive scan of counts (int32/uint32), on the device.  With perm
    (a tensor or device address of n u32), the scan runs in perm order: offsets[perm[t]] is the
    sum of counts[perm[t']] for t' < t (rows laid out in traversal order); offsets[n] = total."""
    n = counts.shape[0]
    if offsets is None:
        offsets = torch.empty(n + 1, dtype=torch.int64, device=counts.device)
    nbytes = lib().nb_scan_scratch_bytes(n)
    # one scratch buffer per (device, stream): scans on different streams may run concurrently
    key = (counts.device, _stream(stream).value)
    buf = _scan_scratch.get(key)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(max(nbytes, 1024), dtype=torch.uint8, device=counts.device)
        _scan_scratch[key] = buf
    if perm is None:
        call("nb_exclusive_scan_u32_u64", _ptr(counts), n, _ptr(offsets), _ptr(buf), _stream(stream))
    else:
        call("nb_exclusive_scan_permuted_u32_u64", _ptr(counts), _ptr(perm), n, _ptr(offsets), _ptr(buf), _stream(stream))
    return offsets
