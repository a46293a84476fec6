"""Snapshot benchmark harness with the reference benchmark's protocol
(mphowardlab/neighbor benchmark/lbvh_benchmark.cc:80-246).

    python -m neighbor_amd.harness <input.gsd|input.npz> <Nframes> <rcut> <output> [--calls K]

For each frame, starting from 0, as the reference harness does:
- a new LBVH over the frame's positions in its periodic box (lo = -L/2, hi = L/2, the HOOMD
  convention the reference reads through GSDReader);
- the build's block size tuned over 32..1024 step 32, then the median of 5 runs of K
  back-to-back builds;
- one sphere of radius rcut per particle, listed in the tree's primitive (Morton) order, and
  the 27 periodic images listed center first (lbvh_benchmark.cc:183-207);
- the count traversal of the reference's structure (quantized per-thread ropes, SphereQueryOp),
  its block size tuned the same way, then the median of 5 x K calls.

The output table keeps the reference's columns (frame, build (ms), traverse (ms)) and adds this
library's exact wave walk (FP64 pair test) as a fourth column.

Input: HOOMD-schema GSD files, as the reference reads through HOOMD's GSDReader
(neighbor_amd/gsd.py: configuration/box with zero tilt, particles/N, particles/position; a
chunk missing from a frame comes from frame 0), or an .npz holding `positions` (F x N x 3 or
N x 3, float32) and `box` (F x 3 or 3: the periodic box lengths).  The tuner samples each block
size 5 times rather than running the reference's 200 x 200 warm-up.
"""
from __future__ import annotations

import argparse
import sys
from typing import Dict, List

import numpy as np

#: block sizes the reference's tunables accept (Tunable.h: 32..1024 step 32)
BLOCKS = list(range(32, 1025, 32))


def read_frames(path: str):
    """(positions list of N x 3 float32 arrays, box list of 3-vectors) from a GSD file (by its
    .gsd suffix or its magic number) or an .npz file."""
    with open(path, "rb") as f:
        head = f.read(8)
    if path.endswith(".gsd") or (len(head) == 8 and int.from_bytes(head, "little") == 0x65DF65DF65DF65DF):
        from .gsd import read_hoomd_frames
        return read_hoomd_frames(path)
    with np.load(path, allow_pickle=False) as z:
        pos = np.asarray(z["positions"], dtype=np.float32)
        box = np.asarray(z["box"], dtype=np.float64)
    if pos.ndim == 2:
        pos = pos[None]
    if pos.ndim != 3 or pos.shape[2] != 3:
        raise ValueError("positions must be N x 3 or F x N x 3")
    if box.ndim == 1:
        box = np.broadcast_to(box, (pos.shape[0], 3))
    if box.shape != (pos.shape[0], 3) or not np.all(box > 0):
        raise ValueError("box must hold three positive lengths per frame")
    return [np.ascontiguousarray(p) for p in pos], [b.copy() for b in box]


def write_frames(path: str, positions, box) -> None:
    """Store frames for read_frames (positions: F x N x 3 or N x 3; box: F x 3 or 3)."""
    np.savez(path, positions=np.asarray(positions, dtype=np.float32), box=np.asarray(box, dtype=np.float64))


def image_offsets() -> List[tuple]:
    """The 27 periodic image offsets in box lengths, center first (lbvh_benchmark.cc:190-207)."""
    return [(0, 0, 0)] + [(ix, iy, iz) for ix in (-1, 0, 1) for iy in (-1, 0, 1) for iz in (-1, 0, 1)
                          if (ix, iy, iz) != (0, 0, 0)]


def frame_protocol(nb, torch, pts: np.ndarray, L, rcut: float, calls: int, bucket: int = 48, lo=None,
                   blocks=BLOCKS, keep_hits: bool = False) -> Dict[str, object]:
    """One frame of the reference protocol on the current device; returns the medians (ms),
    the tuned block sizes and the hit statistics of both traversals (keep_hits: also the
    per-particle hit counts, raw_hits and exact_hits).  L: the periodic box lengths; lo: the
    box's lower corner (default -L/2, the HOOMD convention)."""
    L = [float(x) for x in L]
    lo = [-0.5 * x for x in L] if lo is None else [float(x) for x in lo]
    hi = [lo[a] + L[a] for a in range(3)]
    N = pts.shape[0]
    d_pts = torch.from_numpy(np.array(pts, dtype=np.float32, order="C", copy=True)).cuda()
    sph = np.empty((N, 4), dtype=np.float32)
    sph[:, :3] = pts
    sph[:, 3] = np.float32(rcut)
    d_sph = torch.from_numpy(sph).cuda()
    offs = image_offsets()
    img32 = torch.tensor([[np.float32(L[a]) * np.float32(v[a]) for a in range(3)] for v in offs], dtype=torch.float32).cuda()
    img64 = torch.tensor([[L[a] * v[a] for a in range(3)] for v in offs], dtype=torch.float64).cuda()
    lb = nb.LBVH()
    lb.build(d_pts, lo, hi)
    hits = torch.empty(N, dtype=torch.int32, device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(f, n):
        ev[0].record()
        for _ in range(n):
            f()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / n

    def tune_and_time(f, candidates):
        best, best_ms = None, None
        for b in candidates:
            f(b)
            ms = float(np.median([timed(lambda: f(b), 1) for _ in range(5)]))
            if best_ms is None or ms < best_ms:
                best, best_ms = b, ms
        runs = sorted(timed(lambda: f(best), calls) for _ in range(5))
        return best, runs[2]

    bb, build_ms = tune_and_time(lambda b: lb.build(d_pts, lo, hi, block=b), blocks)
    lb.build(d_pts, lo, hi, block=bb)
    order = lb.device_pointers()["primitive"]          # spheres served in primitive (Morton) order
    ref = nb.LBVHTraverser(variant=nb.NB_VARIANT_QUANTIZED)
    ref.setup(lb)
    rb, ref_ms = tune_and_time(lambda b: ref.count(lb, d_sph, images=img32, order=order, nneigh=hits, block=b), blocks)
    torch.cuda.synchronize()
    raw = hits.to(torch.int64)
    wave = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
    wave.set_bucket_size(bucket)
    wave.setup(lb)
    wb, wave_ms = tune_and_time(
        lambda b: wave.exact_count(lb, d_sph, d_pts, images=img64, order=order, nneigh=hits, block=b), [256])
    torch.cuda.synchronize()
    exact = hits.to(torch.int64)
    out = {"N": N, "build_ms": build_ms, "build_block": bb, "traverse_reference_structure_ms": ref_ms,
           "traverse_reference_structure_block": rb, "traverse_exact_wave_ms": wave_ms,
           "raw_hits_min_max_mean": [int(raw.min()), int(raw.max()), float(raw.float().mean())],
           "exact_hits_min_max_mean": [int(exact.min()), int(exact.max()), float(exact.float().mean())]}
    if keep_hits:
        out.update(raw_hits=raw.cpu().numpy(), exact_hits=exact.cpu().numpy(), lo=lo, hi=hi,
                   images32=img32.cpu().numpy(), images64=img64.cpu().numpy())
    return out


def run(input_path: str, n_frames: int, rcut: float, output_path: str, calls: int = 500, bucket: int = 48,
        log=print, keep_hits: bool = False) -> List[Dict[str, object]]:
    import torch
    import neighbor_amd as nb
    positions, boxes = read_frames(input_path)
    if n_frames > len(positions):
        raise ValueError(f"{input_path} holds {len(positions)} frames, {n_frames} requested")
    results = []
    with open(output_path, "w") as out:
        out.write("# neighbor_amd snapshot harness (protocol of benchmark/lbvh_benchmark.cc)\n")
        out.write(f"# input: {input_path}  rcut: {rcut}  calls per run: {calls}\n#\n")
        out.write(f"# {'frame':>6}{'build (ms)':>16}{'traverse (ms)':>16}{'exact wave (ms)':>18}\n")
        for frame in range(n_frames):
            log(f"Frame {frame}")
            r = frame_protocol(nb, torch, positions[frame], boxes[frame], rcut, calls, bucket, keep_hits=keep_hits)
            log(f"  build {r['build_ms']:.5f} ms (block {r['build_block']}), traverse {r['traverse_reference_structure_ms']:.5f} ms "
                f"(block {r['traverse_reference_structure_block']}), exact wave {r['traverse_exact_wave_ms']:.5f} ms")
            out.write(f"{frame:>8} {r['build_ms']:>16.5f} {r['traverse_reference_structure_ms']:>16.5f} "
                      f"{r['traverse_exact_wave_ms']:>18.5f}\n")
            results.append(r)
    return results


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="LBVH snapshot benchmark (reference protocol)")
    ap.add_argument("input", help="HOOMD-schema .gsd file, or .npz with positions (F x N x 3) and box (F x 3)")
    ap.add_argument("nframes", type=int, help="frames to benchmark, starting from 0")
    ap.add_argument("rcut", type=float, help="query sphere radius")
    ap.add_argument("output", help="tabulated output file")
    ap.add_argument("--calls", type=int, default=500, help="back-to-back calls per timed run (reference: 500)")
    ap.add_argument("--bucket", type=int, default=48, help="leaves per bucket of the exact wave walk")
    a = ap.parse_args(argv)
    run(a.input, a.nframes, a.rcut, a.output, a.calls, a.bucket)
    return 0


if __name__ == "__main__":
    sys.exit(main())
