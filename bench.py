#!/usr/bin/env python3
"""Benchmark of the LBVH neighbor search on one MI355X.

Workload (one "step"): uniform N=1,000,000 points at density 0.85 in a periodic cube,
query radius 2.8, 27 periodic images, FP32 positions:
  full LBVH build (Morton codes + radix sort + Karras hierarchy + AABB bubble)
  + bucket records + capacity features + MLP capacity prediction + scan of the capacities
  + one exact FP64 (neighbor, image) fill pass into the predicted rows (rows that overflow
  are moved and walked again inside the same launch, so every row is exact).
The network is trained once beforehand on an independent configuration (seed 41).
Queries are the points themselves, served in the tree's Morton order; query q's pairs are
rows [start[q], start[q] + count[q]).  The fill's counts are checked against an exact count
pass after the timed steps.

value = exact FP64 pairs (self included) produced per second of whole step.
Also reported: the count-then-fill step (two tree walks) timed the same way, per-stage
times, the reference-semantics path (quantized rope traversal, CountNeighborsOp +
NeighborListOp) per variant, the roofline of the dominant kernel, and the oracle's CPU
timings (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic bytes per stage (neighbor_amd/roofline.json, the one machine-readable copy)
with open(os.path.join(ROOT, "neighbor_amd", "roofline.json")) as _f:
    ROOFLINE = json.load(_f)
HBM_PEAK_GBS = float(ROOFLINE["hbm_peak_GBs"])
BUILD_STAGE_BYTES = ROOFLINE["build_bytes_per_primitive"]
BUILD_BYTES_PER_PRIM = BUILD_STAGE_BYTES["total"]           # Morton + sort + Karras + bubble + compress
COUNT_BYTES_PER_QUERY = ROOFLINE["count_bytes_per_query"]["total"]     # query, count, tree once, positions once
FILL_EXTRA_BYTES_PER_QUERY = ROOFLINE["fill_bytes_per_query_extra"]["total"]   # row start, capacity
FILL_BYTES_PER_PAIR = ROOFLINE["fill_bytes_per_pair"]["total"]          # neighbor index + image byte
PREDICT_BYTES_PER_QUERY = ROOFLINE["predict_bytes_per_query"]["total"]
PREDICT_BYTES_PER_GROUP = ROOFLINE["predict_bytes_per_group_of_256"]["total"]
assert sum(v for k, v in BUILD_STAGE_BYTES.items() if k != "total") == BUILD_BYTES_PER_PRIM


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--N", type=int, default=1_000_000)
    ap.add_argument("--density", type=float, default=0.85)
    ap.add_argument("--r", type=float, default=2.8)
    ap.add_argument("--variant", type=int, default=3, help="traversal variant of the headline step")
    ap.add_argument("--bucket", type=int, default=48, help="leaves per bucket (wave variants)")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--tail-groups", type=int, default=None,
                    help="exact wave walks: last groups of 256 queries per XCD run as split blocks "
                         "(default: the library's, 32)")
    ap.add_argument("--tail-split", type=int, default=2, help="blocks per split group (1, 2 or 4)")
    ap.add_argument("--pad", type=int, default=8, help="capacity rows: predicted count + pad")
    ap.add_argument("--quantile", type=float, default=0.999,
                    help="capacity margin: this fraction of the training counts fit their predicted rows")
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the per-variant extras")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-capacity leg (clustered N=4M)")
    ap.add_argument("--learned-N", type=int, default=4_000_000)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched leg (4096 boxes) and its per-box loop")
    ap.add_argument("--batched-boxes", type=int, default=4096)
    ap.add_argument("--batch-z", type=float, default=4.0,
                    help="batched single pass: rows of ceil(lambda + z sqrt(lambda)) + pad positions")
    ap.add_argument("--no-loop", action="store_true", help="batched leg without the per-box loop baseline")
    ap.add_argument("--no-colloid", action="store_true", help="skip the colloid-mixture leg (per-type trees)")
    ap.add_argument("--colloid-N", type=int, default=2_000_000)
    ap.add_argument("--no-md", action="store_true", help="skip the MD-loop leg (skin test, refit, rebuild)")
    ap.add_argument("--no-harness", action="store_true", help="skip the reference-protocol leg (median of 5 x K calls)")
    ap.add_argument("--harness-calls", type=int, default=500)
    ap.add_argument("--md-steps", type=int, default=100)
    ap.add_argument("--no-build-scaling", action="store_true", help="skip the build-vs-N leg (uniform N up to 64M)")
    ap.add_argument("--build-sizes", type=str, default="1048576,4194304,16777216,67108864")
    ap.add_argument("--cpu-sample", type=int, default=20000, help="queries in the CPU baseline sample")
    ap.add_argument("--no-cpu-configs", action="store_true",
                    help="CPU baseline without the FP64 cell list at the clustered, colloid and batched configurations")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run 3 device-to-device copies of this many MiB (HBM counter calibration)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; multi-GPU sharding (queries by Morton "
                                   "range over a replicated tree, or slabs with ghost halos) is not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch
    import neighbor_amd as nb
    from neighbor_amd.api import exact_rows
    from neighbor_amd.learned import LearnedFill, exact_rows_placed
    from neighbor_amd.workloads import box_length, images27, spheres_from_points, uniform_points

    if args.calib_copy_mib > 0:
        src = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    N, r = args.N, args.r
    L = box_length(N, args.density)
    log(f"generating N={N} L={L:.4f}")
    pts = uniform_points(N, L, seed=42)
    sph = spheres_from_points(pts, r)
    d_pts = torch.from_numpy(pts).cuda()
    d_sph = torch.from_numpy(sph).cuda()
    d_img32 = torch.from_numpy(images27(L)).cuda()
    d_img64 = torch.from_numpy(images27(L, dtype=np.float64)).cuda()
    lo, hi = [0.0, 0.0, 0.0], [L, L, L]
    blk = args.block

    lb = nb.LBVH()
    tr = nb.LBVHTraverser(variant=args.variant)
    tr.set_bucket_size(args.bucket)
    if args.tail_groups is not None:
        tr.set_tail_split(args.tail_groups, args.tail_split)
    counts = torch.empty(N, dtype=torch.int32, device="cuda")
    fill_counts = torch.empty(N, dtype=torch.int32, device="cuda")
    offs = torch.empty(N + 1, dtype=torch.int64, device="cuda")

    # ---- the capacity network: trained once beforehand on an independent configuration of the
    # same workload (seed 41), by the project's own forward/backward kernels on the GPU
    pred, margin, train_info = train_predictor(args, nb, torch, np, N, L, r, blk)

    # size the pair buffers once (fixed input): count pass, then capacity = total
    lb.build(d_pts, lo, hi, block=blk)
    tr.setup(lb, block=blk)
    prim_ptr = lb.device_pointers()["primitive"]
    tr.exact_count(lb, d_sph, d_pts, images=d_img64, order=prim_ptr, nneigh=counts, block=blk)
    nb.exclusive_scan(counts, offs)
    torch.cuda.synchronize()
    total_pairs = int(offs[-1].item())
    counts_ref = counts.clone()
    cap = total_pairs
    neigh = torch.empty(cap, dtype=torch.int32, device="cuda")
    img = torch.empty(cap, dtype=torch.uint8, device="cuda")
    # single pass: rows of predicted capacity (about 1.3 positions per pair) plus room for the
    # rows that overflow theirs, moved and walked again inside the fill launch
    buf = LearnedFill(N, 2 * total_pairs)
    log(f"exact pairs per step: {total_pairs} ({total_pairs / N:.2f} per particle)")

    def build():
        lb.build(d_pts, lo, hi, block=blk)

    def compress():
        tr.setup(lb, block=blk)

    def count():
        tr.exact_count(lb, d_sph, d_pts, images=d_img64, order=prim_ptr, nneigh=counts, block=blk)

    def scan():
        # rows in traversal (Morton) order: the 64 queries of a wave write adjacent rows
        nb.exclusive_scan(counts, offs, perm=prim_ptr)

    def fill():
        exact_rows(tr, lb, d_sph, d_pts, offs, counts, neigh, img, images=d_img64, order=prim_ptr, nneigh=fill_counts,
                   block=blk)

    def predict():
        # features and MLP in one kernel (no feature array), which also sums the capacities of
        # each block of 256 queries; one scan over those sums gives the rows' block bases and the
        # spill cursor (no scan over the queries)
        pred.predict_rows(lb, lo, hi, r, margin, args.pad, buf.cap, buf.base, buf.spill)

    def single_fill():
        # each lane's row start = its block's base + the capacities before it in the block
        exact_rows_placed(tr, lb, d_sph, d_pts, buf.base, buf.start, buf.cap, buf.neigh, buf.image, buf.count,
                        spill=buf.spill, images=d_img64)

    # the timed single-pass step overlaps the prediction (which needs only the sorted codes and
    # points) with the rest of the build: a side stream (default priority) waits for the build's
    # sorted-data event, predicts, and the fill waits for it
    lb.enable_sorted_event(True)
    side = torch.cuda.Stream()
    ev_pred = torch.cuda.Event()

    def predict_side():
        lb.wait_sorted(stream=side)
        with torch.cuda.stream(side):
            pred.predict_rows(lb, lo, hi, r, margin, args.pad, buf.cap, buf.base, buf.spill, stream=side)
        ev_pred.record(side)

    def fill_after_predict():
        torch.cuda.current_stream().wait_event(ev_pred)
        single_fill()

    two_pass = [("build", build), ("compress", compress), ("count", count), ("scan", scan), ("fill", fill)]
    one_pass = [("build", build), ("predict", predict_side), ("compress", compress), ("fill", fill_after_predict)]
    # the same stages one after the other on one stream, for the per-stage times
    one_pass_seq = [("build", build), ("compress", compress), ("predict", predict), ("fill", single_fill)]

    def timed_steps(stages, steps, warmup):
        for _ in range(warmup):
            for _, f in stages:
                f()
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        t0.record()
        for _ in range(steps):
            for _, f in stages:
                f()
        t1.record()
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / steps, time.perf_counter() - wall0

    # ---- timed region: exactly K steps of the single-pass step
    ms_per_step, wall = timed_steps(one_pass, args.steps, args.warmup)
    used, moved = (int(v) for v in buf.spill.tolist())
    assert used <= buf.neigh.numel(), "single-pass rows did not fit their buffer"
    assert torch.equal(buf.count, counts_ref), "single-pass counts differ from the exact count pass"
    lb.check()      # raises if a build of the timed steps gave up a bounded wait (incomplete tree)

    # ---- the count-then-fill step (two tree walks), timed the same way
    ms_two, _ = timed_steps(two_pass, args.steps, args.warmup)
    check_total = int(offs[-1].item())
    assert check_total == total_pairs, "pair count changed between steps"
    assert torch.equal(fill_counts, counts), "fill pass counts differ from the count pass"
    lb.check()

    # ---- per-stage times (events between launches on the same stream)
    reps = max(3, min(args.steps, 10))

    def stage_times(stages):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(stages) + 1)]
        acc = [0.0] * len(stages)
        for _ in range(reps):
            ev[0].record()
            for i, (_, f) in enumerate(stages):
                f()
                ev[i + 1].record()
            torch.cuda.synchronize()
            for i in range(len(stages)):
                acc[i] += ev[i].elapsed_time(ev[i + 1])
        return {name: acc[i] / reps for i, (name, _) in enumerate(stages)}

    stage_ms = stage_times(one_pass_seq)
    stage_two = stage_times(two_pass)
    dominant = max(stage_ms, key=stage_ms.get)

    # algorithmic bytes per stage of the single-pass step (one launch group each)
    alg_bytes = {
        "build": float(BUILD_BYTES_PER_PRIM - BUILD_STAGE_BYTES["compress"]) * N,
        "compress": float(BUILD_STAGE_BYTES["compress"]) * N,
        "predict": float(PREDICT_BYTES_PER_QUERY * N + PREDICT_BYTES_PER_GROUP * ((N + 255) // 256)),
        "fill": float((COUNT_BYTES_PER_QUERY + FILL_EXTRA_BYTES_PER_QUERY) * N + FILL_BYTES_PER_PAIR * total_pairs),
    }
    achieved = alg_bytes[dominant] / (stage_ms[dominant] * 1e-3) / 1e9
    # build + compression back to back (no events between their launches)
    eb = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    eb[0].record()
    for _ in range(reps):
        build()
        compress()
    eb[1].record()
    torch.cuda.synchronize()
    build_total_ms = eb[0].elapsed_time(eb[1]) / reps
    build_gbs = BUILD_BYTES_PER_PRIM * N / (build_total_ms * 1e-3) / 1e9
    traffic, traffic_note = profiled_traffic(dominant)

    # wave-walk statistics of one untimed count pass (per-wave visit log)
    walk_stats = None
    if args.variant in (nb.NB_VARIANT_WAVE, nb.NB_VARIANT_WAVE_QUANTIZED):
        vlog = torch.zeros(nb.LBVHTraverser.visit_log_words(N), dtype=torch.int32, device="cuda")
        tr.set_visit_log(vlog)
        count()
        torch.cuda.synchronize()
        tr.set_visit_log(None)
        v = vlog.view(-1, 3).to(torch.int64).sum(dim=0).tolist()
        cnt_s = stage_two["count"] * 1e-3
        walk_stats = {"wave_node_records_loaded": v[0], "wave_bucket_visits": v[1], "wave_bucket_leaves": v[2],
                      "node_tests_per_query": v[0] * 64 / N, "leaf_tests_per_query": v[2] * 64 / N,
                      "bucket_visits_per_wave": v[1] / ((N + 63) // 64),
                      "node_records_per_s_count": v[0] / cnt_s, "node_tests_per_s_count": v[0] * 64 / cnt_s,
                      "note": "a wave loads each visited node record once for its 64 queries (node and leaf "
                              "tests count 64 lanes each); rates over the count stage time"}

    result = {
        "metric": "exact neighbor pairs/s (N=1M uniform rho=0.85 r=2.8, 27 images; full LBVH build + neighbor "
                  "list per step)",
        "value": total_pairs / (ms_per_step * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (exact refine f64)",
        "data": "synthetic (splitmix64 uniform points, seed 42)",
        "config": {"workload": f"uniform periodic cube N={N} rho={args.density} r={r} 27 images",
                   "N": N, "box_L": L, "pairs_per_step": total_pairs, "variant": args.variant, "block": blk,
                   "bucket": args.bucket, "parallelism": "single GPU",
                   "step": "LBVH build + bucket records + capacity prediction (cell-occupancy features and "
                           "the MLP in one kernel, which also sums the capacities of each block of 256 queries) "
                           "+ scan of those sums + one exact (neighbor, image) fill into the predicted rows, "
                           "each row's start placed by the walk, overflowed rows moved and walked again inside "
                           "the fill (every row exact); the prediction runs on a second stream once the build's "
                           "sorted data exist, beside the bubble passes and the records"},
        "build_ms": build_total_ms,
        "stage_ms": stage_ms,
        "single_pass": {"overflowed_rows": moved, "row_positions_used": used,
                        "row_positions_per_pair": used / total_pairs, "margin_log2": margin,
                        "margin_quantile": args.quantile, "pad": args.pad,
                        **train_info},
        "count_then_fill": {"ms_per_step": ms_two, "pairs_per_s": total_pairs / (ms_two * 1e-3),
                            "stage_ms": stage_two,
                            "step": "LBVH build + bucket records + exact count + scan (traversal order) + exact "
                                    "(neighbor, image) fill: two tree walks"},
        "wave_walk": walk_stats,
        "build_roofline": {"bound": "hbm", "achieved": build_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": build_gbs / HBM_PEAK_GBS, "bytes_per_prim": BUILD_BYTES_PER_PRIM},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "algorithmic_bytes": alg_bytes[dominant]},
        "wall_s_timed": wall,
    }

    # ---- reference-semantics path (raw quantized hits), per variant
    if not args.no_extra:
        extra = {}
        rcounts = torch.empty(N, dtype=torch.int32, device="cuda")
        max_neigh = 128
        nl = torch.empty((N, max_neigh), dtype=torch.int32, device="cuda")
        for var in (0, 1, 4):
            tv = nb.LBVHTraverser(variant=var)
            tv.set_bucket_size(args.bucket)
            tv.setup(lb, block=blk)
            from neighbor_amd._capi import call, P
            from neighbor_amd.api import _stream

            def raw_count():
                tv.count(lb, d_sph, images=d_img32, order=prim_ptr, nneigh=rcounts, block=blk)

            def raw_list():
                call("nb_traverse_list", tv._h, lb.handle, P(d_sph.data_ptr()), N, P(d_img32.data_ptr()), 27, None,
                     P(prim_ptr), P(nl.data_ptr()), P(rcounts.data_ptr()), max_neigh, blk, _stream(None))

            def raw_compress():
                tv.setup(lb, block=blk)

            for _ in range(2):
                raw_compress(); raw_count(); raw_list()
            torch.cuda.synchronize()
            e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            a = [0.0, 0.0, 0.0]
            for _ in range(reps):
                e[0].record(); raw_compress(); e[1].record(); raw_count(); e[2].record(); raw_list(); e[3].record()
                torch.cuda.synchronize()
                for i in range(3):
                    a[i] += e[i].elapsed_time(e[i + 1])
            raw_hits = int(rcounts.to(torch.int64).sum().item())
            extra[f"variant{var}"] = {"compress_ms": a[0] / reps, "count_ms": a[1] / reps, "list_ms": a[2] / reps,
                                      "raw_hits": raw_hits,
                                      "exact_pairs_per_s_traversal": total_pairs / ((a[1] + a[2]) / reps * 1e-3)}
        result["reference_semantics"] = extra

    # ---- learned capacity: single-pass fill vs count-then-fill on the clustered workload
    if not args.no_learned:
        try:
            result["learned"] = learned_leg(args, nb, torch, np)
        except Exception as exc:  # reported, the headline stands on its own
            result["learned"] = {"error": repr(exc)}

    # ---- batched boxes: one graph launch vs looping the single-tree path per box
    if not args.no_batched:
        try:
            result["batched"] = batched_leg(args, nb, torch, np)
        except Exception as exc:  # reported, the headline stands on its own
            result["batched"] = {"error": repr(exc)}

    # ---- size-asymmetric colloid mixture: one tree per type, a radius per type pair
    if not args.no_colloid:
        try:
            result["colloid"] = colloid_leg(args, nb, torch, np)
        except Exception as exc:  # reported, the headline stands on its own
            result["colloid"] = {"error": repr(exc)}

    # ---- the reference benchmark's own protocol (tuned block size, median of 5 x 500 calls)
    if not args.no_harness:
        try:
            result["reference_protocol"] = harness_leg(args, nb, torch, np)
        except Exception as exc:  # reported, the headline stands on its own
            result["reference_protocol"] = {"error": repr(exc)}

    # ---- the build alone at growing N (HBM fraction of the build + compression)
    if not args.no_build_scaling:
        try:
            result["build_scaling"] = build_scaling_leg(args, nb, torch, np)
        except Exception as exc:  # reported, the headline stands on its own
            result["build_scaling"] = {"error": repr(exc)}

    # ---- molecular-dynamics loop: keep / refit / rebuild the list by the skin test
    if not args.no_md:
        try:
            result["md_loop"] = md_leg(args, nb, torch, np)
        except Exception as exc:  # reported, the headline stands on its own
            result["md_loop"] = {"error": repr(exc)}

    # ---- CPU baselines (the oracle, timed on this host's cores): FP64 cell list and brute force,
    # and the restatement of the reference's LBVH path on one core
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline_leg(args, np, pts, sph, L, lo, hi, total_pairs, result)
        except Exception as exc:  # the GPU result stands on its own
            result["cpu_baseline"] = {"error": repr(exc)}
        if not args.no_cpu_configs and "error" not in result["cpu_baseline"]:
            try:
                result["cpu_baseline"]["celllist_configs"] = cpu_config_leg(args, np, result)
            except Exception as exc:
                result["cpu_baseline"]["celllist_configs"] = {"error": repr(exc)}

    print(json.dumps(result), flush=True)


def train_predictor(args, nb, torch, np, N, L, r, blk):
    """Train the capacity MLP on the exact counts of an independent configuration (seed 41) of the
    headline workload; returns (predictor, margin, info).  Not timed as part of the step."""
    from neighbor_amd.learned import CapacityPredictor
    from neighbor_amd.workloads import images27, spheres_from_points, uniform_points
    pts = uniform_points(N, L, seed=41)
    d_pts = torch.from_numpy(pts).cuda()
    d_sph = torch.from_numpy(spheres_from_points(pts, r)).cuda()
    d_img = torch.from_numpy(images27(L, dtype=np.float64)).cuda()
    lb = nb.LBVH()
    lb.build(d_pts, [0.0, 0.0, 0.0], [L, L, L], block=blk)
    tr = nb.LBVHTraverser(variant=args.variant)
    tr.set_bucket_size(args.bucket)
    if args.tail_groups is not None:
        tr.set_tail_split(args.tail_groups, args.tail_split)
    tr.setup(lb, block=blk)
    counts = tr.exact_count(lb, d_sph, d_pts, images=d_img, order=lb.device_pointers()["primitive"], block=blk)
    pred = CapacityPredictor(seed=1)
    pred.features(lb, [0.0, 0.0, 0.0], [L, L, L], r)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = pred.train(lb, counts, steps=300, batch=16384, lr=3e-3, seed=0)
    train_s = time.perf_counter() - t0
    margin = pred.margin(lb, args.quantile)
    return pred, margin, {"trained_on": "seed 41 (an independent configuration of the same workload)",
                          "train_steps": 300, "train_batch": 16384, "train_s": train_s, "train_loss": loss}


def host_threads():
    """Threads of this process's CPU share: OMP_NUM_THREADS when set (16 per GPU on the GPU
    hosts), else the visible CPUs."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)
    return max(1, min(n, os.cpu_count() or n))


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_config_leg(args, np, result):
    """The oracle's FP64 cell list (all threads of this process's CPU share) at the other
    configurations, each over its whole workload, with the pair total checked against the GPU
    leg's where that leg ran: clustered N=4M (the learned leg's applied configuration, seed 43),
    the colloid mixture N=2M (four type pairs) and the 4096 batched boxes (box by box)."""
    import oracle
    from neighbor_amd.workloads import batched_boxes, clustered_points, colloid_mixture, spheres_from_points
    threads = host_threads()
    out = {"threads": threads}

    def timed(f):
        t = time.perf_counter()
        res = f()
        return time.perf_counter() - t, res

    log("cpu baseline: FP64 cell list, clustered N=4000000")
    pts, L = clustered_points(4_000_000, 0.85, n_clusters=1000, sigma=5.0, background=0.2, seed=43)
    sph = spheres_from_points(pts, 2.8)
    t, ref = timed(lambda: oracle.exact_celllist(pts, sph, [L, L, L], 2.8, pairs=False, nthreads=threads))
    gpu = result.get("learned", {})
    if "pairs" in gpu:
        assert ref["total"] == gpu["pairs"], f"clustered: cell list {ref['total']} pairs, GPU {gpu['pairs']}"
    out["clustered_n4m"] = {"pairs": ref["total"], "s": t, "pairs_per_s": ref["total"] / t,
                            "gpu_pairs_per_s": gpu.get("pairs_per_s_single_pass")}
    del pts, sph

    log("cpu baseline: FP64 cell list, colloid N=2000000")
    d = [1.0, 10.0]
    R = [[0.5 * (d[a] + d[b]) + 0.4 for b in range(2)] for a in range(2)]
    pts, types, L = colloid_mixture(2_000_000, seed=7)
    total, ts = 0, 0.0
    for a in range(2):
        qa = pts[types == a]
        for b in range(2):
            xb = np.ascontiguousarray(pts[types == b])
            t, ref = timed(lambda: oracle.exact_celllist(xb, spheres_from_points(qa, R[a][b]), [L, L, L], R[a][b],
                                                         pairs=False, nthreads=threads))
            total += ref["total"]
            ts += t
    gpu = result.get("colloid", {})
    if "pairs" in gpu:
        assert total == gpu["pairs"], f"colloid: cell list {total} pairs, GPU {gpu['pairs']}"
    out["colloid_n2m"] = {"pairs": total, "s": ts, "pairs_per_s": total / ts, "gpu_pairs_per_s": gpu.get("pairs_per_s")}
    del pts, types

    log("cpu baseline: FP64 cell list, 4096 batched boxes")
    pts, offs, blo, bhi = batched_boxes(nbox=args.batched_boxes, nmin=256, nmax=65536, density=0.85, seed=42)
    total, ts = 0, 0.0
    for bx in range(offs.shape[0] - 1):
        p = pts[offs[bx]:offs[bx + 1]]
        if p.shape[0] == 0:
            continue
        t, ref = timed(lambda: oracle.exact_celllist(p, spheres_from_points(p, 2.8), bhi[bx], 2.8, pairs=False,
                                                     nthreads=threads))
        total += ref["total"]
        ts += t
    gpu = result.get("batched", {})
    if "pairs" in gpu:
        assert total == gpu["pairs"], f"batched: cell list {total} pairs, GPU {gpu['pairs']}"
    out["batched_4096"] = {"pairs": total, "s": ts, "pairs_per_s": total / ts,
                           "gpu_pairs_per_s": gpu.get("pairs_per_s_batched")}
    return out


def cpu_baseline_leg(args, np, pts, sph, L, lo, hi, total_pairs, result):
    """The reference has no CPU neighbor search beyond its test's O(N^2) loop
    (test/lbvh_test.cu:459-490).  Timed here on the host, all from the oracle:
    - value: the FP64 cell list (cell >= r, 27 images) over the whole headline workload on all
      threads of this process's CPU share, median of 3;
    - the same on one core, over the first 200,000 queries, scaled to N;
    - the FP64 brute force over 27 explicit images (the test loop restated) at N=10,000, rho=1,
      on all threads and on one core (1,000 queries scaled to 10,000);
    - the oracle's restatement of the reference's LBVH path (build + count and list traversal)
      on one core, a query sample scaled to N."""
    import oracle
    from neighbor_amd.workloads import box_length, images27, spheres_from_points, uniform_points
    threads = host_threads()
    N = pts.shape[0]
    r = float(sph[0, 3])
    log(f"cpu baseline: FP64 cell list on {threads} threads")

    def timed(f, reps):
        ts = []
        out = None
        for _ in range(reps):
            t = time.perf_counter()
            out = f()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts)), out

    t_all, ref = timed(lambda: oracle.exact_celllist(pts, sph, [L, L, L], r, pairs=False, nthreads=threads), 3)
    assert ref["total"] == total_pairs, f"cell list found {ref['total']} pairs, the GPU {total_pairs}"
    n1 = min(N, 200_000)
    t_one, ref1 = timed(lambda: oracle.exact_celllist(pts, sph[:n1], [L, L, L], r, pairs=False, nthreads=1), 1)
    t_one_full = t_one * N / n1
    # brute force at the small configuration (N=10^4, rho=1)
    Ns = 10_000
    Ls = box_length(Ns, 1.0)
    ps = uniform_points(Ns, Ls, seed=42)
    ss = spheres_from_points(ps, r)
    im = images27(Ls, dtype=np.float64)
    log(f"cpu baseline: FP64 brute force N={Ns}")
    t_bf, bf = timed(lambda: oracle.exact_bruteforce(ps, ss, im, pairs=False, nthreads=threads), 3)
    t_bf1, _ = timed(lambda: oracle.exact_bruteforce(ps, ss[:1000], im, pairs=False, nthreads=1), 1)
    t_bf1 *= Ns / 1000
    oracle.lib().orc_set_threads(threads)
    # the reference's LBVH path restated (one core)
    log("cpu baseline: LBVH restatement (1 core)")
    t_b = time.perf_counter()
    comp = oracle.compress(oracle.build(pts, lo, hi))
    t_build = time.perf_counter() - t_b
    sample = min(args.cpu_sample, N)
    idx = np.random.default_rng(0).choice(N, sample, replace=False)
    t_t = time.perf_counter()
    rr = oracle.traverse(N, comp, sph[idx], images27(L), max_neigh=128)
    t_trav = (time.perf_counter() - t_t) * 2.0       # count + list passes
    per_query = t_trav / sample
    # node visits of the reference-structure walk from the oracle's exact count on the sample
    ref_sem = result.get("reference_semantics", {})
    for var in ("variant0", "variant1"):
        if var in ref_sem:
            ref_sem[var]["node_visits_per_query"] = rr["visits"] / sample
            ref_sem[var]["node_visits_per_s_count"] = rr["visits"] / sample * N / (ref_sem[var]["count_ms"] * 1e-3)
    lbvh_step = t_build + per_query * N
    return {
        "value": total_pairs / t_all, "unit": "pairs/s", "cores": threads, "kind": "port",
        "sample": f"the whole headline workload ({N} queries, {total_pairs} exact pairs): FP64 cell list "
                  f"(oracle/oracle.c orc_exact_celllist, cell >= r, 27 images) on {threads} threads, "
                  f"median of 3, {t_all:.3f} s",
        "host_cpus": os.cpu_count(), "threads": threads, "cpu_model": cpu_model(),
        "celllist_all_threads_s": t_all,
        "celllist_one_core": {"pairs_per_s": total_pairs / t_one_full, "s_scaled": t_one_full,
                              "sample": f"first {n1} queries ({t_one:.2f} s) scaled to {N}"},
        "bruteforce_small": {"workload": f"N={Ns} rho=1 r={r:.6g} 27 images (test/lbvh_test.cu:459-490 restated in FP64)",
                             "pairs": bf["total"], "all_threads_s": t_bf, "pairs_per_s_all_threads": bf["total"] / t_bf,
                             "one_core_s_scaled": t_bf1, "pairs_per_s_one_core": bf["total"] / t_bf1},
        "lbvh_restatement_one_core": {"pairs_per_s": total_pairs / lbvh_step, "build_s": t_build,
                                      "traverse_s_per_query": per_query, "raw_hits_sample": int(rr["total"]),
                                      "sample": f"build of all {N} points + count and list traversal of {sample} "
                                                f"sampled queries, scaled to {N}"},
    }


#: kernel names of the stages in the committed rocprof summary of this step
PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "exact_step_uniform_n1m_kernels.md")
PROFILE_KERNELS = {"fill": "kernel::lbvh_traverse_wave_pairs<PlacedRowPairOp",
                   "count": "kernel::lbvh_traverse_wave_pairs<CountNeighborsOp"}


def profiled_traffic(stage):
    """HBM bytes per launch of a stage's kernel (calibrated FETCH_SIZE + WRITE_SIZE) from the
    committed rocprof summary of this step, or None when it has no row for the kernel."""
    key = PROFILE_KERNELS.get(stage)
    if key is None or not os.path.exists(PROFILE):
        return None, None
    for line in open(PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 6 and key in cells[0]:
            try:
                fetch, write = float(cells[4]), float(cells[5])
            except ValueError:
                return None, None
            return (fetch + write) * 2.0 ** 20, (f"profiles/{os.path.basename(PROFILE)}: FETCH_SIZE x2 + WRITE_SIZE x1 "
                                                  f"per dispatch ({fetch:.1f} + {write:.1f} MiB), avg {cells[2]} us")
    return None, None


def learned_leg(args, nb, torch, np, uniform=False):
    """Clustered N points (1000 Gaussian clusters of sigma 5 per 4M points + 20% uniform, rho 0.85,
    r 2.8, 27 images, exact pairs), or (uniform) the headline configuration.  The MLP is
    trained on the counts of one configuration and applied to an independent one; reported:
    the exact traversal time of count -> scan -> fill against features -> predict -> scan ->
    single fill (overflowed rows redone inside the fill)."""
    from neighbor_amd.api import exact_rows
    from neighbor_amd.learned import CapacityPredictor, LearnedFill, learned_exact_pairs
    from neighbor_amd.workloads import box_length, clustered_points, images27, spheres_from_points, uniform_points
    N, r = (args.N, args.r) if uniform else (args.learned_N, 2.8)
    ncl = max(1, round(1000 * N / 4_000_000))
    blk = args.block
    log(f"learned leg: {'uniform' if uniform else 'clustered'} N={N}")

    def frame(seed):
        if uniform:
            L = box_length(N, args.density)
            pts = uniform_points(N, L, seed=seed)
        else:
            pts, L = clustered_points(N, 0.85, n_clusters=ncl, sigma=5.0, background=0.2, seed=seed)
        d_pts = torch.from_numpy(pts).cuda()
        d_sph = torch.from_numpy(spheres_from_points(pts, r)).cuda()
        lb = nb.LBVH()
        lb.build(d_pts, [0, 0, 0], [L, L, L], block=blk)
        tr = nb.LBVHTraverser(variant=args.variant)
        tr.set_bucket_size(args.bucket)
        if args.tail_groups is not None:
            tr.set_tail_split(args.tail_groups, args.tail_split)
        tr.setup(lb, block=blk)
        return L, d_pts, d_sph, lb, tr

    train_seed, apply_seed = (41, 42) if uniform else (42, 43)
    L, d_pts, d_sph, lb, tr = frame(train_seed)
    d_img = torch.from_numpy(images27(L, dtype=np.float64)).cuda()
    order = lb.device_pointers()["primitive"]
    counts = torch.empty(N, dtype=torch.int32, device="cuda")
    tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order, nneigh=counts, block=blk)
    pred = CapacityPredictor(seed=1)
    pred.features(lb, [0, 0, 0], [L, L, L], r)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = pred.train(lb, counts, steps=300, batch=16384, lr=3e-3, seed=0)
    train_s = time.perf_counter() - t0
    margin = pred.margin(lb, args.quantile)

    # an independent configuration of the same workload
    L, d_pts, d_sph, lb, tr = frame(apply_seed)
    order = lb.device_pointers()["primitive"]
    pred.features(lb, [0, 0, 0], [L, L, L], r)
    offs = torch.empty(N + 1, dtype=torch.int64, device="cuda")
    tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order, nneigh=counts, block=blk)
    nb.exclusive_scan(counts, offs)
    total = int(offs[-1].item())
    neigh = torch.empty(total, dtype=torch.int32, device="cuda")
    img = torch.empty(total, dtype=torch.uint8, device="cuda")
    buf = LearnedFill(N, int(total * 3))

    fill_counts = torch.empty_like(counts)

    def two_pass():
        # as the headline step: rows in traversal order
        tr.exact_count(lb, d_sph, d_pts, images=d_img, order=order, nneigh=counts, block=blk)
        nb.exclusive_scan(counts, offs, perm=order)
        exact_rows(tr, lb, d_sph, d_pts, offs, counts, neigh, img, images=d_img, order=order, nneigh=fill_counts,
                   block=blk)

    def one_pass():
        # prediction for this configuration (with the per-block capacity sums), their scan, one
        # fill with rows placed by the wave walk
        return learned_exact_pairs(tr, lb, pred, d_sph, d_pts, d_img, margin, pad=args.pad, buffers=buf, block=blk,
                                   placed=([0, 0, 0], [L, L, L], r))

    def timed(f, reps):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            f()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts)) * 1e3

    for f in (two_pass, one_pass):
        f()
    res = one_pass()
    assert int(res.count.to(torch.int64).sum().item()) == total, "learned fill lost pairs"
    t2 = timed(two_pass, 7)
    t1 = timed(one_pass, 7)
    workload = (f"uniform periodic cube N={N} rho={args.density} r={r}, 27 images, exact pairs (trained on seed "
                f"{train_seed}, applied to the headline configuration, seed {apply_seed})") if uniform else (
        f"clustered periodic cube N={N} rho=0.85 {ncl} Gaussian clusters sigma=5 + 20% uniform, r={r}, 27 images, "
        f"exact pairs")
    return {"workload": workload,
            "pairs": total, "train_steps": 300, "train_batch": 16384, "train_s": train_s, "train_loss": loss,
            "margin_log2": margin, "count_then_fill_ms": t2, "single_pass_ms": t1, "gain": t2 / t1,
            "pairs_per_s_count_then_fill": total / (t2 * 1e-3), "pairs_per_s_single_pass": total / (t1 * 1e-3),
            "overflowed_rows": res.n_overflow, "overflow_fraction": res.n_overflow / N,
            "row_positions_per_pair": res.capacity_total / total}


def batched_leg(args, nb, torch, np):
    """Independent periodic boxes (N per box log-uniform in [256, 65536], rho 0.85, r 2.8, 27 images
    per box, exact pairs).  Batched: one graph launch of forest build + records + count + scan +
    pair fill with autotuned block / bucket / variant.  Loop: the single-tree path box by box
    (build, records, count, scan, host read of the total, fill)."""
    from neighbor_amd.batch import BatchPlan, LBVHBatch
    from neighbor_amd.workloads import batched_boxes, images27, spheres_from_points
    r = 2.8
    nbox = args.batched_boxes
    log(f"batched leg: {nbox} boxes")
    pts, offs, lo, hi = batched_boxes(nbox=nbox, nmin=256, nmax=65536, density=0.85, seed=42)
    N = pts.shape[0]
    d_pts = torch.from_numpy(pts).cuda()
    d_sph = torch.from_numpy(spheres_from_points(pts, r)).cuda()
    d_offs = torch.from_numpy(offs.astype(np.int32)).cuda()
    d_lo, d_hi = torch.from_numpy(lo).cuda(), torch.from_numpy(hi).cuda()

    bt = LBVHBatch()
    t0 = time.perf_counter()
    choice = bt.tune(d_pts, d_offs, d_lo, d_hi, d_sph)
    tune_s = time.perf_counter() - t0
    counts = bt.count(d_sph)
    total = int(counts.to(torch.int64).sum().item())
    plan = BatchPlan(bt, d_pts, d_offs, d_lo, d_hi, d_sph, capacity=total)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    plan.launch()
    reps = 3
    e0.record()
    for _ in range(reps):
        plan.launch()
    e1.record()
    torch.cuda.synchronize()
    batched_ms = e0.elapsed_time(e1) / reps
    assert int(plan.offsets[-1].item()) == total, "batched pair count changed"
    bt.check()      # raises if the forest build gave up a bounded wait
    del plan

    # the single pass: forest build + records + density capacities + their permuted scan + one
    # fill, no count pass (eager launches on the stream, no host read inside the step)
    from neighbor_amd.batch import BatchRows
    rows = BatchRows(N, int(total * 1.8) + 64 * nbox)

    def single_step():
        bt.build(d_pts, d_offs, d_lo, d_hi)
        bt.setup()
        bt.single_pass(d_sph, rows, z=args.batch_z, pad=args.pad)

    single_step()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        single_step()
    e1.record()
    torch.cuda.synchronize()
    single_ms = e0.elapsed_time(e1) / reps
    used, moved = (int(v) for v in rows.spill.tolist())
    assert used <= rows.neigh.numel(), "batched single-pass rows did not fit their buffer"
    assert torch.equal(rows.count, counts), "batched single-pass counts differ from the count pass"
    bt.check()
    single = {"ms": single_ms, "pairs_per_s": total / (single_ms * 1e-3), "gain": batched_ms / single_ms,
              "moved_rows": moved, "row_positions_per_pair": used / max(total, 1), "z": args.batch_z, "pad": args.pad,
              "step": "forest build + bucket records + row capacities from each box's density "
                      "(ceil(lambda + z sqrt(lambda)) + pad) + their scan in forest order + one exact fill, "
                      "overflowed rows moved and walked again inside the fill"}
    del rows
    # per-box pair totals from the batched count (host), so that the loop below needs no host read
    c_host = counts.to(torch.int64).cpu().numpy()
    box_total = np.add.reduceat(c_host, offs[:-1].astype(np.int64)) if N else np.zeros(nbox, np.int64)
    box_total[np.diff(offs) == 0] = 0

    # the single-tree path, box by box, into one pair buffer
    d_img = torch.from_numpy(np.concatenate([images27(float(hi[b, 0]), dtype=np.float64) for b in range(nbox)])).cuda()
    neigh = torch.empty(max(total, 1), dtype=torch.int32, device="cuda")
    img = torch.empty(max(total, 1), dtype=torch.uint8, device="cuda")
    cnt = torch.empty(max(int(np.diff(offs).max()), 1), dtype=torch.int32, device="cuda")
    poff = torch.empty(cnt.numel() + 1, dtype=torch.int64, device="cuda")
    lb = nb.LBVH()
    tr = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
    tr.set_bucket_size(args.bucket)

    def loop(host_reads):
        """host_reads: each box's pair total is read back after its scan (as a caller without
        the totals must); otherwise it comes from box_total and the loop never waits."""
        g0 = 0
        for b in range(nbox):
            s, e = int(offs[b]), int(offs[b + 1])
            n = e - s
            L = float(hi[b, 0])
            p, q, im = d_pts[s:e], d_sph[s:e], d_img[27 * b:27 * b + 27]
            lb.build(p, [0.0, 0.0, 0.0], [L, L, L], block=args.block)
            tr.setup(lb, block=args.block)
            order = lb.device_pointers()["primitive"]
            tr.exact_count(lb, q, p, images=im, order=order, nneigh=cnt[:n], block=args.block)
            nb.exclusive_scan(cnt[:n], poff[:n + 1])
            t = int(poff[n].item()) if host_reads else int(box_total[b])
            tr.exact_pairs(lb, q, p, poff[:n + 1], images=im, order=order, total=t, neigh=neigh[g0:g0 + t],
                           img=img[g0:g0 + t], block=args.block)
            g0 += t
        return g0

    loop_ms = loop_reads_ms = float("nan")
    if not args.no_loop:
        loop(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loop_total = loop(False)
        torch.cuda.synchronize()
        loop_ms = (time.perf_counter() - t0) * 1e3
        assert loop_total == total, f"per-box loop found {loop_total} pairs, batched {total}"
        t0 = time.perf_counter()
        loop_total = loop(True)
        torch.cuda.synchronize()
        loop_reads_ms = (time.perf_counter() - t0) * 1e3
        assert loop_total == total, f"per-box loop found {loop_total} pairs, batched {total}"
        lb.check()
    return {"workload": f"{nbox} periodic boxes, N per box log-uniform in [256, 65536] (seed 42), rho=0.85, "
                        f"r={r}, 27 images per box, exact pairs", "boxes": nbox, "points": N, "pairs": total,
            "batched_ms": batched_ms, "loop_ms": loop_ms, "speedup": loop_ms / batched_ms,
            "loop_host_reads_ms": loop_reads_ms, "speedup_vs_loop_host_reads": loop_reads_ms / batched_ms,
            "pairs_per_s_batched": total / (batched_ms * 1e-3), "pairs_per_s_loop": total / (loop_ms * 1e-3),
            "autotune": choice, "autotune_s": tune_s, "single_pass": single,
            "batched_step": "one graph launch: forest build + bucket records + count + scan + pair fill",
            "loop_step": "per box: build + records + count + scan + pair fill, each box's pair total taken from "
                         "the batched count (no host synchronization inside the loop)",
            "loop_host_reads_step": "the same loop reading each box's pair total back after its scan"}


def colloid_leg(args, nb, torch, np):
    """Colloid mixture (N spheres, 1% of diameter 10 and 99% of diameter 1, volume fraction 0.3,
    random sequential addition without overlap, periodic cube).  One LBVH per type over points,
    cutoffs (d_a + d_b)/2 + 0.4 per type pair, exact pairs (27 images) in per-type Morton-order
    rows; the traversal of each (query type, tree type) pair autotuned once beforehand."""
    from neighbor_amd.typed import TypedNeighborList
    from neighbor_amd.workloads import colloid_mixture
    N = args.colloid_N
    d = [1.0, 10.0]
    R = [[0.5 * (d[a] + d[b]) + 0.4 for b in range(2)] for a in range(2)]
    log(f"colloid leg: N={N}")
    t0 = time.perf_counter()
    pts, types, L = colloid_mixture(N, seed=7)
    gen_s = time.perf_counter() - t0
    d_pts = torch.from_numpy(pts).cuda()
    nl = TypedNeighborList(2)
    nl.build(d_pts, L, types=types)
    t0 = time.perf_counter()
    choice = nl.tune(R)
    tune_s = time.perf_counter() - t0

    def step():
        nl.build(d_pts, L)
        return nl.compute(R)

    res = step()
    total = res["total"]
    per = {f"{a}{b}": int(v.to(torch.int64).sum().item()) for (a, b), v in res["per_pair"].items()}
    torch.cuda.synchronize()
    ts = []
    for _ in range(7):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = step()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
        assert r["total"] == total, "colloid pair count changed"
    ms = float(np.median(ts)) * 1e3
    # the same step over 16-byte (x, y, z, type) records read in place (HOOMD's postype layout):
    # each type's tree reads its records through the type's index list, no per-type copy
    rec = torch.cat([d_pts, torch.from_numpy(types.astype(np.float32)).cuda()[:, None]], 1).contiguous()
    nl4 = TypedNeighborList(2)
    nl4.build(rec, L, types=types)
    nl4.choice = dict(nl.choice)

    def step_records():
        nl4.build(rec, L)
        return nl4.compute(R)

    r4 = step_records()
    assert r4["total"] == total, "records-in-place pair count differs"
    ts4 = []
    for _ in range(7):
        torch.cuda.synchronize()
        t = time.perf_counter()
        step_records()
        torch.cuda.synchronize()
        ts4.append(time.perf_counter() - t)
    ms4 = float(np.median(ts4)) * 1e3
    names = {nb.NB_VARIANT_WAVE: "wave", nb.NB_VARIANT_FLOAT_ROPE_EXACT: "lane float-rope"}
    return {"workload": f"colloid mixture N={N} (1% d=10, 99% d=1, phi=0.3, RSA seed 7), L={L:.2f}, "
                        f"per-type LBVHs, R_ab=(d_a+d_b)/2+0.4, 27 images, exact pairs",
            "pairs": total, "pairs_by_type_pair": per, "step_ms": ms, "pairs_per_s": total / (ms * 1e-3),
            "step_ms_postype_records": ms4,
            "postype_records": "the same step over (x, y, z, type) 16-byte records read in place by the per-type "
                               "trees (index lists) and the exact tests (stride 4): no per-type position copy",
            "generate_s": gen_s, "n_large": int((types == 1).sum()), "tune_s": tune_s,
            "traversal_by_type_pair": {f"{a}{b}": f"{names.get(v, v)} bucket {k}" for (a, b), (v, k) in choice.items()},
            "step": "both type trees built + set up, 4 (query type, tree type) exact counts, 2 permuted scans, "
                    "1 host read of the total, 4 exact row fills"}


def build_scaling_leg(args, nb, torch, np):
    """LBVH build (codes, sort, Karras, bubble), quantized compression and wave-walk setup of
    uniform points in a periodic cube (density 0.85) at each N of --build-sizes, 5 timed
    repetitions each after one untimed; HBM fraction of build + compression at
    BUILD_BYTES_PER_PRIM algorithmic bytes per primitive."""
    from neighbor_amd.workloads import box_length
    out = []
    for N in [int(x) for x in args.build_sizes.split(",") if x]:
        log(f"build scaling N={N}")
        L = box_length(N, args.density)
        g = torch.Generator(device="cuda")
        g.manual_seed(1)
        d = (torch.rand((N, 3), generator=g, device="cuda") * L).contiguous()
        lb = nb.LBVH()
        lb.build(d, [0, 0, 0], [L] * 3)
        tq = nb.LBVHTraverser(variant=nb.NB_VARIANT_QUANTIZED)
        tw = nb.LBVHTraverser(variant=nb.NB_VARIANT_WAVE)
        tw.set_bucket_size(args.bucket)
        tq.setup(lb)
        tw.setup(lb)
        torch.cuda.synchronize()
        ms = {}
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        for name, f in (("build", lambda: lb.build(d, [0, 0, 0], [L] * 3)), ("quantized_compress", lambda: tq.setup(lb)),
                        ("wave_setup", lambda: tw.setup(lb))):
            ev[0].record()
            for _ in range(5):
                f()
            ev[1].record()
            torch.cuda.synchronize()
            ms[name] = ev[0].elapsed_time(ev[1]) / 5
        bc = ms["build"] + ms["quantized_compress"]
        out.append({"N": N, "build_ms": ms["build"], "quantized_compress_ms": ms["quantized_compress"],
                    "wave_setup_ms": ms["wave_setup"], "build_compress_ms": bc,
                    "hbm_frac": BUILD_BYTES_PER_PRIM * N / (bc * 1e-3) / 1e9 / HBM_PEAK_GBS})
        del lb, d, tq, tw
        torch.cuda.empty_cache()
    return {"bytes_per_prim": BUILD_BYTES_PER_PRIM, "peak_GBs": HBM_PEAK_GBS, "sizes": out}


def md_leg(args, nb, torch, np):
    """A molecular-dynamics loop over the headline configuration: N uniform points at rho 0.85,
    cutoff 2.5 + skin 0.3 (lists at 2.8, as the headline), each step every particle moves by
    a random step of up to 0.01 per axis (periodic wrap).  Per step the list is kept while no
    particle moved more than skin/2 since it was computed, else refit (at most 8 times in a
    row) or rebuilt; against rebuilding the list at the bare cutoff every step.  Timed: the
    list updates only (the motion is not)."""
    from neighbor_amd.typed import TypedNeighborList
    from neighbor_amd.workloads import box_length, uniform_points
    N, rho, R, skin, amp = args.N, 0.85, 2.5, 0.3, 0.01
    L = box_length(N, rho)
    steps = args.md_steps
    log(f"md leg: N={N}, {steps} steps")
    pts = torch.from_numpy(uniform_points(N, L, seed=42)).cuda()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)

    def move(p):
        d = (torch.rand(p.shape, generator=gen, device="cuda") * 2.0 - 1.0) * amp
        q = torch.remainder(p + d, L)
        return torch.where(q >= L, torch.zeros_like(q), q).contiguous()

    def run(refit, reuse):
        nl = TypedNeighborList(1)
        p = pts.clone()
        sk = skin if reuse else 0.0
        nl.update(p, L, [[R]], sk, types=np.zeros(N, np.int64))
        torch.cuda.synchronize()
        times, acts = {"kept": [], "refit": [], "built": []}, []
        for _ in range(steps):
            p = move(p)
            torch.cuda.synchronize()
            t = time.perf_counter()
            res = nl.update(p, L, [[R]], sk, refit=refit)
            torch.cuda.synchronize()
            times[res["action"]].append(time.perf_counter() - t)
            acts.append(res["action"])
        total = sum(sum(v) for v in times.values())
        return {"ms_per_step": total / steps * 1e3, "actions": {k: len(v) for k, v in times.items()},
                "median_ms": {k: (float(np.median(v)) * 1e3 if v else None) for k, v in times.items()}}

    reuse = run(refit=True, reuse=True)
    every = run(refit=False, reuse=False)
    return {"workload": f"uniform N={N} rho={rho}, cutoff {R} + skin {skin} (lists at {R + skin}), random steps "
                        f"of up to {amp} per axis, {steps} steps, exact pairs, 27 images",
            "skin": reuse, "rebuild_every_step": every,
            "speedup_vs_every_step": every["ms_per_step"] / reuse["ms_per_step"],
            "note": "a recompute refits only when no particle crossed the box since the last build; "
                    "with periodic wrapping at this N some always have, so every recompute rebuilds"}


def harness_leg(args, nb, torch, np):
    """The reference benchmark's protocol (benchmark/lbvh_benchmark.cc:123-229) on the headline
    configuration: the block size tuned over 32..1024 step 32 (5 samples each, lowest median),
    then the median of 5 runs of K back-to-back calls, for the build and for the count
    traversal (the 27 images listed center first, queries in Morton order).  Traversals: the
    reference's structure (quantized per-thread ropes, SphereQueryOp, float images) and this
    library's exact wave walk (FP64 pair test)."""
    from neighbor_amd.harness import frame_protocol
    from neighbor_amd.workloads import box_length, uniform_points
    N, r, K = args.N, args.r, args.harness_calls
    L = box_length(N, args.density)
    log(f"reference-protocol leg: N={N}, median of 5 x {K} calls")
    pts = uniform_points(N, L, seed=42)
    res = frame_protocol(nb, torch, pts, [L, L, L], r, K, bucket=args.bucket, lo=[0.0, 0.0, 0.0])
    return {"workload": f"uniform N={N} rho={args.density} r={r}, 27 images (center first), queries in Morton order",
            "protocol": f"block tuned over 32..1024 step 32 (5 samples each), median of 5 x {K} calls",
            "build_ms": res["build_ms"], "build_block": res["build_block"],
            "build_hbm_frac": BUILD_BYTES_PER_PRIM * N / (res["build_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traverse_reference_structure_ms": res["traverse_reference_structure_ms"],
            "traverse_reference_structure_block": res["traverse_reference_structure_block"],
            "raw_hits_min_max_mean": res["raw_hits_min_max_mean"],
            "traverse_exact_wave_ms": res["traverse_exact_wave_ms"],
            "exact_hits_min_max_mean": res["exact_hits_min_max_mean"]}


if __name__ == "__main__":
    main()
