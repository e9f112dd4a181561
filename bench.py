#!/usr/bin/env python3
"""Detection throughput of the plain pipeline on one MI355X.

Workload: one seeded synthetic 1920x1080 uint8 frame with 16 three-ring
markers (outer radius 25-200 px, tilt <= 60 deg, 20 % partially occluded,
blur 0.5-2 px, noise 6 DN), default parameters, 3 crowns.  The frame is
loaded into device memory once (cctagx_load); a step is one full
cctagx_detect_loaded: front end on the GPU, candidate stage on the host,
identification on the GPU, markers back on the host.

Prints one JSON line with the throughput, the roofline position of the
kernel that takes the most device time per frame (durations from HIP events
recorded around that kernel on the library stream) and the CPU baseline: the
oracle's single-threaded restatement of the reference CPU path timed on a
bounded sample of the same frame.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3    # MI355X fp32

# Floating-point operations per bilinear cut sample of the identification
# kernels: homography (6 mul + 6 add + 2 div), bilinear (4 sub, 8 mul,
# 3 add, 1 div), step (2 add), per-sample cost terms (3 fp64 per cut).
FLOP_PER_REFINE_SAMPLE = 14 + 16 + 2 + 3
# Cut walk of select: bilinear 16, step 2, running mean/variance 10.
FLOP_PER_SELECT_SAMPLE = 16 + 2 + 10


def level_counts(det, n_levels):
    out = []
    for l in range(n_levels):
        v = det.level(l)
        out.append({"N": v["width"] * v["height"], "E": v["n_edges"],
                    "V": int(v["voters_index"][-1]), "S": len(v["seeds"])})
    return out


def frontend_bytes(name, lv):
    """Algorithmic HBM bytes of one front-end kernel over all levels: what
    its contract must read and write once, from the per-level pixel (N),
    edge point (E), vote (V) and seed (S) counts."""
    tot = 0
    for l, c in enumerate(lv):
        N, E, V, S = c["N"], c["E"], c["V"], c["S"]
        per = {
            "pyramid": N,                                  # level 0 read, coarser levels written
            "filter": N + 4 * N,                           # u8 in, dx/dy int16 out
            "nms": 4 * N + N + 4 * N + N,                  # dx,dy in; cls, lab, root flag out
            "cc_merge": N + 4 * N,                         # cls, labels
            "cc_flatten": N + 4 * N + 4 * N + N,           # cls, labels in/out, root flag
            "thin1": N + 4 * N + N + 2 * N,                # cls, labels, roots; map + pass-1 out
            "thin2": 2 * N,                                # pass-1 in, thinned out
            "compact_count": N,                            # thinned map
            "compact_scan": 0,
            "compact_write": N + 4 * E + 4 * N + 12 * E + 4 * E,  # map, grads at edges; emap, list, flags
            "descent": 8 * E + 8 * E,                      # list in, before/after out
            "vote": 16 * E + 12 * E,                       # list + links in; chosen, dist, counts out
            "voters_scan": 4 * E + 8 * E,
            "voters_fill": 4 * E + 4 * V,
            "voters_finish": 4 * E + 8 * V + 12 * E,
            "seeds": 4 * E + 4 * S,
        }[name]
        tot += per
    return tot


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="bound of the CPU baseline sample (whole frames, at least one)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the detection path has no multi-GPU form here "
              "(--gpus %d, WORLD_SIZE %d requested)" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    import torch
    from cctag_amd import synth
    from cctag_amd.detector import Detector

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    img, gts = synth.config_frame(2, 0)
    h, w = img.shape
    det = Detector(w, h)
    host_threads = det.host_threads()
    det.load(img)
    markers = []
    for _ in range(args.warmup):
        markers = det.detect_loaded()

    K = args.steps
    step_markers = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        step_markers.append(det.detect_loaded())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    fps = K / elapsed
    markers = step_markers[-1]
    assert all(m == markers for m in step_markers), "detection differs between steps"

    # per-kernel device times: a second pass of K steps, reading the HIP
    # event times the library records on its stream after each step
    kern = {}
    sums = {k: 0.0 for k in ("front_end_us", "stage_us", "host_us", "load_us", "loop_one_us",
                             "loop_two_us", "labels_us", "loop_three_us", "rescale_us",
                             "ident_us", "finish_us", "ident_prep_us", "ident_select_us",
                             "ident_refine_us", "ident_post_us")}
    work = {"select_samples": 0, "refine_samples": 0}
    for _ in range(K):
        det.detect_loaded()
        ds = det.detect_stats()
        for n, us in det.stage_times():
            kern.setdefault(n, []).append(us)
        kern.setdefault("select_cuts", []).append(ds["select_us"])
        kern.setdefault("refine", []).append(ds["refine_us"])
        for k in sums:
            sums[k] += ds[k]
        for k in work:
            work[k] += ds[k]

    # accuracy against the synthetic ground truth
    errs, ids_ok = [], 0
    for m in markers:
        if m["status"] != 1:
            continue
        d, g = min((math.hypot(m["x"] - g.center[0], m["y"] - g.center[1]), g) for g in gts)
        errs.append(d)
        ids_ok += int(g.id == m["id"])

    # dominant kernel and its roofline
    avg = {n: sum(v) / len(v) for n, v in kern.items()}
    dom = max(avg, key=avg.get)
    lv = level_counts(det, 4)
    if dom == "refine":
        flops = work["refine_samples"] / K * FLOP_PER_REFINE_SAMPLE
        achieved = flops / (avg[dom] * 1e-6) / 1e12
        roof = {"kernel": "k_refine", "bound": "mfma", "achieved": achieved,
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                "traffic": None, "avg_us": avg[dom], "algorithmic_per_launch": flops}
    elif dom == "select_cuts":
        flops = work["select_samples"] / K * FLOP_PER_SELECT_SAMPLE
        achieved = flops / (avg[dom] * 1e-6) / 1e12
        roof = {"kernel": "k_select_cuts", "bound": "mfma", "achieved": achieved,
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                "traffic": None, "avg_us": avg[dom], "algorithmic_per_launch": flops}
    else:
        b = frontend_bytes(dom, lv)
        achieved = b / (avg[dom] * 1e-6) / 1e9
        roof = {"kernel": "k_" + dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "avg_us": avg[dom], "algorithmic_per_launch": b}
    kernels = {}
    for n, t in sorted(avg.items(), key=lambda kv: -kv[1]):
        e = {"avg_us": round(t, 2)}
        if n not in ("refine", "select_cuts"):
            e["GB/s"] = round(frontend_bytes(n, lv) / (t * 1e-6) / 1e9, 1) if t > 0 else None
        kernels["k_" + n] = e

    cpu = None
    if not args.no_cpu_baseline:
        from tests import oracle_lib
        t_one = oracle_lib.time_detect(img, 1)
        iters = max(1, int(args.cpu_seconds / t_one) - 1)
        t_avg = oracle_lib.time_detect(img, iters) if iters > 0 else t_one
        cpu = {"value": 1.0 / t_avg, "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": "%d+1 full detections (front end, candidates, identification) of the "
                         "same 1920x1080 frame, one thread" % iters}

    line = {
        "metric": "1080p frames/s (16 markers), plain pipeline",
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "plain 1920x1080 uint8 frame, 16 three-ring markers, "
                               "4-level pyramid, default parameters, frame resident in HBM",
                   "width": w, "height": h, "markers": len(gts), "pipes": 1,
                   "host_threads": host_threads},
        "roofline": roof,
        "cpu_baseline": cpu,
        "vs_cpu": (fps / cpu["value"]) if cpu else None,
        "accuracy": {"identified": len(errs), "ids_correct": ids_ok, "gt_markers": len(gts),
                     "center_err_px_mean": (sum(errs) / len(errs)) if errs else None,
                     "center_err_px_max": max(errs) if errs else None},
        "breakdown_us": {k[:-3]: round(v / K, 1) for k, v in sums.items()},
        "kernels": kernels,
    }
    det.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
