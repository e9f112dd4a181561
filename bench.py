#!/usr/bin/env python3
"""Detection throughput of the plain pipeline on one MI355X.

Workload: one seeded synthetic 1920x1080 uint8 frame with 16 three-ring
markers (outer radius 25-200 px, tilt <= 60 deg, 20 % partially occluded,
blur 0.5-2 px, noise 6 DN), default parameters, 3 crowns.  A step is one
drop-in call, cctagx_detect: the frame from host memory to the device, front
end on the GPU, candidate stage on the host, identification on the GPU,
markers back on the host.  The same detection on the frame already resident
in device memory (cctagx_detect_loaded) is reported beside it.

Prints one JSON line with the throughput, the roofline position of the
kernel that takes the most device time per frame (durations from HIP events
recorded around that kernel on the library stream) and the CPU baseline:
the oracle's restatement of the reference CPU path timed on a bounded
sample of the same frame.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3    # MI355X fp32

# Cut walk of select: bilinear 16, step 2, running mean/variance 10.  (The
# grid search's algorithmic FLOPs come from the library: per marker,
# iteration and grid point 25 per cut sample and 3 per sample of every cut
# pair, 3.1 MFLOP for 22 cuts of 100 samples.)
FLOP_PER_SELECT_SAMPLE = 16 + 2 + 10


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def counter_traffic(path: str) -> dict:
    """Per-kernel HBM bytes per launch from a committed rocprofv3 counter
    summary (lines `kernel_bytes: <kernel> <fetch> <write>`), or {}."""
    out = {}
    try:
        for line in open(os.path.join(ROOT, path)):
            if line.startswith("kernel_bytes:"):
                _, k, f, w = line.split()
                out[k] = (int(f), int(w))
    except OSError:
        pass
    return out


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
            "pyramid": N + (4 * N if l == 0 else 0),       # level 0 read + its quads written, coarser levels written
            "filter": N + 4 * N,                           # u8 in, dx/dy int16 out
            "nms_local": 4 * N + N + 4 * N + N,            # dx,dy in; cls, labels, root flag out
            "cc_border": 0,                                # tile edges only
            "cc_flatten": N + 4 * N + 4 * N + N,           # cls, labels in/out, root flag
            "thin": N + 4 * N + N + N,                     # cls, labels, roots in; thinned map out
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


def batched_comparison(frames, Detector, BatchDetector, reps=3):
    """Uneven-size batch: the plain path frame by frame (one handle per
    size) against one batched detection; both include the frames' uploads."""
    if not frames:
        return None
    import collections
    imgs = [f[0] for f in frames]
    sizes = [(im.shape[1], im.shape[0]) for im in imgs]
    dets = {sz: Detector(*sz) for sz in set(sizes)}
    ref = [dets[sz].detect(im) for sz, im in zip(sizes, imgs)]   # warm-up and reference
    t0 = time.perf_counter()
    for _ in range(reps):
        for sz, im in zip(sizes, imgs):
            dets[sz].detect(im)
    plain_s = (time.perf_counter() - t0) / reps
    for d in dets.values():
        d.close()
    b = BatchDetector(sizes)
    for f, im in enumerate(imgs):
        b.load(f, im)
    got = b.detect()                                             # autotune + graph capture
    st0 = b.stats()
    t0 = time.perf_counter()
    for _ in range(reps):
        for f, im in enumerate(imgs):
            b.load(f, im)
        got = b.detect()
    batch_s = (time.perf_counter() - t0) / reps
    st = b.stats()
    b.close()
    n = len(imgs)
    return {"workload": "%d frames of uneven sizes (640x480 .. 3840x2160), 1-32 markers and "
                        "0-100 distractors each" % n,
            "sizes": {"%dx%d" % k: v for k, v in sorted(collections.Counter(sizes).items())},
            "markers": sum(len(r) for r in ref),
            "plain_frames_per_s": n / plain_s, "batched_frames_per_s": n / batch_s,
            "speedup": plain_s / batch_s, "identical_to_plain": got == ref,
            "partition": {"filter_tile_h": [16, 32, 64][st["filter_variant"]],
                          "edge_blocks": st["edge_blocks"], "tune_ms": st0["tune_us"] / 1e3},
            "last_batch_us": {k: round(st[k], 1) for k in ("front_end_us", "stage_us",
                                                            "candidates_us", "ident_us")}}


def measured_traffic(path: str):
    """HBM bytes per detection of the refinement sequence from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (calibrated against a copy), or
    None when the summary is absent."""
    try:
        for line in open(os.path.join(ROOT, path)):
            if line.startswith("traffic_bytes_per_detection:"):
                return int(line.split(":")[1])
    except OSError:
        return None
    return None


def pipes_comparison(img, Detector, ref, configs=((2, 8), (4, 4)), steps=40):
    """Several pipes detecting at once from their own host threads (the
    reference's pipeId: one pipe per camera or stream), the frame resident in
    each; whole-job frames/s and whether every detection equals `ref`."""
    import threading
    h, w = img.shape
    out = []
    for pipes, threads in configs:
        dets = [Detector(w, h) for _ in range(pipes)]
        for d in dets:
            d.set_host_threads(threads)
            d.load(img)
            d.detect_loaded()
        results = [[] for _ in dets]

        def run(d, res):
            for _ in range(steps):
                res.append(d.detect_loaded())
        ths = [threading.Thread(target=run, args=(d, r)) for d, r in zip(dets, results)]
        t0 = time.perf_counter()
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        dt = time.perf_counter() - t0
        for d in dets:
            d.close()
        out.append({"pipes": pipes, "host_threads_per_pipe": threads,
                    "frames_per_s": pipes * steps / dt,
                    "identical": all(x == ref for r in results for x in r)})
    return out


def gt_centre_errors(markers, gts):
    """Distance of every identified marker to the nearest ground-truth centre,
    and how many carry that marker's ID."""
    errs, ids_ok = [], 0
    for m in markers:
        if m["status"] != 1:
            continue
        d, g = min((math.hypot(m["x"] - g.center[0], m["y"] - g.center[1]), g) for g in gts)
        errs.append(d)
        ids_ok += int(g.id == m["id"])
    return errs, ids_ok


def other_configs(Detector, synth, steps=10):
    """The plain drop-in call on the other single-frame configurations: the
    1280x720 frame with 4 markers and the 3840x2160 frame with 32 markers and
    200 distractors; frames/s with the upload, and the markers against the
    synthetic ground truth."""
    out = {}
    for cfg, name in ((1, "1280x720, 4 markers"), (3, "3840x2160, 32 markers, 200 distractors")):
        img, gts = synth.config_frame(cfg, 0)
        h, w = img.shape
        det = Detector(w, h)
        for _ in range(2):
            ref = det.detect(img)
        t0 = time.perf_counter()
        for _ in range(steps):
            got = det.detect(img)
        dt = (time.perf_counter() - t0) / steps
        det.close()
        errs, ids_ok = gt_centre_errors(got, gts)
        out[name] = {"frames_per_s": 1.0 / dt, "ms_per_frame": 1e3 * dt,
                     "identical_across_steps": got == ref, "identified": len(errs),
                     "ids_correct": ids_ok, "gt_markers": len(gts),
                     "center_err_px_mean": (sum(errs) / len(errs)) if errs else None}
    return out


def gradient_comparison(frames, Detector, BatchDetector, steps, reps=3):
    """1920x1080 frames with 16 markers: the plain path frame by frame
    against the batched path, with the reference grid search and with the
    gradient refinement (its first grid iteration, then `steps` descent
    steps); frames/s with the frames' uploads and centre error against the
    synthetic ground truth."""
    if not frames:
        return None
    imgs = [f[0] for f in frames]
    gts = [f[1] for f in frames]
    h, w = imgs[0].shape
    det = Detector(w, h)
    plain = [det.detect(im) for im in imgs]
    t0 = time.perf_counter()
    for _ in range(reps):
        for im in imgs:
            det.detect(im)
    plain_s = (time.perf_counter() - t0) / reps
    det.close()
    out = {"workload": "%d frames 1920x1080, 16 three-ring markers each" % len(imgs),
           "gradient_steps": steps}

    def run_batch(gradient_steps):
        b = BatchDetector([(w, h)] * len(imgs))
        b.set_refinement(gradient_steps)
        for f, im in enumerate(imgs):
            b.load(f, im)
        got = b.detect()
        t0 = time.perf_counter()
        for _ in range(reps):
            for f, im in enumerate(imgs):
                b.load(f, im)
            got = b.detect()
        dt = (time.perf_counter() - t0) / reps
        refine_us = b.stats()["refine_us"]
        b.close()
        return got, dt, refine_us

    def accuracy(res):
        errs, ids = [], 0
        for markers, g in zip(res, gts):
            e, ok = gt_centre_errors(markers, g)
            errs += e
            ids += ok
        return {"identified": len(errs), "ids_correct": ids,
                "center_err_px_mean": float(np.mean(errs)) if errs else None,
                "center_err_px_max": float(np.max(errs)) if errs else None}

    grid, grid_s, grid_us = run_batch(0)
    grad, grad_s, grad_us = run_batch(steps)
    n = len(imgs)
    out.update({
        "plain_frames_per_s": n / plain_s, "batched_frames_per_s": n / grid_s,
        "batched_gradient_frames_per_s": n / grad_s, "speedup_vs_plain": plain_s / grad_s,
        "refine_us_per_batch": {"grid": grid_us, "gradient": grad_us},
        "batched_identical_to_plain": grid == plain,
        "same_ids_as_plain": all([m["id"] for m in a] == [m["id"] for m in b]
                                 for a, b in zip(grad, plain)),
        "accuracy": {"plain": accuracy(plain), "batched_gradient": accuracy(grad)}})
    return out


def refine_comparison(mosaic, Refiner, steps=20, reps=3):
    """Centre refinement of many candidate markers: batched descent on the
    differentiable score against the reference grid search; centre error
    against the synthetic ground truth."""
    if mosaic is None:
        return None
    import numpy as np
    img, ms = mosaic
    r = Refiner(len(ms))
    r.set_image(img)
    r.set_markers(ms)
    r.descend(steps)
    r.refine(steps)
    r.grid()
    d_us, f_us, g_us = [], [], []
    for _ in range(reps):
        c_desc, _ = r.descend(steps)
        d_us.append(r.times()["descend_us"])
        c_ref, _ = r.refine(steps)
        f_us.append(r.times()["descend_us"])
        c_grid, ok = r.grid()
        g_us.append(r.times()["grid_us"])
    r.close()
    gt = np.array([m["center_gt"] for m in ms])
    c0 = np.array([m["center0"] for m in ms])

    def err(c, sel=slice(None)):
        e = np.hypot(*(np.asarray(c, np.float64)[sel] - gt[sel]).T)
        return {"mean": float(e.mean()), "median": float(np.median(e)),
                "p95": float(np.percentile(e, 95))}
    return {"workload": "%d markers, %d cuts x 100 samples, fp32 sampling" % (len(ms), len(ms[0]["stops"])),
            "steps": steps,
            "refine": "the grid search's first iteration (5 x 5 points around the start), "
                      "then %d descent steps from the point it picked" % steps,
            "refine_ms": min(f_us) / 1e3, "descend_ms": min(d_us) / 1e3, "grid_ms": min(g_us) / 1e3,
            "speedup": min(g_us) / min(f_us), "grid_ok": int(ok.sum()),
            "center_err_px": {"initial": err(c0), "refine": err(c_ref), "descend": err(c_desc),
                              "grid": err(c_grid, ok)}}


def progress(msg: str) -> None:
    """One line per finished stage on stderr (stdout carries only the JSON)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="bound of the CPU baseline sample (whole frames, at least one)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-pipes", action="store_true",
                    help="skip the concurrent-pipes comparison")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="skip the 1280x720 and 3840x2160 plain-path rates")
    ap.add_argument("--batch-frames", type=int, default=64,
                    help="frames of the uneven-size batch (0 skips the batched comparison)")
    ap.add_argument("--refine-markers", type=int, default=4096,
                    help="markers of the refinement comparison (0 skips it)")
    ap.add_argument("--gradient-frames", type=int, default=32,
                    help="1920x1080 frames of the plain / batched / batched+gradient "
                         "comparison (0 skips it)")
    ap.add_argument("--gradient-steps", type=int, default=30,
                    help="descent steps of the batched+gradient detection (after the grid "
                         "search's first iteration)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the detection path has no multi-GPU form here "
              "(--gpus %d, WORLD_SIZE %d requested)" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    # synthetic inputs of the batched and refinement comparisons, made in
    # worker processes before this process touches the GPU
    from cctag_amd import synth
    side = {}
    if args.batch_frames > 0 or args.refine_markers > 0 or args.gradient_frames > 0:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(max(1, min(16, os.cpu_count() or 1))) as pool:
            mosaic = (pool.apply_async(synth.refine_mosaic, (args.refine_markers,))
                      if args.refine_markers > 0 else None)
            side["batch"] = pool.map(synth.batch_frame, range(args.batch_frames))
            side["hd"] = pool.map(synth.frame_1080p, range(args.gradient_frames))
            side["mosaic"] = mosaic.get() if mosaic else None
        progress("synthetic inputs ready")

    import torch
    from cctag_amd.detector import BatchDetector, Detector, Refiner

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    img, gts = synth.config_frame(2, 0)
    h, w = img.shape
    det = Detector(w, h)
    host_threads = det.host_threads()
    markers = []
    for _ in range(args.warmup):
        markers = det.detect(img)

    # the timed steps: the drop-in call with the frame in host memory
    K = args.steps
    step_markers = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        step_markers.append(det.detect(img))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    fps = K / elapsed
    progress("timed steps done: %.1f frames/s" % fps)
    markers = step_markers[-1]
    assert all(m == markers for m in step_markers), "detection differs between steps"

    # the same detection on the frame resident in device memory
    det.load(img)
    det.detect_loaded()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        det.detect_loaded()
    torch.cuda.synchronize()
    fps_resident = K / (time.perf_counter() - t0)

    # per-kernel device times: a second pass of K steps, reading the HIP
    # event times the library records on its stream after each step
    kern = {}
    sums = {k: 0.0 for k in ("front_end_us", "stage_us", "host_us", "load_us", "loop_one_us",
                             "loop_two_us", "labels_us", "loop_three_us", "rescale_us",
                             "ident_us", "finish_us", "ident_prep_us", "ident_select_us",
                             "ident_refine_us", "ident_post_us")}
    work = {"select_samples": 0, "refine_samples": 0, "refine_flops": 0}
    det.set_stage_timing(True)
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
    det.set_stage_timing(False)

    pipes = None if args.no_pipes else pipes_comparison(img, Detector, markers)
    progress("per-kernel times and concurrent pipes done")

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
        flops = work["refine_flops"] / K
        achieved = flops / (avg[dom] * 1e-6) / 1e12
        roof = {"kernel": "k_refine", "bound": "fp32-valu", "achieved": achieved,
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                "traffic": measured_traffic("profiles/refine_1080p16_pmc.md"),
                "traffic_source": "profiles/refine_1080p16_pmc.md",
                "avg_us": avg[dom], "algorithmic_per_launch": flops}
    elif dom == "select_cuts":
        flops = work["select_samples"] / K * FLOP_PER_SELECT_SAMPLE
        achieved = flops / (avg[dom] * 1e-6) / 1e12
        roof = {"kernel": "k_select_cuts", "bound": "fp32-valu", "achieved": achieved,
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                "traffic": None, "avg_us": avg[dom], "algorithmic_per_launch": flops}
    else:
        b = frontend_bytes(dom, lv)
        achieved = b / (avg[dom] * 1e-6) / 1e9
        roof = {"kernel": "k_" + dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "avg_us": avg[dom], "algorithmic_per_launch": b}
    kernels = {}
    measured = counter_traffic("profiles/frontend_1080p16_pmc.md")
    for n, t in sorted(avg.items(), key=lambda kv: -kv[1]):
        e = {"avg_us": round(t, 2)}
        if n not in ("refine", "select_cuts"):
            e["GB/s"] = round(frontend_bytes(n, lv) / (t * 1e-6) / 1e9, 1) if t > 0 else None
            if "k_" + n in measured and t > 0:
                f, wr = measured["k_" + n]
                e["counter_GB/s"] = round((f + wr) / (t * 1e-6) / 1e9, 1)
                e["counter_bytes"] = f + wr
        kernels["k_" + n] = e

    cpu = None
    if not args.no_cpu_baseline:
        from tests import oracle_lib
        # the host's CPU share of one GPU (16 cores), as the reference runs
        # its candidate loops under tbb::parallel_for; one thread beside it
        threads = max(1, min(16, os.cpu_count() or 1))
        t_serial = oracle_lib.time_detect(img, 1, threads=1)
        t_one = oracle_lib.time_detect(img, 1, threads=threads)
        iters = max(1, int((args.cpu_seconds - t_serial) / t_one) - 1)
        t_avg = oracle_lib.time_detect(img, iters, threads=threads)
        cpu = {"value": 1.0 / t_avg, "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": "%d+1 full detections (front end, candidates, identification) of the "
                         "same 1920x1080 frame on %d threads (front end serial; candidate "
                         "stage and the identification of the candidates parallel)"
                         % (iters, threads),
               "single_thread_value": 1.0 / t_serial,
               "cpu_model": cpu_model(), "nproc": os.cpu_count()}
        progress("CPU baseline done: %.2f frames/s on %d threads" % (cpu["value"], threads))

    batched = batched_comparison(side.get("batch") or [], Detector, BatchDetector)
    progress("batched comparison done")
    refine = refine_comparison(side.get("mosaic"), Refiner)
    progress("refinement comparison done")
    gradient = gradient_comparison(side.get("hd") or [], Detector, BatchDetector,
                                   args.gradient_steps)
    progress("batched gradient comparison done")
    others = None if args.no_other_configs else other_configs(Detector, synth)
    progress("other configurations done")

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
                               "4-level pyramid, default parameters, frame uploaded each step",
                   "width": w, "height": h, "markers": len(gts), "pipes": 1,
                   "host_threads": host_threads},
        "roofline": roof,
        "cpu_baseline": cpu,
        "vs_cpu": (fps / cpu["value"]) if cpu else None,
        "resident_frames_per_s": fps_resident,
        "accuracy": {"identified": len(errs), "ids_correct": ids_ok, "gt_markers": len(gts),
                     "center_err_px_mean": (sum(errs) / len(errs)) if errs else None,
                     "center_err_px_max": max(errs) if errs else None},
        "breakdown_us": {k[:-3]: round(v / K, 1) for k, v in sums.items()},
        "kernels": kernels,
        "batched": batched,
        "grad_refine": refine,
        "batched_gradient": gradient,
        "concurrent_pipes": pipes,
        "other_configs": others,
    }
    det.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
