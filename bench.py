#!/usr/bin/env python3
"""Throughput benchmark of the MI355X WEMA pipeline (one JSON line on stdout).

A step is one pass of the hot path (DWT -> temporal filter -> amplification -> IDWT ->
guided filter / colour transfer -> RGB24) over one batch of consecutive frames of one
stream; the stream's IIR state carries from step to step. Frames are synthetic
(wema_amd.Synth, generated in HBM before the timed region).

Workloads (the default is the single-stream drop-in case, 1920x1080 colour):
  c2   1920x1080 RGB24 @30, 5-level DWT, band 0.5-3 Hz, amp 120, colour (IIR; the
       bilateral option is inert in IIR mode), 300 frames = (warmup + steps) x 60
  c3   3840x2160 RGB24 @60, 6-level DWT, band 0.8-2.0 Hz, amp 50, guided filter,
       240 frames = (warmup + steps) x 48  (reported under "also" by default)
  c1   320x240 RGB24 @30, 3-level DWT, band 0.5-3 Hz, amp 120, guided filter
  c2w  as c2 in sliding-window mode with bilateral weights (window 32)

roofline: the kernel with the largest share of the timed region; achieved = its
algorithmic bytes per launch / its mean launch duration, from HIP events recorded on the
launch stream between the launches of the timed region. traffic: HBM bytes per launch of
that kernel from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary committed in profiles/
(scaled by the same counters on a copy kernel of known size), or null.
cpu_baseline: the reference pipeline compiled from its sources (oracle/_ref), on this
host's cores, over a bounded sample of the same frames.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mpixel/s & fps (1080p/4K) at % HBM peak; p50/p99 frame latency, mixed streams"

CONFIGS = {
    "c1": dict(preset=1, w=320, h=240, fps=30.0, levels=3, amp=120.0, f_low=0.5, f_high=3.0,
               window=32, color=False, use_iir=True, bilateral=False, batch=32,
               desc="320x240 RGB24@30 disc, L3, IIR 0.5-3 Hz, amp 120, guided filter"),
    "c2": dict(preset=2, w=1920, h=1080, fps=30.0, levels=5, amp=120.0, f_low=0.5, f_high=3.0,
               window=32, color=True, use_iir=True, bilateral=False, batch=60,
               desc="1920x1080 RGB24@30 three discs, L5, IIR 0.5-3 Hz, amp 120, colour"),
    "c2w": dict(preset=2, w=1920, h=1080, fps=30.0, levels=5, amp=120.0, f_low=0.5, f_high=3.0,
                window=32, color=True, use_iir=False, bilateral=True, batch=60,
                desc="1920x1080 RGB24@30, L5, sliding window 32 + bilateral, amp 120, colour"),
    "c3": dict(preset=3, w=3840, h=2160, fps=60.0, levels=6, amp=50.0, f_low=0.8, f_high=2.0,
               window=64, color=False, use_iir=True, bilateral=False, batch=48,
               desc="3840x2160 RGB24@60 checkerboard shear, L6, IIR 0.8-2 Hz, amp 50, guided filter"),
}


def read_traffic(kernel, workload):
    """HBM bytes per launch from profiles/traffic_<workload>.json (written from rocprofv3
    --pmc summaries), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("kernels", {}).get(kernel, {}).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def run_gpu(cfg, name, steps, warmup, batch):
    import torch
    import wema_amd

    n = (steps + warmup) * batch
    w, h = cfg["w"], cfg["h"]
    synth = wema_amd.Synth(cfg["preset"], w, h, cfg["fps"])
    frames = torch.empty((n, h, w, 3), dtype=torch.uint8, device="cuda")
    for t0 in range(0, n, 64):
        synth.frames_device(t0, min(64, n - t0), frames[t0:t0 + 64])
    out = torch.empty_like(frames)
    sess = wema_amd.Session(w, h, cfg["fps"], amp=cfg["amp"], f_low=cfg["f_low"],
                            f_high=cfg["f_high"], window=cfg["window"], levels=cfg["levels"],
                            use_iir=cfg["use_iir"], color=cfg["color"],
                            bilateral=cfg["bilateral"], max_batch=batch)
    torch.cuda.synchronize()
    for k in range(warmup):
        sess.process_device(frames[k * batch:(k + 1) * batch], d_out_u8=out[k * batch:(k + 1) * batch], n=batch)
    torch.cuda.synchronize()
    sess.timing(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    e0.record()
    for k in range(warmup, warmup + steps):
        sess.process_device(frames[k * batch:(k + 1) * batch], d_out_u8=out[k * batch:(k + 1) * batch], n=batch)
    e1.record()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms = e0.elapsed_time(e1)
    kernels = sess.timing_report()
    sess.timing(False)
    frames_timed = steps * batch
    mpix = frames_timed * w * h / 1e6
    res = dict(ms=ms, wall_ms=t_wall * 1e3, frames=frames_timed, mpix_s=mpix / (ms / 1e3),
               fps=frames_timed / (ms / 1e3), kernels=kernels)
    dom = max(kernels, key=lambda k: k["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
    res["roofline"] = dict(bound="hbm", kernel=dom["name"], achieved=round(achieved, 1),
                           peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                           traffic=read_traffic(dom["name"], name),
                           algorithmic_bytes_per_launch=bytes_per_launch,
                           mean_launch_us=round(avg_ms * 1e3, 2))
    tot_bytes = sum(k["bytes"] for k in kernels)
    res["pipeline_gbs"] = round(tot_bytes / (ms / 1e3) / 1e9, 1)
    res["out_sample"] = out[warmup * batch: warmup * batch + 2].cpu().numpy()
    res["first_timed_frame"] = warmup * batch
    return res


def cpu_baseline(cfg, budget_s=20.0, first_frame=0, gpu_sample=None):
    """The compiled reference on this host: colour and window modes run the per-frame path
    (src/main.c:355: colour never batches; single-threaded), luma IIR the OpenMP batch path."""
    from oracle import oracle as O
    import wema_amd

    if not O.ref_available():
        return None
    w, h = cfg["w"], cfg["h"]
    synth = wema_amd.Synth(cfg["preset"], w, h, cfg["fps"])
    ref = O.RefPipe(w, h, cfg["fps"], cfg["amp"], cfg["f_low"], cfg["f_high"], cfg["window"],
                    cfg["levels"], cfg["use_iir"], cfg["color"], True, cfg["bilateral"])
    batch_path = cfg["use_iir"] and not cfg["color"]
    threads = os.cpu_count() if batch_path else 1
    if batch_path:
        threads = min(threads, 16)
        bsz = 8
        first = np.stack([synth.frame(t) for t in range(bsz)])
        t0 = time.perf_counter()
        ref.batch(first, bsz, threads)                      # includes the serial preseed
        t_first = time.perf_counter() - t0
        done, t_sum, t = 0, 0.0, bsz
        while t_sum < budget_s and done < 64:
            fr = np.stack([synth.frame(t + i) for i in range(bsz)])
            t0 = time.perf_counter()
            ref.batch(fr, bsz, threads)
            t_sum += time.perf_counter() - t0
            done += bsz
            t += bsz
        sample = "%d frames (after a first batch of %d incl. preseed, %.1f s), batch path, " \
                 "%d OpenMP threads" % (done, bsz, t_first, threads)
    else:
        f0 = synth.frame(0)
        t0 = time.perf_counter()
        ref.frame(f0)
        t_first = time.perf_counter() - t0
        done, t_sum, t = 0, 0.0, 1
        while t_sum < budget_s and done < 64:
            fr = synth.frame(t)
            t0 = time.perf_counter()
            ref.frame(fr)
            t_sum += time.perf_counter() - t0
            done += 1
            t += 1
        sample = "%d frames after frame 0 (incl. preseed, %.1f s), per-frame path, 1 thread" % (done, t_first)
    mpix = done * w * h / 1e6
    ref.close()
    return dict(value=round(mpix / t_sum, 3), unit="Mpixel/s", cores=threads, kind="reference",
                sample=sample + "; u8<->float conversion included",
                fps=round(done / t_sum, 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--batch", type=int, default=0, help="frames per step (default per config)")
    ap.add_argument("--also", default="c3", help="comma list of extra workloads ('' = none)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (for rocprofv3 counter scaling)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs one stream pipeline per GPU and "
              "shards sessions across GPUs without exchange; --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()

    cfg = CONFIGS[args.config]
    batch = args.batch or cfg["batch"]
    r = run_gpu(cfg, args.config, args.steps, args.warmup, batch)
    line = {
        "metric": METRIC,
        "value": round(r["mpix_s"], 1),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (deterministic generator, frames resident in HBM)",
        "config": {"workload": args.config, "desc": cfg["desc"], "width": cfg["w"],
                   "height": cfg["h"], "frames_per_step": batch, "levels": cfg["levels"]},
        "fps": round(r["fps"], 1),
        "pipeline_algorithmic_GBs": r["pipeline_gbs"],
        "roofline": r["roofline"],
        "kernels": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}
                    for d in r["kernels"]],
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cfg, args.cpu_budget)
    also = {}
    for name in [a for a in args.also.split(",") if a and a != args.config]:
        c = CONFIGS[name]
        rr = run_gpu(c, name, args.steps, args.warmup, c["batch"])
        also[name] = {"desc": c["desc"], "value": round(rr["mpix_s"], 1), "unit": "Mpixel/s",
                      "fps": round(rr["fps"], 1), "ms_per_step": round(rr["ms"] / args.steps, 3),
                      "frames_per_step": c["batch"], "pipeline_algorithmic_GBs": rr["pipeline_gbs"],
                      "roofline": rr["roofline"],
                      "kernels": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}
                                  for d in rr["kernels"]]}
    if also:
        line["also"] = also
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
