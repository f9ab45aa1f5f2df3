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
  c1   320x240 RGB24 @30, 3-level DWT, band 0.5-3 Hz, amp 120, guided filter, 128 frames
       (the clip) per step; also reported
  c2w  as c2 in sliding-window mode with bilateral weights (window 32); also reported
  multistream  24 concurrent streams of mixed resolution, batched ticks vs one at a time
  serving      Poisson sessions at real-time pace: p50/p99 latency, capacity at p99 <= 33 ms

roofline: the kernel with the largest share of the timed region; achieved = its
algorithmic bytes per launch / its mean launch duration, from HIP events recorded on the
launch stream between the launches of a second pass over the timed steps (the timed pass
itself runs without them; ms_per_step_with_launch_markers is the second pass's time
per step). traffic: HBM bytes per launch of
that kernel from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary committed in profiles/
(scaled by the same counters on a copy kernel of known size), or null.
cpu_baseline: the reference pipeline compiled from its sources (oracle/_ref), on this
host's cores, over a bounded sample of the same frames.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mpixel/s & fps (1080p/4K) at % HBM peak; p50/p99 frame latency, mixed streams"

CONFIGS = {
    "c1": dict(preset=1, w=320, h=240, fps=30.0, levels=3, amp=120.0, f_low=0.5, f_high=3.0,
               window=32, color=False, use_iir=True, bilateral=False, batch=128,
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
# The same workloads one frame per call (the path of wema_process_frame, src/wema.c:746-952,
# and of the serving loop): the IIR state is read and written for every frame.
CONFIGS["c2_n1"] = dict(CONFIGS["c2"], batch=1, desc=CONFIGS["c2"]["desc"] + ", one frame per call")
CONFIGS["c3_n1"] = dict(CONFIGS["c3"], batch=1, desc=CONFIGS["c3"]["desc"] + ", one frame per call")
# steps x frames of the per-frame legs (the c3 clip: 240 frames; c2: 300)
PER_FRAME_STEPS = {"c2_n1": 300, "c3_n1": 240}


def log(msg):
    """A progress line on stderr as each measurement ends (stdout carries only the result)."""
    print("bench: %s" % msg, file=sys.stderr, flush=True)


def read_traffic(kernel, workload):
    """HBM bytes per launch from profiles/traffic_<workload>.json (written from rocprofv3
    --pmc summaries), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    try:
        with open(path) as f:
            ks = json.load(f).get("kernels", {})
    except (OSError, ValueError):
        return None
    names = [n for n in ks if n in _kernel_names(kernel)]
    return ks[names[0]].get("hbm_bytes_per_launch") if len(names) == 1 else None


def _kernel_names(kernel):
    """rocprofv3 names of the kernels timed here under one name: the sliding window's
    variants (k_win_r32 ...) as k_win, the one-frame IIR launches (k_iir1) as k_iir."""
    if kernel == "k_win":
        return ("k_win", "k_win_r16", "k_win_r32", "k_win_r64", "k_win_l32", "k_win_l64")
    if kernel == "k_iir":
        return ("k_iir", "k_iir1")
    return (kernel,)


def read_valu(kernel, workload):
    """VALU issue fraction of a kernel from profiles/valu_<workload>.json (rocprofv3 SQ
    counters of the same workload, wema_amd/profile_summary.py), or None."""
    path = os.path.join(ROOT, "profiles", "valu_%s.json" % workload)
    try:
        with open(path) as f:
            ks = json.load(f)["kernels"]
    except (OSError, KeyError, ValueError):
        return None
    names = [n for n in ks if n in _kernel_names(kernel)]
    return ks[names[0]]["valu_frac"] if len(names) == 1 else None


def run_gpu(cfg, name, steps, warmup, batch, gf_seg=0):
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
    if gf_seg:
        sess.set_tile_rows(gf_seg)
    torch.cuda.synchronize()
    for k in range(warmup):
        sess.process_device(frames[k * batch:(k + 1) * batch], d_out_u8=out[k * batch:(k + 1) * batch], n=batch)
    torch.cuda.synchronize()
    # the timed steps run without per-launch event markers (an event recorded between two
    # launches costs a few microseconds of stream time); a second pass over the same frames,
    # with a marker after every launch, gives each kernel's mean launch duration
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
    out_sample = out[warmup * batch: warmup * batch + 2].cpu().numpy()
    sess.timing(True)
    e0.record()
    for k in range(warmup, warmup + steps):
        sess.process_device(frames[k * batch:(k + 1) * batch], d_out_u8=out[k * batch:(k + 1) * batch], n=batch)
    e1.record()
    torch.cuda.synchronize()
    ms_markers = e0.elapsed_time(e1)
    kernels = sess.timing_report()
    sess.timing(False)
    seg = sess.tile_rows
    frames_timed = steps * batch
    mpix = frames_timed * w * h / 1e6
    res = dict(ms=ms, wall_ms=t_wall * 1e3, frames=frames_timed, mpix_s=mpix / (ms / 1e3),
               fps=frames_timed / (ms / 1e3), kernels=kernels, gf_seg=seg, ms_markers=ms_markers)
    dom = max(kernels, key=lambda k: k["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
    res["roofline"] = dict(bound="hbm", kernel=dom["name"], achieved=round(achieved, 1),
                           peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                           traffic=read_traffic(dom["name"], name),
                           valu_frac=read_valu(dom["name"], name),
                           algorithmic_bytes_per_launch=bytes_per_launch,
                           mean_launch_us=round(avg_ms * 1e3, 2))
    tot_bytes = sum(k["bytes"] for k in kernels)
    res["pipeline_gbs"] = round(tot_bytes / (ms / 1e3) / 1e9, 1)
    res["out_sample"] = out_sample
    res["first_timed_frame"] = warmup * batch
    return res


def run_host_pipeline(cfg, steps, batch):
    """The same workload from pinned host memory: each step's frames are uploaded, processed
    and downloaded with the asynchronous pipeline (two device slots, copies on their own
    streams), so the rate includes the PCIe transfers both ways."""
    import wema_amd
    w, h = cfg["w"], cfg["h"]
    synth = wema_amd.Synth(cfg["preset"], w, h, cfg["fps"])
    sess = wema_amd.Session(w, h, cfg["fps"], amp=cfg["amp"], f_low=cfg["f_low"],
                            f_high=cfg["f_high"], window=cfg["window"], levels=cfg["levels"],
                            use_iir=cfg["use_iir"], color=cfg["color"],
                            bilateral=cfg["bilateral"], max_batch=batch)
    nbuf = 3   # host buffers: a third batch is queued while two are on the device
    pin_in = [wema_amd.PinnedArray((batch, h, w, 3)) for _ in range(nbuf)]
    pin_out = [wema_amd.PinnedArray((batch, h, w, 3)) for _ in range(nbuf)]
    for k in range(nbuf):
        for i in range(batch):
            pin_in[k].array[i] = synth.frame(k * batch + i)
    # warm-up batch, then `steps` timed batches, rotating through the host buffers
    sess.wait(sess.submit_host(pin_in[0].array, pin_out[0].array, batch))
    t0 = time.perf_counter()
    tickets = []
    for k in range(steps):
        if len(tickets) == nbuf:
            sess.wait(tickets.pop(0))
        tickets.append(sess.submit_host(pin_in[k % nbuf].array, pin_out[k % nbuf].array, batch))
    for t in tickets:
        sess.wait(t)
    dt = time.perf_counter() - t0
    sess.close()
    for a in pin_in + pin_out:
        a.close()
    frames = steps * batch
    return {"desc": "the same frames from pinned host memory through the asynchronous pipeline "
                    "(RGB24 up and down over PCIe, overlapped with processing)",
            "value": round(frames * w * h / 1e6 / dt, 1), "unit": "Mpixel/s",
            "fps": round(frames / dt, 1), "steps": steps, "frames_per_step": batch,
            "link_GBs": round(2 * frames * w * h * 3 / dt / 1e9, 1)}


def cpu_info():
    """lscpu facts of this host, and the cores this process may use."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    def num(k):
        try:
            return int(info.get(k, ""))
        except ValueError:
            return None
    sockets, cps, tpc = num("Socket(s)"), num("Core(s) per socket"), num("Thread(s) per core")
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    physical = sockets * cps if sockets and cps else None
    return {"model": info.get("Model name"), "sockets": sockets, "physical_cores": physical,
            "threads_per_core": tpc, "smt": (tpc or 1) > 1, "cpus_available": avail}


def cpu_threads(ci):
    """All physical cores this process may use (the box gives one GPU's process 16 CPUs)."""
    n = min(x for x in (ci["physical_cores"], ci["cpus_available"], 16) if x)
    if ci["threads_per_core"] and ci["threads_per_core"] > 1:
        n = min(n, max(1, ci["cpus_available"] // ci["threads_per_core"]))
    return n


def cpu_baseline(cfg, budget_s=20.0, batch_size=64):
    """The compiled reference on this host: colour and window modes run the per-frame path
    (src/main.c:355: colour never batches; single-threaded), luma IIR the OpenMP batch path
    (src/wema_batch.c:352-457) with its default batch of 64 frames (src/wema.h:218) on all
    physical cores available, OMP_PROC_BIND=close, OMP_PLACES=cores. The first batch, which
    runs the serial preseed (src/iir_filter.c:376-471), is reported apart from the steady
    state."""
    from oracle import oracle as O
    import wema_amd

    if not O.ref_available():
        return None
    ci = cpu_info()
    w, h = cfg["w"], cfg["h"]
    synth = wema_amd.Synth(cfg["preset"], w, h, cfg["fps"])
    ref = O.RefPipe(w, h, cfg["fps"], cfg["amp"], cfg["f_low"], cfg["f_high"], cfg["window"],
                    cfg["levels"], cfg["use_iir"], cfg["color"], True, cfg["bilateral"])
    batch_path = cfg["use_iir"] and not cfg["color"]
    threads = cpu_threads(ci) if batch_path else 1
    if batch_path:
        bsz = batch_size
        first = np.stack([synth.frame(t) for t in range(bsz)])
        t0 = time.perf_counter()
        ref.batch(first, bsz, threads)                      # includes the serial preseed
        t_first = time.perf_counter() - t0
        done, t_sum, t = 0, 0.0, bsz
        while (t_sum < budget_s and done < 256) or done == 0:
            fr = np.stack([synth.frame(t + i) for i in range(bsz)])
            t0 = time.perf_counter()
            ref.batch(fr, bsz, threads)
            t_sum += time.perf_counter() - t0
            done += bsz
            t += bsz
        sample = "%d frames in batches of %d (after a first batch incl. the preseed, %.1f s), batch " \
                 "path, %d OpenMP threads" % (done, bsz, t_first, threads)
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
                fps=round(done / t_sum, 3), first_batch_s=round(t_first, 2), host=ci)


def run_dropin():
    """The drop-in wema.h path (wema_amd/libwema.so) driven from C as the reference's CLI
    drives it (src/main.c:343-470), with frames from frame_alloc allocated as async_io does
    (2 slots x batch x in/out): c3 through wema_batch_process with the reference's default
    batch of 64 (luma IIR batches, src/main.c:355), c2 through wema_process_frame (colour
    never batches). Host float frames cross PCIe both ways (24 bytes per pixel)."""
    exe = os.path.join(ROOT, "wema_amd", "wema_h_bench")
    if not os.path.exists(exe):
        return None
    res = {}
    for name, batch, batches in (("c3", 64, 3), ("c2", 30, 4)):
        c = CONFIGS[name]
        args = [exe, str(c["w"]), str(c["h"]), str(c["fps"]), str(c["levels"]), str(c["amp"]),
                str(c["f_low"]), str(c["f_high"]), str(int(c["color"])), str(c["preset"]), str(batch),
                str(batches)]
        out = subprocess.run(args, capture_output=True, text=True, timeout=600)
        if out.returncode != 0:
            res[name] = {"error": (out.stderr or out.stdout)[-300:]}
            continue
        res[name] = json.loads(out.stdout.strip().splitlines()[-1])
        log("drop-in %s: %.1f Mpixel/s" % (name, res[name]["mpix_per_s"]))
    return {"desc": "include/wema.h through libwema.so from C: host float Frames (frame_alloc, as "
                    "async_io allocates them), the reference CLI's call pattern", "runs": res}


REALTIME_MIX_MPXS = (8 * 640 * 480 * 30 + 8 * 1280 * 720 * 30 + 6 * 1920 * 1080 * 60
                     + 2 * 3840 * 2160 * 30) / 1e6   # real-time demand of the 24-stream mix


def _serve_summary(st):
    keys = ("frames", "mpix_per_s", "wall_s", "lat_p50_ms", "lat_p99_ms", "lat_max_ms",
            "sessions_arrived", "sessions_admitted", "admission_waits", "peak_sessions",
            "peak_bytes", "ticks", "mean_frames_per_tick", "host_ms_per_tick", "graph_tick_frac",
            "gpu_busy_frac", "link_gbs")
    return {k: (round(st[k], 3) if isinstance(st[k], float) else st[k]) for k in keys}


def run_multistream(frames_per_session=180):
    """24 concurrent streams of the 8:8:6:2 mix (640x480@30, 1280x720@30, 1920x1080@60,
    3840x2160@30), each with its own band in 0.3-5 Hz, amp 20-150 and window 16/32/64, a frame
    of every stream always due. batched: one launch per stage per tick over all 24 streams,
    replayed from a hipGraph; plain: the same frames one stream at a time."""
    import wema_amd
    common = dict(saturate=True, sessions=24, fps_from_mix=True, hbm_budget=200e9,
                  frames_per_session=frames_per_session, warmup_rounds=3, duration_s=120.0, seed=11)
    res = {}
    for name, kw in (("batched_graphs", dict(batched=True, graphs=True)),
                     ("batched", dict(batched=True, graphs=False)),
                     ("plain", dict(batched=False, graphs=False))):
        res[name] = _serve_summary(wema_amd.serve(**kw, **common))
        log("multistream %s: %.1f Mpixel/s" % (name, res[name]["mpix_per_s"]))
    b, p = res["batched_graphs"]["mpix_per_s"], res["plain"]["mpix_per_s"]
    return {"desc": "24 streams x %d frames, mix 8x640x480@30 8x1280x720@30 6x1920x1080@60 "
                    "2x3840x2160@30, per-stream band/amp/window" % frames_per_session,
            "value": round(b, 1), "unit": "Mpixel/s",
            "realtime_demand_mpix_s": round(REALTIME_MIX_MPXS, 1),
            "x_realtime": round(b / REALTIME_MIX_MPXS, 2),
            "batched_over_plain": round(b / p, 2) if p else None, "runs": res}


def run_serving(duration_s=4.0, max_mult=64.0):
    """Poisson sessions at 4/s (life 2-20 s, the 8:8:6:2 resolution mix, 30 or 60 fps),
    frames delivered at real-time pace, admission under a 200 GB HBM budget, started in
    steady state; frames resident in HBM ("device"), or uploaded from and downloaded to
    pinned host memory on copy streams overlapping the ticks ("host_io", the path a real
    ingest takes). paced: that load, batched (continuous batching) and plain (one session at
    a time). capacity: the arrival rate is raised in steps of sqrt(2) until p99 latency
    exceeds 33 ms, then bisected twice; the aggregate Mpixel/s of the highest load that met
    the bound, for each mode."""
    import wema_amd
    base = dict(arrival_rate=4.0, life=(2.0, 20.0), hbm_budget=200e9, warm_start=True,
                duration_s=duration_s, fps_from_mix=False, seed=5)
    out = {"desc": "Poisson 4/s, life 2-20 s, mix 8:8:6:2 of 640x480..3840x2160 at 30/60 fps, "
                   "real-time pace, 200 GB HBM budget, %.0f s runs from steady state; frames in "
                   "HBM (device) or through pinned host memory (host_io)" % duration_s}
    out["paced"] = {"batched": _serve_summary(wema_amd.serve(batched=True, graphs=True, **base)),
                    "plain": _serve_summary(wema_amd.serve(batched=False, graphs=False, **base)),
                    "batched_host_io": _serve_summary(wema_amd.serve(batched=True, graphs=True,
                                                                     host_io=True, **base)),
                    "plain_host_io": _serve_summary(wema_amd.serve(batched=False, graphs=False,
                                                                   host_io=True, **base))}
    log("serving at the paced load done")
    cap = {}
    for mode, kw in (("batched", dict(batched=True, graphs=True)),
                     ("plain", dict(batched=False, graphs=False)),
                     ("batched_host_io", dict(batched=True, graphs=True, host_io=True)),
                     ("plain_host_io", dict(batched=False, graphs=False, host_io=True))):
        best, tried = None, []

        def run(m):
            st = wema_amd.serve(**kw, **dict(base, arrival_rate=4.0 * m))
            log("serving %s at %.1f/s: %.1f Mpixel/s, p99 %.2f ms" % (mode, 4.0 * m, st["mpix_per_s"],
                                                                      st["lat_p99_ms"]))
            tried.append({"arrival_rate": round(4.0 * m, 2), "mpix_per_s": round(st["mpix_per_s"], 1),
                          "lat_p99_ms": round(st["lat_p99_ms"], 2), "peak_sessions": st["peak_sessions"]})
            return st, st["lat_p99_ms"] <= 33.0

        # geometric steps of sqrt(2) up to the first load that misses the bound, then two
        # bisections (geometric midpoints) between the last load that met it and that one
        m_ok, m_bad, m = None, None, 1.0
        while m <= max_mult:
            st, ok = run(m)
            if not ok:
                m_bad = m
                break
            best, m_ok = dict(_serve_summary(st), arrival_rate=round(4.0 * m, 2)), m
            m *= 2 ** 0.5
        for _ in range(2 if (m_ok and m_bad) else 0):
            mid = (m_ok * m_bad) ** 0.5
            st, ok = run(mid)
            if ok:
                best, m_ok = dict(_serve_summary(st), arrival_rate=round(4.0 * mid, 2)), mid
            else:
                m_bad = mid
        cap[mode] = {"best": best, "sweep": tried}
    out["capacity"] = cap
    for suffix in ("", "_host_io"):
        b = (cap["batched" + suffix]["best"] or {}).get("mpix_per_s")
        p = (cap["plain" + suffix]["best"] or {}).get("mpix_per_s")
        out["capacity_batched_over_plain" + suffix] = round(b / p, 2) if b and p else None
    out["p99_target_ms"] = 33.0
    return out


def state_model(cfg, batch):
    """How the temporal state's traffic enters the algorithmic bytes of a launch."""
    if not cfg["use_iir"]:
        return ("window: the ring samples a launch's windows cover (window - 1 + frames) read once "
                "per launch, one amplified coefficient written per frame")
    if batch == 1:
        return "per frame: IIR state read and written for every frame (40 B per unique position per frame)"
    return ("per launch: the %d frames of a launch run with the IIR state in registers, so the state "
            "is read and written once per launch (8 B per position per frame + 32 B per position per "
            "launch); the one-frame-per-call path is reported separately" % batch)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--batch", type=int, default=0, help="frames per step (default per config)")
    ap.add_argument("--also", default="c1,c3,c2w,c2_n1,c3_n1",
                    help="comma list of extra workloads ('' = none)")
    ap.add_argument("--cpu-also", default="c1,c3",
                    help="workloads whose reference CPU baseline (OpenMP batch path) is also timed")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the multi-stream and serving measurements")
    ap.add_argument("--serve-duration", type=float, default=4.0, help="seconds per serving run")
    ap.add_argument("--gf-seg", type=int, default=0,
                    help="guided-filter segment height in rows, a multiple of 4 (0: the "
                         "autotuner's pick; a fixed height skips the tuner's timing launches, "
                         "e.g. under a profiler)")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the pinned-host pipeline and drop-in wema.h measurements "
                         "(e.g. under a profiler)")
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

    if args.gf_seg < 0 or args.gf_seg % 4:
        print("--gf-seg: a multiple of 4 (0: the autotuner's pick)", file=sys.stderr)
        sys.exit(2)
    if args.gf_seg:
        import wema_amd
        wema_amd.autotune(False, dwt=True)   # (a fixed guided-filter segment; the strips still tuned)
    cfg = CONFIGS[args.config]
    batch = args.batch or cfg["batch"]
    steps = PER_FRAME_STEPS.get(args.config, args.steps) if batch == 1 and not args.batch else args.steps
    r = run_gpu(cfg, args.config, steps, args.warmup, batch, args.gf_seg)
    r["roofline"]["state_model"] = state_model(cfg, batch)
    log("%s: %.1f Mpixel/s" % (args.config, r["mpix_s"]))
    line = {
        "metric": METRIC,
        "value": round(r["mpix_s"], 1),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"] / steps, 3),
        "ms_per_step_with_launch_markers": round(r["ms_markers"] / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (deterministic generator, frames resident in HBM)",
        "config": {"workload": args.config, "desc": cfg["desc"], "width": cfg["w"],
                   "height": cfg["h"], "frames_per_step": batch, "levels": cfg["levels"],
                   "gf_seg_rows": r["gf_seg"] if not cfg["color"] else None},
        "fps": round(r["fps"], 1),
        "pipeline_algorithmic_GBs": r["pipeline_gbs"],
        "roofline": r["roofline"],
        "kernels": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}
                    for d in r["kernels"]],
        "cpu_baseline": None,
    }
    if not args.no_host:
        line["host_pipeline"] = run_host_pipeline(cfg, max(args.steps, 16), max(batch, 8))
        log("host pipeline done")
        line["dropin_wema_h"] = run_dropin()
        log("drop-in wema.h done")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cfg, args.cpu_budget)
        log("cpu baseline done")
    also = {}
    for name in [a for a in args.also.split(",") if a and a != args.config]:
        c = CONFIGS[name]
        st_n = PER_FRAME_STEPS.get(name, args.steps)
        rr = run_gpu(c, name, st_n, args.warmup * (8 if c["batch"] == 1 else 1), c["batch"], args.gf_seg)
        rr["roofline"]["state_model"] = state_model(c, c["batch"])
        log("%s: %.1f Mpixel/s" % (name, rr["mpix_s"]))
        also[name] = {"desc": c["desc"], "value": round(rr["mpix_s"], 1), "unit": "Mpixel/s",
                      "fps": round(rr["fps"], 1), "steps": st_n, "ms_per_step": round(rr["ms"] / st_n, 3),
                      "ms_per_step_with_launch_markers": round(rr["ms_markers"] / st_n, 3),
                      "frames_per_step": c["batch"], "pipeline_algorithmic_GBs": rr["pipeline_gbs"],
                      "gf_seg_rows": rr["gf_seg"] if not c["color"] else None,
                      "roofline": rr["roofline"],
                      "kernels": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}
                                  for d in rr["kernels"]]}
    if not args.no_cpu:
        for name in [a for a in args.cpu_also.split(",") if a]:
            if name in also and name != args.config:
                also[name]["cpu_baseline"] = cpu_baseline(CONFIGS[name], args.cpu_budget)
                log("%s cpu baseline done" % name)
    if also:
        line["also"] = also
    if not args.no_serve:
        line["multistream"] = run_multistream()
        line["serving"] = run_serving(args.serve_duration)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
