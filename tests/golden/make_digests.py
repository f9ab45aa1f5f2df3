"""Writes tests/golden/oracle_digests.json: per-frame SHA-256 digests of the oracle's output
bytes and its exact thresholds (float32 bit patterns), for workloads too large to run the
oracle on at test time:

  c2      1920x1080 colour IIR, 5 levels, 0.5-3 Hz, amp 120, 20 frames (past the 15-frame
          fade-in of src/wema.c:830-839, so frames 15-19 run at full gain);
  c3      3840x2160 luma + guided filter, 6 levels, 0.8-2 Hz, amp 50, 20 frames;
  c2w_nb  1920x1080 colour, 5 levels, sliding window of 32 frames without bilateral weights,
          amp 120, 36 frames (the window is full from frame 31 on);
  c2w     the same with bilateral weights (the window benchmark): the exponential differs
          from libm's by an ulp, so output bytes may differ by 1 LSB and a digest cannot
          hold them; instead the bytes at 2048 fixed positions of every frame are stored
          (base64), with the exact thresholds;
  c4_*    the 640x480 and 1280x720 streams of the 24-stream mix (8:8:6:2 of 640x480@30,
          1280x720@30, 1920x1080@60, 3840x2160@30; band in 0.3-5 Hz, amp in 20-150 and
          window in {16, 32, 64} drawn per stream from a fixed seed; automatic levels;
          IIR luma + guided filter), 20 frames each. The parameters of all 24 streams are
          stored too, so that the GPU test runs exactly these streams.

The oracle is the C restatement in oracle/ (FP32, the reference's operation order, exact
threshold sums); the frames come from the deterministic generator of wema_amd.Synth. The
GPU tests (tests/test_gpu_pipeline.py, tests/test_gpu_pool.py) compare the HIP path's bytes
and thresholds with these digests.

Run from the repository root:  python tests/golden/make_digests.py   (8 processes, ~2 min)
"""
import base64
import hashlib
import json
import multiprocessing as mp
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(ROOT, "tests", "golden", "oracle_digests.json")
FRAMES = 20

FULL = {
    "c2": dict(preset=2, w=1920, h=1080, fps=30.0, seed_offset=0,
               kw=dict(levels=5, amp=120.0, f_low=0.5, f_high=3.0, color=True)),
    "c3": dict(preset=3, w=3840, h=2160, fps=60.0, seed_offset=0,
               kw=dict(levels=6, amp=50.0, f_low=0.8, f_high=2.0)),
    "c2w_nb": dict(preset=2, w=1920, h=1080, fps=30.0, seed_offset=0, frames=36,
                   kw=dict(levels=5, amp=120.0, f_low=0.5, f_high=3.0, color=True, use_iir=False,
                           window=32)),
    "c2w": dict(preset=2, w=1920, h=1080, fps=30.0, seed_offset=0, frames=36, samples=2048,
                kw=dict(levels=5, amp=120.0, f_low=0.5, f_high=3.0, color=True, use_iir=False,
                        window=32, bilateral=True)),
}

MIX = [(640, 480, 30.0, 8), (1280, 720, 30.0, 8), (1920, 1080, 60.0, 6), (3840, 2160, 30.0, 2)]


def c4_streams():
    """The 24 streams: class in exact 8:8:6:2 proportions, per-stream band, amp, window."""
    rng = np.random.default_rng(0xC4)
    out = []
    for w, h, fps, count in MIX:
        for _ in range(count):
            a, b = sorted(rng.uniform(0.3, 5.0, size=2))
            if b - a < 0.2:
                b = a + 0.2
            out.append(dict(preset=4, w=w, h=h, fps=fps, seed_offset=100 + len(out),
                            kw=dict(levels=0, amp=float(np.float32(rng.uniform(20.0, 150.0))),
                                    f_low=float(np.float32(a)), f_high=float(np.float32(b)),
                                    window=int(rng.choice([16, 32, 64])))))
    return out


def sample_positions(w, h, n):
    """Fixed byte positions of an RGB24 frame for the sampled cases."""
    return np.sort(np.random.default_rng(0x5A).choice(3 * w * h, size=n, replace=False))


def run_case(args):
    name, c = args
    import wema_amd
    from oracle import oracle as O
    syn = wema_amd.Synth(c["preset"], c["w"], c["h"], c["fps"], c["seed_offset"])
    orc = O.OraclePipe(c["w"], c["h"], fps=c["fps"], **c["kw"])
    frames = c.get("frames", FRAMES)
    pos = sample_positions(c["w"], c["h"], c["samples"]) if c.get("samples") else None
    sha, thr, smp = [], [], []
    for t in range(frames):
        out, th = orc.frame(syn.frame(t))
        sha.append(hashlib.sha256(out.tobytes()).hexdigest())
        thr.append([int(x) for x in th.astype(np.float32).view(np.uint32)])
        if pos is not None:
            smp.append(base64.b64encode(out.reshape(-1)[pos].tobytes()).decode())
    res = dict(params=c, frames=frames, out_sha256=sha, thr_bits=thr)
    if pos is not None:
        res["samples_b64"] = smp
    return name, res


def main():
    jobs = list(FULL.items())
    streams = c4_streams()
    for i, s in enumerate(streams):
        if s["w"] <= 1280:
            jobs.append(("c4_%02d" % i, s))
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        res = dict(pool.map(run_case, jobs, chunksize=1))
    doc = {"about": "oracle output digests; written by tests/golden/make_digests.py",
           "c4_streams": streams, "cases": {k: res[k] for k in sorted(res)}}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote %s: %d cases" % (OUT, len(res)))


if __name__ == "__main__":
    main()
