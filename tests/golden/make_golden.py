"""Writes tests/golden/pipeline_ref.npz: input frames, output bytes, per-frame thresholds,
and the last frame's coefficients and temporal outputs of the reference pipeline, compiled
from its own sources (oracle/_ref/libwema_ref.so, built by `make -C oracle ref`).

The reference's per-frame path is wema_process_frame (src/wema.c:746-952); thresholds are
the reference's own FP32 sum (src/wema.c:714-744, src/phase_amp.c:12-56) read back by
oracle/ref_harness.c; coefficients are its 6-slot flat layout (src/dwt.c:679-687) reduced
to the 3 unique orientations [level][LH, HH, HL].

Run from the repository root:  python tests/golden/make_golden.py [--out PATH]
The reference reads freshly allocated memory where an odd level size leaves a subband row
or column unwritten (src/dwt.c:459-472, 648-653); the script runs itself with
MALLOC_PERTURB_=255, under which glibc hands out zero-filled blocks, so those entries are 0
and the fixtures are reproducible. The fixtures are data only (uint8 frames, float32
vectors); nothing in them executes.
"""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "pipeline_ref.npz")

# name -> (width, height, frames, seed, pipeline settings). Pixel counts are multiples of
# 8: the reference's colour-conversion loops compute the last (pixels mod 8) pixels in a
# scalar epilogue with another FMA contraction (tests/test_golden.py characterises it).
CASES = {
    "iir": (40, 30, 8, 21, dict(use_iir=True)),
    "iir_odd_width": (37, 24, 8, 22, dict(use_iir=True)),
    "iir_color_odd_height": (40, 29, 8, 27, dict(use_iir=True, color=True)),
    "iir_color": (40, 30, 8, 23, dict(use_iir=True, color=True)),
    "window": (40, 30, 10, 24, dict(use_iir=False)),
    "window_bilateral_color": (40, 30, 10, 25, dict(use_iir=False, bilateral=True, color=True)),
    "gauss": (40, 30, 8, 26, dict(use_iir=True, edge_aware=False)),
    "window32_bilateral": (32, 24, 36, 28, dict(use_iir=False, bilateral=True, window=32)),
}
COMMON = dict(fps=30.0, amp=120.0, f_low=0.5, f_high=3.0, window=6, levels=3)


def frames_for(w, h, n, seed, color):
    """A static noise texture with a global brightness oscillation and per-frame noise."""
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 256, (h, w, 3 if color else 1)).astype(np.float32)
    f = np.stack([np.clip(base + rng.normal(0, 6, base.shape) + 30 * np.sin(t * 0.7), 0, 255)
                  .astype(np.uint8) for t in range(n)])
    return f if color else np.repeat(f, 3, axis=3)   # luma cases: grey frames, R = G = B


def make_case(name):
    w, h, n, seed, extra = CASES[name]
    kw = dict(COMMON, **extra)
    color = kw.get("color", False)
    nch = 3 if color else 1
    frames = frames_for(w, h, n, seed, color)
    ref = O.RefPipe(w, h, **kw)
    outs, thrs = [], []
    for t in range(n):
        o, thr = ref.frame(frames[t])
        outs.append(o)
        thrs.append(thr.copy())
    coef = np.stack([O.flat6_to_planes(ref.buffer(0, c), w, h, kw["levels"]) for c in range(nch)])
    filt = np.stack([O.flat6_to_planes(ref.buffer(1, c), w, h, kw["levels"]) for c in range(nch)])
    ref.close()
    d = {
        "params": np.array(json.dumps(dict(width=w, height=h, **kw))),
        "in": frames if color else frames[..., 0].copy(),
        "out": np.stack(outs) if color else np.stack(outs)[..., 0].copy(),
        "thr": np.stack(thrs).astype(np.float32),
        "coef": coef.astype(np.float32),
        "filt": filt.astype(np.float32),
    }
    if not color:   # luma fixtures store one channel of R = G = B frames
        assert np.array_equal(np.stack(outs)[..., 0], np.stack(outs)[..., 1])
    return {name + "/" + k: v for k, v in d.items()}


def load(path=OUT):
    """{case: dict(params, in, out, thr, coef, filt)}; frames come back as (n, h, w, 3)."""
    z = np.load(path, allow_pickle=False)
    cases = {}
    for key in z.files:
        name, field = key.split("/")
        cases.setdefault(name, {})[field] = z[key]
    for c in cases.values():
        c["params"] = json.loads(str(c["params"]))
        for f in ("in", "out"):
            if c[f].ndim == 3:
                c[f] = np.repeat(c[f][..., None], 3, axis=3)
    return cases


def generate():
    data = {}
    for name in CASES:
        data.update(make_case(name))
    return data


if __name__ == "__main__":
    if os.environ.get("MALLOC_PERTURB_") != "255":
        env = dict(os.environ, MALLOC_PERTURB_="255")
        sys.exit(subprocess.run([sys.executable, os.path.abspath(__file__)] + sys.argv[1:],
                                env=env).returncode)
    if not O.ref_available():
        sys.exit("oracle/_ref/libwema_ref.so is not built (make -C oracle ref)")
    out = sys.argv[sys.argv.index("--out") + 1] if "--out" in sys.argv else OUT
    np.savez_compressed(out, **generate())
    print("wrote", out, os.path.getsize(out), "bytes")
