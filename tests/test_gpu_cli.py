"""The command-line front end end to end on the GPU: raw RGB24 frames through the FFmpeg
stand-ins of tests/cli/ (FFmpeg is absent from this image), the CLI's pinned ring, the
asynchronous device pipeline and back. The bar: every output byte equals the oracle's for
the same options, the CLI's mode matrix included (luma IIR runs the batch path's semantics,
guided filter always on, src/main.c:355; colour and window modes the per-frame path's), the
frame rate's Nyquist clamp of --fh (src/main.c:247-250), and --levels (which the reference
cannot set)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import wema_amd  # noqa: E402
from oracle import oracle as O  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "wema_amd", "wema")
FAKES = os.path.join(ROOT, "tests", "cli")
W, H, N, FPS = 96, 64, 23, 30.0

CASES = {
    # name: (CLI options, oracle options)
    "luma_iir_levels4": (["--levels", "4", "--batch", "5", "-a", "90", "--fl", "0.5", "--fh", "3"],
                         dict(levels=4, amp=90.0, f_low=0.5, f_high=3.0, edge_aware=True)),
    "luma_iir_no_edge_is_batch_path": (["--no-edge-aware", "--batch", "4", "--fh", "2.5"],
                                       dict(levels=0, f_low=0.5, f_high=2.5, edge_aware=True)),
    "colour_levels2_nyquist": (["--color", "--levels", "2", "--batch", "3"],
                               dict(levels=2, f_low=0.5, color=True,
                                    f_high=float(np.float32(FPS) / np.float32(2) - np.float32(0.0001)))),
    "window_bilateral": (["--haar", "--temporal-window", "8", "--bilateral-filter", "--no-edge-aware",
                          "--levels", "3"],
                         dict(levels=3, use_iir=False, window=8, bilateral=True, edge_aware=False)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_cli_end_to_end_equals_oracle(case, tmp_path):
    args, okw = CASES[case]
    syn = wema_amd.Synth(2 if okw.get("color") else 4, W, H, FPS, 31)
    frames = np.stack([syn.frame(t) for t in range(N)])
    src, dst = tmp_path / "in.raw", tmp_path / "out.mkv"
    frames.tofile(src)
    env = dict(os.environ, PATH=FAKES + os.pathsep + os.environ["PATH"], FAKE_W=str(W), FAKE_H=str(H),
               FAKE_FPS="30/1", FAKE_N=str(N))
    r = subprocess.run([CLI, "-i", str(src), "-o", str(dst), "-v"] + args, capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(dst, np.uint8)
    assert got.size == frames.size, (got.size, frames.size)
    got = got.reshape(frames.shape)
    kw = dict(okw)
    kw.setdefault("amp", 120.0)
    orc = O.OraclePipe(W, H, fps=FPS, **kw)
    want = np.stack([orc.frame(f)[0] for f in frames])
    tol = 1 if kw.get("bilateral") else 0
    assert int(np.abs(got.astype(int) - want.astype(int)).max()) <= tol
