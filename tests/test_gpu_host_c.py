"""The C11 host API (include/wema.h, wema_amd/libwema.so), driven from C the way the
reference's CLI drives its wema.h (tests/c/wema_c_driver.c): per-frame and batch paths,
luma/colour, IIR/window; pinned and pageable frames (frame_alloc past its pinned-byte
cap); batches larger than the host's sub-batch (several sub-batches in flight); calls
passing more frames than the batch size, which are clamped to it as the reference does
(src/wema_batch.c:355-357). The bar: output bytes identical to the oracle run with the same
options (the bilateral mode within 1 LSB: expf)."""
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

CASES = {
    "iir_frame": dict(use_iir=1, color=0, edge=1, bil=0, batch=1),
    "iir_color_frame": dict(use_iir=1, color=1, edge=1, bil=0, batch=1),
    "iir_batch": dict(use_iir=1, color=0, edge=0, bil=0, batch=5),
    "window_frame": dict(use_iir=0, color=0, edge=1, bil=0, batch=1),
    "window_bilateral_color": dict(use_iir=0, color=1, edge=1, bil=1, batch=1),
    "gauss_frame": dict(use_iir=1, color=0, edge=0, bil=0, batch=1),
    "iir_batch_pageable": dict(use_iir=1, color=0, edge=1, bil=0, batch=5, pinned_cap=0),
    "iir_batch_over": dict(use_iir=1, color=0, edge=1, bil=0, batch=5, over=3),
    "iir_batch_sub_batches": dict(use_iir=1, color=0, edge=1, bil=0, batch=20),
}


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("drv") / "wema_c_driver")
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "wema_c_driver.c"), "-L" + os.path.join(ROOT, "wema_amd"),
                    "-lwema", "-Wl,-rpath," + os.path.join(ROOT, "wema_amd"), "-lm", "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("case", sorted(CASES))
def test_c_host_equals_oracle(case, driver, tmp_path):
    c = CASES[case]
    w, h, n, fps, window, levels = 64, 48, 12, 30.0, 8, 3
    syn = wema_amd.Synth(2 if c["color"] else 4, w, h, fps, 5)
    frames = np.stack([syn.frame(t) for t in range(n)])
    fin, fout = tmp_path / "in.rgb", tmp_path / "out.rgb"
    frames.tofile(fin)
    args = [driver, str(fin), str(fout), str(w), str(h), str(n), str(fps), "90", "0.5", "3.0",
            str(window), str(levels), str(c["use_iir"]), str(c["color"]), str(c["edge"]), str(c["bil"]),
            str(c["batch"]), str(c.get("over", 0))]
    env = dict(os.environ)
    if "pinned_cap" in c:
        env["WEMA_PINNED_MAX_BYTES"] = str(c["pinned_cap"])
    res = subprocess.run(args, capture_output=True, text=True, timeout=120, env=env)
    assert res.returncode == 0, res.stderr
    got = np.fromfile(fout, np.uint8).reshape(n, h, w, 3)
    kw = dict(fps=fps, amp=90.0, f_low=0.5, f_high=3.0, window=window, levels=levels,
              use_iir=bool(c["use_iir"]), color=bool(c["color"]), bilateral=bool(c["bil"]),
              edge_aware=bool(c["edge"]) or c["batch"] > 1)
    orc = O.OraclePipe(w, h, **kw)
    worst = 0
    for t in range(n):
        o_out, _ = orc.frame(frames[t])
        worst = max(worst, int(np.abs(got[t].astype(int) - o_out.astype(int)).max()))
    assert worst <= (1 if c["bil"] else 0), worst
    words = res.stdout.split()
    assert int(words[words.index("ready") + 1]) == 1
    assert int(words[words.index("pinned") + 1]) == (0 if "pinned_cap" in c else 1)
