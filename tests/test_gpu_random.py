"""Randomized parity sweep of the HIP path against the oracle: frame sizes from 1 pixel to a
few hundred, DWT levels, temporal modes, window sizes (register and LDS kernels), colour,
Gaussian transfer and batch sizes, drawn from a fixed seed. Each case opens its own session
(so session set-up is exercised many times in one process) and compares every output frame
with the oracle: identical bytes, or within 1 LSB with bilateral weights (exponential)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import wema_amd  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _case(rng):
    w = int(rng.integers(1, 200)) if rng.random() < 0.8 else int(rng.integers(1, 24))
    h = int(rng.integers(1, 150)) if rng.random() < 0.8 else int(rng.integers(1, 24))
    maxl = max(1, min(5, int(np.log2(max(2, min(w, h))))))
    mode = str(rng.choice(["iir", "iir_color", "window", "window_bil", "window_color", "gauss"]))
    ws = int(rng.choice([4, 6, 16, 24, 32, 64]))
    kw = dict(fps=float(rng.choice([24.0, 30.0, 60.0])), amp=float(rng.uniform(20, 150)), f_low=0.5,
              f_high=3.0, levels=int(rng.integers(1, maxl + 1)), window=ws)
    if mode == "iir_color":
        kw.update(color=True)
    elif mode == "window":
        kw.update(use_iir=False)
    elif mode == "window_bil":
        kw.update(use_iir=False, bilateral=True)
    elif mode == "window_color":
        kw.update(use_iir=False, color=True, bilateral=bool(rng.random() < 0.5))
    elif mode == "gauss":
        kw.update(edge_aware=False)
    n = ws + 3 if not kw.get("use_iir", True) else int(rng.integers(2, 8))
    return w, h, n, int(rng.integers(1, 9)), kw


@pytest.mark.parametrize("chunk", range(4))
def test_random_cases_vs_oracle(chunk):
    rng = np.random.default_rng(1000 + chunk)
    for _ in range(25):
        w, h, n, batch, kw = _case(rng)
        frames = rng.integers(0, 256, size=(n, h, w, 3), dtype=np.uint8)
        sess = wema_amd.Session(w, h, max_batch=batch, **kw)
        d_in = torch.from_numpy(frames).cuda()
        d_out = torch.empty_like(d_in)
        for s0 in range(0, n, batch):
            k = min(batch, n - s0)
            sess.process_device(d_in[s0:s0 + k], d_out_u8=d_out[s0:s0 + k], n=k)
        sess.sync()
        got = d_out.cpu().numpy()
        sess.close()
        orc = O.OraclePipe(w, h, **kw)
        worst = 0
        for t in range(n):
            o, _ = orc.frame(frames[t])
            worst = max(worst, int(np.abs(got[t].astype(int) - o.astype(int)).max()))
        orc.close()
        assert worst <= (1 if kw.get("bilateral") else 0), (w, h, n, batch, kw, worst)
