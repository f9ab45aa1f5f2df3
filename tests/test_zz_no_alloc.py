"""wema_dev_process_u8_noalloc, through the C ABI: the kernels' per-call buffers come from
the library's static device scratch instead of the session's memory. Bar: output bytes and
thresholds identical to wema_dev_process_u8 on a twin session, call after call, where
consecutive calls use different sessions and the later ones smaller inputs (so the scratch
holds the previous call's contents); a call that needs more scratch than the array holds is
refused with an error that says so."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import wema_amd  # noqa: E402

SESSIONS = {
    # name: (synth preset, w, h, fps, session options, frames per call)
    "luma_gf": (1, 320, 240, 30.0, dict(amp=120.0, f_low=0.5, f_high=3.0, levels=3, max_batch=4)),
    "colour": (2, 96, 54, 30.0, dict(amp=60.0, f_low=0.4, f_high=4.0, color=True, max_batch=2)),
    "window_bilateral": (4, 80, 60, 30.0, dict(amp=90.0, use_iir=False, window=8, bilateral=True,
                                               max_batch=3)),
}
# (session, frames in the call): each call is smaller than the one before it or uses another
# session, so stale scratch from the previous call would show
CALLS = [("luma_gf", 4), ("colour", 2), ("luma_gf", 1), ("window_bilateral", 3), ("colour", 1),
         ("window_bilateral", 3), ("window_bilateral", 3), ("luma_gf", 3), ("window_bilateral", 1),
         ("luma_gf", 2)]


def test_noalloc_equals_process_u8():
    pairs, clips, pos = {}, {}, {}
    for name, (preset, w, h, fps, kw) in SESSIONS.items():
        pairs[name] = (wema_amd.Session(w, h, fps, **kw), wema_amd.Session(w, h, fps, **kw))
        syn = wema_amd.Synth(preset, w, h, fps, seed_offset=7)
        n_all = sum(k for s, k in CALLS if s == name)
        clips[name] = torch.from_numpy(np.stack([syn.frame(t) for t in range(n_all)])).cuda()
        pos[name] = 0
    for i, (name, k) in enumerate(CALLS):
        a, b = pairs[name]
        t0 = pos[name]
        x = clips[name][t0:t0 + k]
        out_a = torch.empty_like(x)
        out_b = torch.full_like(x, 77)
        print("call %d: %s, %d frame(s) from frame %d" % (i, name, k, t0), flush=True)
        a.process_device(x, d_out_u8=out_a, n=k)
        b.process_device_noalloc(x, out_b, n=k)
        a.sync()
        b.sync()
        assert torch.equal(out_a, out_b), (i, name)
        assert np.array_equal(a.thresholds(k), b.thresholds(k), equal_nan=True), (i, name)
        pos[name] += k
    for a, b in pairs.values():
        a.close()
        b.close()


def test_noalloc_refuses_more_than_the_scratch_holds():
    w, h = 3840, 2160   # one 4K guided-filter frame needs about 145 MB of per-call buffers
    s = wema_amd.Session(w, h, 60.0, amp=50.0, f_low=0.8, f_high=2.0, levels=6)
    x = torch.zeros((1, h, w, 3), dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    with pytest.raises(wema_amd.WemaError, match="scratch"):
        s.process_device_noalloc(x, y, n=1)
    assert s.frames == 0
    s.close()
