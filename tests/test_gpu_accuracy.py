"""Per-stage error of the HIP pipeline at the benchmark workloads' full sizes, against the
FP64 oracle and against the compiled reference (src/wema.c:746-952 built -Ofast).

Each stage is compared on the same input as the device stage received (stage-local
error). Stages and bars:
  * ingest, RGB24 -> Y (Cb, Cr) in FP32 (src/ffmpeg_io.c:292-294, src/wema.c:42-78): within
    3e-7 (absolute, values in [0, 1]) of the FP64 conversion with the same constants;
  * DWT coefficients (src/dwt.c:645-703): the FP64 transform of the FP32 channel plane,
    norm-wise (max |error| / max |input|) within 1e-5; also reported relative to the
    largest coefficient (dwt_vs_f64_coef_scale: the chroma planes' detail coefficients are
    small beside the LL path that carries the rounding); bit-identical to the reference's.
  * IIR output (src/iir_filter.c:181-264, preseed 376-471): the FP64 recursion run on the
    same FP32 coefficients and filter taps, norm-wise (max |error| / max |input|) within the
    band's bound (1e-5 for 0.5-3 Hz at 30 fps, 2e-5 for 0.8-2 Hz at 60 fps); bit-identical
    to the reference's.
  * amplified coefficients (src/wema.c:714-744) and output pixels (src/wema.c:289-335,
    src/ffmpeg_io.c:304-310), with the reference's thresholds injected (its FP32 threshold
    sum is vectorised in a compiler-chosen order): amplified coefficients bit-identical,
    output bytes within 1 LSB.
With WEMA_ACCURACY_OUT set, the measured errors are written there as JSON.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import wema_amd  # noqa: E402
from oracle import oracle as O  # noqa: E402

L = O.orc()
ptr = O.ptr

WORKLOADS = {
    # name: preset, size, fps, settings, frames, IIR bound
    "c1": (1, 320, 240, 30.0, dict(levels=3, amp=120.0, f_low=0.5, f_high=3.0), 4, 1e-5),
    "c2": (2, 1920, 1080, 30.0, dict(levels=5, amp=120.0, f_low=0.5, f_high=3.0, color=True), 3, 1e-5),
    "c3": (3, 3840, 2160, 60.0, dict(levels=6, amp=50.0, f_low=0.8, f_high=2.0), 2, 2e-5),
}
RESULTS = {}


def channel_f64(rgb, ch):
    """Y, Cb or Cr of an RGB24 frame in FP64 (src/wema.c:42-78, constants as written)."""
    r, g, b = (rgb[..., k].astype(np.float64) / 255.0 for k in range(3))
    if ch == 0:
        v = 0.299 * r + 0.587 * g + 0.114 * b
    elif ch == 1:
        v = -0.169 * r - 0.331 * g + 0.5 * b + 0.5
    else:
        v = 0.5 * r - 0.419 * g - 0.081 * b + 0.5
    return np.ascontiguousarray(v.reshape(-1))


def dwt_f64(img, w, h, levels):
    planes = np.zeros(O.num_unique(w, h, levels))
    ll = np.zeros(w * h)
    L.orc_dwt_fwd_f64(ptr(img), w, h, levels, ptr(planes), ptr(ll))
    return planes


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_stage_errors_vs_fp64_and_reference(name):
    preset, w, h, fps, kw, n, iir_bound = WORKLOADS[name]
    levels, nch = kw["levels"], 3 if kw.get("color") else 1
    synth = wema_amd.Synth(preset, w, h, fps)
    frames = np.stack([synth.frame(t) for t in range(n)])
    sess = wema_amd.Session(w, h, fps, max_batch=1, capture=True, **kw)
    ref = O.RefPipe(w, h, fps, **kw) if O.ref_available() else None
    c32, warm = sess.iir_coeffs()
    c64 = c32.astype(np.float64)
    d_in = torch.from_numpy(frames).cuda()
    d_out = torch.empty_like(d_in)
    st64 = [None] * nch
    err = dict(ingest_vs_f64=0.0, dwt_vs_f64=0.0, dwt_vs_f64_coef_scale=0.0, iir_vs_f64=0.0, dwt_vs_ref=0.0,
               iir_vs_ref=0.0, amplified_vs_ref=0.0, output_lsb_vs_ref=0)
    for t in range(n):
        if ref is not None:
            r_out, r_thr = ref.frame(frames[t])
            sess.set_thresholds(r_thr[None])
        sess.process_device(d_in[t:t + 1], d_out_u8=d_out[t:t + 1], n=1)
        sess.sync()
        for ch in range(nch):
            g_coef, g_filt, g_amp = sess.read(0, 0, ch), sess.read(1, 0, ch), sess.read(2, 0, ch)
            # the channel plane the device transformed: the FP32 ingest, which the device's
            # fused ingest equals bit for bit (the coefficients are bit-identical to the
            # reference's and the oracle's)
            chan32 = np.zeros(w * h, np.float32)
            L.orc_ingest_f32(ptr(np.ascontiguousarray(frames[t])), w * h, ch, ptr(chan32))
            err["ingest_vs_f64"] = max(err["ingest_vs_f64"],
                                       float(np.abs(chan32 - channel_f64(frames[t], ch)).max()))
            p64 = dwt_f64(np.ascontiguousarray(chan32, np.float64), w, h, levels)
            e = float(np.abs(g_coef - p64).max())
            err["dwt_vs_f64"] = max(err["dwt_vs_f64"], e / float(np.abs(chan32).max()))
            err["dwt_vs_f64_coef_scale"] = max(err["dwt_vs_f64_coef_scale"], e / float(np.abs(p64).max()))
            x64 = np.ascontiguousarray(g_coef, np.float64)
            nu = x64.size
            if t == 0:
                st64[ch] = np.zeros(4 * nu)
                L.orc_iir_preseed_f64(ptr(c64), ptr(x64), nu, warm, ptr(st64[ch]))
            y64 = np.zeros(nu)
            L.orc_iir_step_f64(ptr(c64), ptr(x64), nu, ptr(st64[ch]), ptr(y64))
            err["iir_vs_f64"] = max(err["iir_vs_f64"], float(np.abs(g_filt - y64).max() / np.abs(x64).max()))
            if ref is not None:
                rc = O.flat6_to_planes(ref.buffer(0, ch), w, h, levels)
                rf = O.flat6_to_planes(ref.buffer(1, ch), w, h, levels)
                ra = O.flat6_to_planes(ref.amplified(ch), w, h, levels)
                err["dwt_vs_ref"] = max(err["dwt_vs_ref"], float(np.abs(g_coef - rc).max()))
                err["iir_vs_ref"] = max(err["iir_vs_ref"], float(np.abs(g_filt - rf).max()))
                err["amplified_vs_ref"] = max(err["amplified_vs_ref"], float(np.abs(g_amp - ra).max()))
        if ref is not None:
            g = d_out[t].cpu().numpy()
            err["output_lsb_vs_ref"] = max(err["output_lsb_vs_ref"],
                                           int(np.abs(g.astype(int) - r_out.astype(int)).max()))
    sess.close()
    RESULTS[name] = dict(err, frames=n, width=w, height=h, levels=levels, channels=nch,
                         reference=ref is not None)
    print(name, json.dumps(RESULTS[name]))
    out = os.environ.get("WEMA_ACCURACY_OUT")
    if out:
        with open(out, "w") as f:
            json.dump(RESULTS, f, indent=1)
    assert err["ingest_vs_f64"] <= 3e-7, err
    assert err["dwt_vs_f64"] <= 1e-5, err
    assert err["iir_vs_f64"] <= iir_bound, err
    if ref is not None:
        assert err["dwt_vs_ref"] == 0.0 and err["iir_vs_ref"] == 0.0, err
        assert err["amplified_vs_ref"] == 0.0, err
        assert err["output_lsb_vs_ref"] <= 1, err
