"""The oracle against the reference pipeline compiled from its own sources
(oracle/_ref/libwema_ref.so, see oracle/Makefile). CPU only.

Bit-exact: DWT forward and inverse, IIR design, IIR filtering (preseed, per-frame and
batch), amplification given the threshold, colour conversion. Within tolerance: the box
filter (the reference uses FP32 prefix sums whose rounding the oracle reproduces only in
vectorised loop bodies) and the sliding-window mean on frames whose window starts at ring
slot 0 (a vectorised sum). The threshold sum is compared separately: the reference sums
|c| in FP32 in a compiler-vectorised order, the oracle (like the GPU) exactly.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.ref

ptr = O.ptr


def test_reference_built_with_raised_level_cap():
    assert O.ref().refh_max_levels() == 8
    assert O.ref().refh_auto_levels(320, 240) == 4
    assert O.ref().refh_auto_levels(1920, 1080) == 4


def test_u8_conversions_exhaustive():
    u = np.arange(256, dtype=np.uint8)
    f = np.zeros(256, np.float32)
    O.ref().refh_u8_to_f32(ptr(u), ptr(f), 256)
    assert np.array_equal(f, u.astype(np.float32) * np.float32(1.0 / 255.0))
    back = np.zeros(256, np.uint8)
    O.ref().refh_f32_to_u8(ptr(f), ptr(back), 256)
    assert np.array_equal(back, u)            # passthrough frames round-trip exactly


@pytest.mark.parametrize("w,h,levels", [(64, 48, 3), (80, 36, 4), (320, 240, 3), (96, 70, 4)])
def test_dwt_bitexact(w, h, levels):
    """Level-0 dimensions even; (80,36,4) and (96,70,4) have odd parents at deeper levels,
    where the reference keeps the previous level's scratch values (reproduced exactly)."""
    img = np.random.default_rng(w).random(w * h).astype(np.float32)
    n6 = O.ref().refh_dwt_forward(ptr(img), w, h, levels, None, None)
    f6 = np.zeros(n6, np.float32)
    ll = np.zeros(w * h, np.float32)
    O.ref().refh_dwt_forward(ptr(img), w, h, levels, ptr(f6), ptr(ll))
    planes = np.zeros(n6 // 2, np.float32)
    oll = np.zeros(w * h, np.float32)
    O.orc().orc_dwt_fwd_f32(ptr(img), w, h, levels, ptr(planes), ptr(oll))
    assert np.array_equal(O.planes_to_flat6(planes, w, h, levels), f6)
    assert np.array_equal(ll, oll)
    p2 = (planes + np.random.default_rng(1).normal(0, 0.01, planes.size)).astype(np.float32)
    f62 = O.planes_to_flat6(p2, w, h, levels)
    r = np.zeros(w * h, np.float32)
    o = np.zeros(w * h, np.float32)
    O.ref().refh_dwt_inverse(ptr(f62), ptr(ll), w, h, levels, ptr(r))
    O.orc().orc_dwt_inv_f32(ptr(p2), ptr(ll), w, h, levels, ptr(o))
    assert np.array_equal(r, o)


def test_dwt_odd_level0_differs_only_at_unwritten_entries():
    """Odd image sizes leave level-0 subband entries the reference never writes (it reads
    fresh malloc memory there); every other coefficient is identical."""
    w, h, levels = 37, 29, 3
    img = np.random.default_rng(2).random(w * h).astype(np.float32)
    n6 = O.ref().refh_dwt_forward(ptr(img), w, h, levels, None, None)
    f6 = np.zeros(n6, np.float32)
    O.ref().refh_dwt_forward(ptr(img), w, h, levels, ptr(f6), ptr(np.zeros(w * h, np.float32)))
    planes = np.zeros(n6 // 2, np.float32)
    O.orc().orc_dwt_fwd_f32(ptr(img), w, h, levels, ptr(planes), ptr(np.zeros(w * h, np.float32)))
    ref_planes = O.flat6_to_planes(f6, w, h, levels)
    sw, sh = (w + 1) // 2, (h + 1) // 2
    mask = np.ones(planes.size, bool)
    lh_, hh, hl = (np.arange(k * sw * sh, (k + 1) * sw * sh).reshape(sh, sw) for k in range(3))
    mask[lh_[h // 2:].ravel()] = False
    mask[hh[h // 2:].ravel()] = False
    mask[hh[:, w // 2:].ravel()] = False
    mask[hl[:, w // 2:].ravel()] = False
    # deeper levels inherit level-0 scratch only through these entries' flat indices
    n0 = 3 * sw * sh
    assert np.array_equal(planes[:n0][mask[:n0]], ref_planes[:n0][mask[:n0]])


def _fma32(a, b, c):
    """fmaf on float32 operands: the product is exact in float64 and, for operands in
    [0, 1], so is the sum, so one rounding to float32 remains."""
    return (np.float64(a) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


@pytest.mark.parametrize("npx", [800, 801, 805, 807, 808, 813, 3263])
def test_luma_loop_epilogue_contraction(npx):
    """frame_rgb_to_gray (src/wema.c:42-51) as compiled (-Ofast, AVX-512) runs 16- and
    8-wide vector bodies, whose contraction fma(0.114 b, fma(0.299 r, 0.587 g)) the oracle
    and the kernels use at every pixel, and a scalar epilogue for the last npx mod 8
    pixels, which evaluates fma(0.299 r, fma(0.587 g, 0.114 b)). Frame sizes whose pixel
    count is not a multiple of 8 therefore differ from the reference by an ulp of Y at up
    to 7 pixels of the last row."""
    f = np.float32
    rng = np.random.default_rng(npx)
    tail = npx - npx % 8
    seen_diff = False
    for _ in range(20):
        u8 = rng.integers(0, 256, npx * 3).astype(np.uint8)
        rgb = u8.astype(f) * f(1.0 / 255.0)
        ref = np.zeros(npx, f)
        O.ref().refh_rgb_to_gray(ptr(rgb), npx, 1, ptr(ref))
        orc = np.zeros(npx, f)
        O.orc().orc_ingest_f32(ptr(u8), npx, 0, ptr(orc))
        assert np.array_equal(orc[:tail], ref[:tail])
        r, g, b = rgb[0::3][tail:], rgb[1::3][tail:], rgb[2::3][tail:]
        scalar = _fma32(f(0.299), r, _fma32(f(0.587), g, (f(0.114) * b).astype(f)))
        assert np.array_equal(scalar, ref[tail:])
        seen_diff |= not np.array_equal(orc[tail:], ref[tail:])
    assert seen_diff == (npx % 8 != 0)


def test_iir_design_bitexact_sweep():
    bad = 0
    for fs in (24.0, 25.0, 29.97, 30.0, 50.0, 59.94, 60.0, 120.0):
        for fl in np.arange(0.1, 5.0, 0.37):
            for fh in np.arange(fl + 0.2, fs / 2 - 0.01, 1.3):
                a = np.zeros(10, np.float32)
                b = np.zeros(10, np.float32)
                wa, wb = C.c_int(), C.c_int()
                O.ref().refh_iir_design(C.c_float(fl), C.c_float(fh), C.c_float(fs), ptr(a), C.byref(wa))
                O.orc().orc_iir_design(C.c_float(fl), C.c_float(fh), C.c_float(fs), ptr(b), C.byref(wb))
                bad += (not np.array_equal(a, b)) or wa.value != wb.value
    assert bad == 0


@pytest.mark.parametrize("batch", [0, 1])
def test_iir_run_bitexact(batch):
    fl, fh, fs = 0.5, 3.0, 30.0
    n, T = 4096, 12
    x = (np.random.default_rng(4).standard_normal((T, n)) * 0.05).astype(np.float32)
    out = np.zeros((T, n), np.float32)
    st = np.zeros((4, n), np.float32)
    O.ref().refh_iir_run(C.c_float(fl), C.c_float(fh), C.c_float(fs), ptr(x), T, n, 1, batch, ptr(out), ptr(st))
    c = np.zeros(10, np.float32)
    wu = C.c_int()
    O.orc().orc_iir_design(C.c_float(fl), C.c_float(fh), C.c_float(fs), ptr(c), C.byref(wu))
    s = np.zeros(4 * n, np.float32)
    O.orc().orc_iir_preseed_f32(ptr(c), ptr(np.ascontiguousarray(x[0])), n, wu.value, ptr(s))
    for t in range(T):
        y = np.zeros(n, np.float32)
        O.orc().orc_iir_step_f32(ptr(c), ptr(np.ascontiguousarray(x[t])), n, ptr(s), ptr(y))
        assert np.array_equal(y, out[t]), t
    assert np.array_equal(s.reshape(n, 4).T, st)


def clip(w, h, n, seed, color=False):
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 256, (h, w, 3)).astype(np.float32)
    if not color:
        base[..., 1] = base[..., 2] = base[..., 0]
    return [np.clip(base + rng.normal(0, 6, base.shape) + 30 * np.sin(t * 0.7), 0, 255).astype(np.uint8)
            for t in range(n)]


MODES = {
    "iir": dict(use_iir=True),
    "iir_color": dict(use_iir=True, color=True),
    "window": dict(use_iir=False),
    "window_bilateral_color": dict(use_iir=False, bilateral=True, color=True),
    "gauss": dict(use_iir=True, edge_aware=False),
}


@pytest.mark.parametrize("mode", sorted(MODES))
def test_pipeline_vs_reference(mode):
    """Coefficients and temporal outputs bit-exact (window-mode delta within 4 ulp of the
    coefficient scale: vectorised mean, libmvec expf); output bytes within 1 LSB when the
    oracle is given the reference's thresholds."""
    w, h, n = 64, 48, 12
    kw = dict(fps=30.0, amp=120.0, f_low=0.5, f_high=3.0, window=8, levels=3, **MODES[mode])
    ref = O.RefPipe(w, h, **kw)
    orc = O.OraclePipe(w, h, box_prefix=True, **kw)
    nch = 3 if kw.get("color") else 1
    worst = 0
    for t, fr in enumerate(clip(w, h, n, 11, kw.get("color", False))):
        r_out, r_thr = ref.frame(fr)
        o_out, _ = orc.frame(fr, thr_override=r_thr)
        worst = max(worst, int(np.abs(r_out.astype(int) - o_out.astype(int)).max()))
        passthrough = not kw["use_iir"] and t + 1 < kw["window"]
        if passthrough:
            assert np.array_equal(r_out, fr)
            continue
        for ch in range(nch):
            rc = O.flat6_to_planes(ref.buffer(0, ch), w, h, 3)
            assert np.array_equal(rc, orc.buffer(0, ch)), (t, ch)
            rf = O.flat6_to_planes(ref.buffer(1, ch), w, h, 3)
            of = orc.buffer(1, ch)
            if kw["use_iir"]:
                assert np.array_equal(rf, of), (t, ch)
            else:
                # the mean may differ by an ulp of the samples (vectorised sum order)
                scale = np.abs(rc).max()
                assert np.abs(rf - of).max() <= 4 * np.finfo(np.float32).eps * scale, (t, ch)
    assert worst <= 1, worst


@pytest.mark.parametrize("color,amp,fps,fl,fh", [(False, 120.0, 30.0, 0.5, 3.0), (True, 120.0, 30.0, 0.5, 3.0),
                                                 (False, 50.0, 60.0, 0.8, 2.0)])
def test_amplified_bitexact_vs_reference(color, amp, fps, fl, fh):
    """Amplified coefficients (src/wema.c:714-744 with the fade-in of 830-839) bit-identical
    to the compiled reference's over the 15 fade-in frames and past them, given its
    thresholds: the fade's FP32 order is the -Ofast build's (t = (g + 1) * fl(1/15),
    amp * (t * t)), and each masked coefficient is fma(alpha_l, y, c)."""
    w, h, n, levels = 64, 48, 18, 3
    kw = dict(fps=fps, amp=amp, f_low=fl, f_high=fh, levels=levels, color=color)
    ref = O.RefPipe(w, h, **kw)
    orc = O.OraclePipe(w, h, **kw)
    for t, fr in enumerate(clip(w, h, n, 21, color)):
        _, r_thr = ref.frame(fr)
        orc.frame(fr, thr_override=r_thr)
        for ch in range(3 if color else 1):
            ra = O.flat6_to_planes(ref.amplified(ch), w, h, levels)
            assert np.array_equal(ra, orc.buffer(2, ch)), (t, ch, int((ra != orc.buffer(2, ch)).sum()))


def test_threshold_sum_relative_error():
    """The reference's FP32 running sum vs the exact sum on a 1080p-sized pyramid."""
    w, h, levels = 1920, 1080, 5
    img = np.random.default_rng(8).random(w * h).astype(np.float32)
    nu = O.num_unique(w, h, levels)
    planes = np.zeros(nu, np.float32)
    O.orc().orc_dwt_fwd_f32(ptr(img), w, h, levels, ptr(planes), ptr(np.zeros(w * h, np.float32)))
    f6 = O.planes_to_flat6(planes, w, h, levels)
    t_ref = O.ref().refh_threshold(ptr(f6), f6.size)
    s = int(sum(int(O.orc().orc_abs_fixed(C.c_float(float(v)))) for v in planes[::97]))  # spot
    exact = float(np.abs(f6.astype(np.float64)).sum()) / f6.size * 0.2
    assert abs(t_ref - exact) / exact < 1e-4
    assert s > 0


def test_oracle_independent_of_loading_the_reference():
    """The compiled reference's start-up code sets FTZ/DAZ in whatever process loads it. The
    oracle sets that mode itself per call: in a fresh process, one oracle case gives the same
    bytes and thresholds before and after the reference library is loaded."""
    import subprocess
    import sys
    code = r'''
import sys, numpy as np
sys.path.insert(0, %r)
from oracle import oracle as O
rng = np.random.default_rng(5)
frames = rng.integers(0, 256, size=(6, 29, 37, 3), dtype=np.uint8)
def run():
    o = O.OraclePipe(37, 29, fps=30.0, amp=120.0, f_low=0.5, f_high=3.0, levels=3)
    outs = [o.frame(f) for f in frames]
    return np.stack([a for a, _ in outs]), np.stack([t for _, t in outs])
a0, t0 = run()
O.ref()   # loads libwema_ref.so (FTZ/DAZ set by its start-up code)
a1, t1 = run()
assert np.array_equal(a0, a1) and np.array_equal(t0, t1, equal_nan=True)
print("same")
''' % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "same" in out.stdout, out.stderr
