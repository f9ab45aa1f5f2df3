"""Pins the oracle with checks independent of the reference code: published CDF 9/7
filter taps, brute-force symmetric-extension convolution, perfect reconstruction,
scipy.signal's Butterworth design and sosfilt, brute-force box / guided filters, and
closed forms. CPU only."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import signal

from oracle import oracle as O

L = O.orc()
ptr = O.ptr

# ISO/IEC 15444-1 Table F.4: irreversible 9/7 analysis filters (DC gain 1 low, Nyquist
# gain 2 high), with the standard normalisation constant of the lifting form.
H0 = {0: 0.6029490182363579, 1: 0.2668641184428723, 2: -0.07822326652898785,
      3: -0.01686411844287495, 4: 0.02674875741080976}
G0 = {0: 1.115087052456994, 1: -0.5912717631142470, 2: -0.05754352622849957,
      3: 0.09127176311424948}
K_JPEG2000 = 1.230174104914001
K_REF = 1.149604398860241        # the reference's normalisation (src/dwt.c:35)


def fwd1d_f64(x):
    y = np.ascontiguousarray(x, np.float64).copy()
    L.orc_fwd1d_f64(ptr(y), len(y))
    return y


def inv1d_f64(x):
    y = np.ascontiguousarray(x, np.float64).copy()
    L.orc_inv1d_f64(ptr(y), len(y))
    return y


def ws_ext(x, i):
    """Whole-sample symmetric extension."""
    n = len(x)
    if n == 1:
        return x[0]
    p = 2 * n - 2
    i %= p
    return x[i] if i < n else x[p - i]


def conv_analysis(x):
    """Brute-force analysis with the published taps and symmetric extension."""
    n = len(x)
    lo = [sum(c * ws_ext(x, 2 * i + k) for k0, c in H0.items() for k in ([k0, -k0] if k0 else [0]))
          for i in range((n + 1) // 2)]
    hi = [sum(c * ws_ext(x, 2 * i + 1 + k) for k0, c in G0.items() for k in ([k0, -k0] if k0 else [0]))
          for i in range(n // 2)]
    return np.array(lo), np.array(hi)


@pytest.mark.parametrize("n", [9, 10, 16, 17, 33, 64, 101])
def test_lifting_equals_published_filters_with_symmetric_extension(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    y = fwd1d_f64(x)
    half = (n + 1) // 2
    lo, hi = conv_analysis(x)
    # the lifting form scales lows by 1/K and highs by K; the published taps correspond to
    # K_JPEG2000, so the reference's outputs are those scaled by K_JPEG2000/K and K/K_JPEG2000
    np.testing.assert_allclose(y[:half], lo * (K_JPEG2000 / K_REF), rtol=0, atol=1e-9)
    np.testing.assert_allclose(y[half:], hi * (K_REF / K_JPEG2000), rtol=0, atol=1e-9)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 8, 31, 64, 127])
def test_1d_perfect_reconstruction(n):
    x = np.random.default_rng(n).standard_normal(n)
    np.testing.assert_allclose(inv1d_f64(fwd1d_f64(x)), x, atol=1e-12)


def test_1d_short_signals_unchanged():
    x = np.array([0.7])
    assert fwd1d_f64(x)[0] == 0.7  # n < 2 is left as is (no scaling)


@pytest.mark.parametrize("w,h,levels", [(64, 48, 3), (37, 29, 3), (50, 33, 4), (17, 1, 2), (1, 9, 2)])
def test_2d_perfect_reconstruction_f64(w, h, levels):
    img = np.random.default_rng(w * h).random(w * h)
    nu = O.num_unique(w, h, levels)
    planes = np.zeros(nu)
    ll = np.zeros(w * h)
    L.orc_dwt_fwd_f64(ptr(img), w, h, levels, ptr(planes), ptr(ll))
    rec = np.zeros(w * h)
    L.orc_dwt_inv_f64(ptr(planes), ptr(ll), w, h, levels, ptr(rec))
    np.testing.assert_allclose(rec, img, atol=1e-12)


def test_constant_image_closed_form():
    """A constant image has zero details and LL = c * (K_JPEG2000 / K)^(2L) (the published
    lowpass has DC gain 1)."""
    w, h, levels, c = 64, 48, 3, 0.37
    img = np.full(w * h, c)
    planes = np.zeros(O.num_unique(w, h, levels))
    ll = np.zeros(w * h)
    L.orc_dwt_fwd_f64(ptr(img), w, h, levels, ptr(planes), ptr(ll))
    assert np.abs(planes).max() < 1e-12
    lw, lh = O.level_dims(w, h, levels)
    np.testing.assert_allclose(ll[:lw[-1] * lh[-1]], c * (K_JPEG2000 / K_REF) ** (2 * levels), rtol=1e-12)


def test_f32_dwt_close_to_f64():
    w, h, levels = 64, 48, 3
    img = np.random.default_rng(5).random(w * h)
    nu = O.num_unique(w, h, levels)
    p64, l64 = np.zeros(nu), np.zeros(w * h)
    L.orc_dwt_fwd_f64(ptr(img), w, h, levels, ptr(p64), ptr(l64))
    img32 = img.astype(np.float32)
    p32, l32 = np.zeros(nu, np.float32), np.zeros(w * h, np.float32)
    L.orc_dwt_fwd_f32(ptr(img32), w, h, levels, ptr(p32), ptr(l32))
    # coefficients agree within 1e-5 relative to the coefficient scale
    assert np.abs(p32 - p64).max() <= 1e-5 * np.abs(p64).max()


def design(fl, fh, fs):
    c = np.zeros(10, np.float32)
    wu = C.c_int()
    assert L.orc_iir_design(C.c_float(fl), C.c_float(fh), C.c_float(fs), ptr(c), C.byref(wu)) == 0
    return c, wu.value


@pytest.mark.parametrize("fl,fh,fs", [(0.5, 3.0, 30.0), (0.8, 2.0, 60.0), (0.3, 5.0, 30.0),
                                      (0.3, 5.0, 60.0), (1.0, 12.0, 25.0)])
def test_iir_design_matches_scipy_butter(fl, fh, fs):
    c, wu = design(fl, fh, fs)
    bh, ah = signal.butter(2, fl, "highpass", fs=fs)
    bl, al = signal.butter(2, fh, "lowpass", fs=fs)
    np.testing.assert_allclose(c[:3], bh, rtol=2e-6)
    np.testing.assert_allclose(c[3:5], ah[1:], rtol=2e-6)
    np.testing.assert_allclose(c[5:8], bl, rtol=2e-6, atol=1e-9)
    np.testing.assert_allclose(c[8:10], al[1:], rtol=2e-6)
    assert wu == max(30, int(7.0 * fs / (2 * math.pi * fl)))


def test_iir_rejections():
    c = np.zeros(10, np.float32)
    for fl, fh, fs in [(0.0, 3.0, 30.0), (3.0, 3.0, 30.0), (0.5, 15.0, 30.0)]:
        assert L.orc_iir_design(C.c_float(fl), C.c_float(fh), C.c_float(fs), ptr(c), None) == -1


def test_iir_f64_equals_sosfilt_with_preseed():
    fl, fh, fs = 0.5, 3.0, 30.0
    c = np.zeros(10)
    L.orc_iir_design_f64(fl, fh, fs, ptr(c))
    sos = np.array([[c[0], c[1], c[2], 1.0, c[3], c[4]], [c[5], c[6], c[7], 1.0, c[8], c[9]]])
    rng = np.random.default_rng(0)
    n, T, warm = 7, 40, 66
    x = rng.standard_normal((T, n))
    # preseed = the filter run `warm` times on frame 0 from zero state
    _, zf = signal.sosfilt(sos, np.repeat(x[:1], warm, axis=0), axis=0, zi=np.zeros((2, 2, n)))
    y_ref, _ = signal.sosfilt(sos, x, axis=0, zi=zf)
    st = np.zeros(n * 4)
    L.orc_iir_preseed_f64(ptr(c), ptr(np.ascontiguousarray(x[0])), n, warm, ptr(st))
    ys = []
    for t in range(T):
        y = np.zeros(n)
        L.orc_iir_step_f64(ptr(c), ptr(np.ascontiguousarray(x[t])), n, ptr(st), ptr(y))
        ys.append(y)
    np.testing.assert_allclose(np.array(ys), y_ref, atol=1e-12)


@pytest.mark.parametrize("fl,fh,fs,bound", [
    (0.5, 3.0, 30.0, 1e-5),    # the 320x240 / 1080p band
    (0.8, 2.0, 60.0, 2e-5),    # the 4K band
    (0.3, 5.0, 60.0, 1e-4),    # slowest pole of the multi-stream mix (radius ~0.978)
])
def test_iir_f32_close_to_f64(fl, fh, fs, bound):
    """FP32 band-pass (the reference's arithmetic, kept by the GPU) vs FP64 on the same FP32
    coefficients, norm-wise: max |err| / max |input|. Near the unit circle FP32 state
    rounding grows; the bound per band is the measured error with margin."""
    c32, warm = design(fl, fh, fs)
    c64 = np.zeros(10)
    L.orc_iir_design_f64(fl, fh, fs, ptr(c64))
    rng = np.random.default_rng(1)
    n, T = 64, 300
    x = (rng.standard_normal((T, n)) * 0.05 + np.sin(np.arange(T) / fs * 2 * np.pi)[:, None] * 0.02)
    x32 = x.astype(np.float32)
    s32 = np.zeros(4 * n, np.float32)
    s64 = np.zeros(4 * n)
    L.orc_iir_preseed_f32(ptr(c32), ptr(np.ascontiguousarray(x32[0])), n, warm, ptr(s32))
    L.orc_iir_preseed_f64(ptr(c32.astype(np.float64)), ptr(np.ascontiguousarray(x32[0], np.float64)), n, warm, ptr(s64))
    worst = 0.0
    for t in range(T):
        y32 = np.zeros(n, np.float32)
        y64 = np.zeros(n)
        L.orc_iir_step_f32(ptr(c32), ptr(np.ascontiguousarray(x32[t])), n, ptr(s32), ptr(y32))
        L.orc_iir_step_f64(ptr(c32.astype(np.float64)), ptr(np.ascontiguousarray(x32[t], np.float64)), n, ptr(s64), ptr(y64))
        worst = max(worst, np.abs(y32 - y64).max() / np.abs(x).max())
    assert worst < bound, worst


def test_abs_fixed_and_threshold():
    vals = np.array([0.0, 1.0, -0.5, 3e-10, 1e-12, 123.25, 2.0 ** -33, 7.5e-6], np.float32)
    for v in vals:
        q = L.orc_abs_fixed(C.c_float(float(v)))
        assert q == math.floor(abs(float(v)) * 2 ** 32 + 0.5), v   # round half up
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(10000) * 0.03).astype(np.float32)
    s = sum(L.orc_abs_fixed(C.c_float(float(v))) for v in x)
    exact = math.fsum(abs(float(v)) for v in x)
    assert abs(s / 2 ** 32 - exact) < len(x) * 2 ** -33
    T = L.orc_threshold_from_fixed(C.c_uint64(s), len(x))
    assert abs(T - 0.2 * exact / len(x)) <= 1e-7 * T


def test_fade_closed_form():
    """((g + 1) / 15)^2 amp for g in 0..14, 0 before, amp after, in the FP32 operation order
    of the reference's -Ofast build: t = (g + 1) * fl(1/15), amp * (t * t) (the compiled
    reference's amplified coefficients pin that order: test_amplified_bitexact_vs_reference)."""
    f32 = np.float32
    for amp in (120.0, 50.0, 37.5):
        for g in range(-2, 20):
            a = L.orc_fade_amp(C.c_float(amp), g)
            want = 0.0 if g < 0 else (amp * ((g + 1) / 15.0) ** 2 if g < 15 else amp)
            assert abs(a - want) <= 1e-6 * max(1.0, want), g
            if 0 <= g < 15:
                t = f32(g + 1) * f32(f32(1.0) / f32(15.0))
                assert a == f32(amp) * f32(t * t), (amp, g)
    # the division form, (amp t) t, differs at g = 2: the test tells the two apart
    assert L.orc_fade_amp(C.c_float(120.0), 2) != f32(f32(120.0) * f32(f32(3.0) / f32(15.0))) * f32(f32(3.0) / f32(15.0))


def box_brute(x, w, h, r=4):
    x = x.reshape(h, w)
    out = np.zeros_like(x, dtype=np.float64)
    for y in range(h):
        for xx in range(w):
            out[y, xx] = x[max(0, y - r):y + r + 1, max(0, xx - r):xx + r + 1].mean()
    return out.reshape(-1)


@pytest.mark.parametrize("prefix", [0, 1])
def test_box_and_guided_vs_brute_force(prefix):
    w, h = 23, 17
    rng = np.random.default_rng(9)
    I = rng.random(w * h).astype(np.float32)
    p = (rng.standard_normal(w * h) * 0.05).astype(np.float32)
    out = np.zeros(w * h, np.float32)
    L.orc_box_f32(ptr(I), ptr(out), w, h, 4, prefix)
    np.testing.assert_allclose(out, box_brute(I.astype(np.float64), w, h), atol=2e-6)
    q32 = np.zeros(w * h, np.float32)
    L.orc_guided_f32(ptr(I), ptr(p), ptr(q32), w, h, prefix)
    I64, p64 = I.astype(np.float64), p.astype(np.float64)
    mI, mp = box_brute(I64, w, h), box_brute(p64, w, h)
    cI, cIp = box_brute(I64 * I64, w, h), box_brute(I64 * p64, w, h)
    a = (cIp - mI * mp) / (cI - mI * mI + 0.01)
    b = mp - a * mI
    q = box_brute(a, w, h) * I64 + box_brute(b, w, h)
    np.testing.assert_allclose(q32, q, atol=2e-6)


def box_mean_f64(x, w, h, r=4):
    """Clipped (2r + 1)^2 window means in f64 from an integral image (exact enough in f64)."""
    x = x.reshape(h, w).astype(np.float64)
    S = np.zeros((h + 1, w + 1))
    S[1:, 1:] = x.cumsum(0).cumsum(1)
    ys, xs = np.arange(h), np.arange(w)
    y0, y1 = np.maximum(ys - r, 0), np.minimum(ys + r + 1, h)
    x0, x1 = np.maximum(xs - r, 0), np.minimum(xs + r + 1, w)
    tot = S[y1][:, x1] - S[y0][:, x1] - S[y1][:, x0] + S[y0][:, x0]
    area = np.outer(y1 - y0, x1 - x0)
    return (tot / area).reshape(-1)


@pytest.mark.parametrize("w,h", [(1, 1), (1, 9), (9, 1), (5, 3), (4, 7), (37, 29), (64, 48), (131, 67)])
def test_box_kernel_order_at_edge_sizes(w, h):
    """The kernel's box order (horizontal blocks of 4, vertical groups of 4 rows with prefix and
    suffix sums) at heights and widths below, at and past the window and the group size, on
    values in [0, 1] and on signed ones: within 2e-6 of the f64 means. A row or column taken
    from the wrong group, or one too many or too few, is off by far more."""
    rng = np.random.default_rng(w * 1000 + h)
    for x in (rng.random(w * h), rng.standard_normal(w * h) * 0.05):
        x = x.astype(np.float32)
        out = np.zeros(w * h, np.float32)
        assert L.orc_box_f32(ptr(x), ptr(out), w, h, 4, 0) == 0
        np.testing.assert_allclose(out, box_mean_f64(x, w, h), atol=2e-6)
    # only radius 4 has the kernel's order
    assert L.orc_box_f32(ptr(x), ptr(out), w, h, 3, 0) == -1


def test_fp32_entry_points_flush_denormals():
    """The reference runs with FTZ/DAZ (its -Ofast build), the kernels flush denormals: the
    oracle's FP32 entry points do too, whatever the calling thread's mode. A denormal input
    to the IIR step is read as zero, so the output and state stay exactly zero."""
    c, _ = np.zeros(10, np.float32), None
    O.orc().orc_iir_design(C.c_float(0.5), C.c_float(3.0), C.c_float(30.0), O.ptr(c), C.byref(C.c_int()))
    x = np.full(64, 0x000B_2A4C, np.uint32).view(np.float32)   # denormals (FTZ may be on here)
    st = np.zeros(4 * 64, np.float32)
    y = np.ones(64, np.float32)
    tiny = np.ones(4, np.float32) * np.float32(1e-38)
    before = bool((tiny * np.float32(0.01)).any())     # does this thread keep denormals?
    O.orc().orc_iir_step_f32(O.ptr(c), O.ptr(x), C.c_size_t(64), O.ptr(st), O.ptr(y))
    assert not y.any() and not st.any()
    # and the caller's mode is restored
    assert bool((tiny * np.float32(0.01)).any()) == before
