"""Committed fixtures of the compiled reference (tests/golden/pipeline_ref.npz, written by
tests/golden/make_golden.py from wema_process_frame, src/wema.c:746-952) against the oracle
(CPU) and the HIP pipeline (GPU, through the C ABI).

Bars, as for the reference itself:
  * DWT coefficients of the last frame: bit-exact;
  * temporal output of the last frame: bit-exact for the IIR (src/iir_filter.c:211-219);
    within 4 ulp of the coefficient scale for the sliding window, whose mean the reference
    sums in a compiler-vectorised order (src/temporal_filter.c:170-319);
  * output bytes: within 1 LSB, with the reference's own per-frame thresholds injected (its
    FP32 threshold sum is compiler-vectorised; the oracle's and the kernels' sum is exact);
  * the oracle's exact thresholds within 1e-4 relative of the reference's.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = make_golden.load()
F32_EPS = np.finfo(np.float32).eps


def _setting(p):
    kw = dict(p)
    return kw.pop("width"), kw.pop("height"), kw


def _check_temporal(got, want, use_iir, what):
    if use_iir:
        np.testing.assert_array_equal(got, want, err_msg=what)
    else:
        assert np.abs(got - want).max() <= 4 * F32_EPS * max(np.abs(want).max(), 1e-30), what


def _filled(kw, t):
    return kw.get("use_iir", True) or t + 1 >= kw["window"]


def test_fixture_cases_cover_every_mode():
    modes = {(c["params"].get("use_iir", True), c["params"].get("color", False),
              c["params"].get("bilateral", False), c["params"].get("edge_aware", True))
             for c in CASES.values()}
    assert (True, False, False, True) in modes          # luma IIR + guided filter
    assert (True, True, False, True) in modes           # colour IIR
    assert (False, False, False, True) in modes         # luma window
    assert (False, True, True, True) in modes           # colour window, bilateral
    assert (True, False, False, False) in modes         # Gaussian transfer
    assert any(c["params"]["width"] % 2 for c in CASES.values())   # odd parents (stale entries)


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_reproduces_reference_fixtures(case):
    c = CASES[case]
    w, h, kw = _setting(c["params"])
    orc = O.OraclePipe(w, h, box_prefix=True, **kw)
    worst = 0
    for t in range(c["in"].shape[0]):
        out, thr = orc.frame(c["in"][t], thr_override=c["thr"][t])
        worst = max(worst, int(np.abs(out.astype(int) - c["out"][t].astype(int)).max()))
        if not _filled(kw, t):
            assert np.array_equal(out, c["in"][t])            # window fill: passthrough
    assert worst <= 1, worst
    for ch in range(c["coef"].shape[0]):
        np.testing.assert_array_equal(orc.buffer(0, ch), c["coef"][ch], err_msg="coefficients")
        _check_temporal(orc.buffer(1, ch), c["filt"][ch], kw.get("use_iir", True), "temporal")
    # thresholds: the oracle's exact sum against the reference's FP32 sum
    orc2 = O.OraclePipe(w, h, **kw)
    for t in range(c["in"].shape[0]):
        _, thr = orc2.frame(c["in"][t])
        if not _filled(kw, t):
            continue
        for ch in range(c["coef"].shape[0]):
            r = float(c["thr"][t][ch])
            assert abs(float(thr[ch]) - r) <= 1e-4 * abs(r), (t, ch, thr[ch], r)


@pytest.mark.ref
def test_fixtures_are_what_the_compiled_reference_writes(tmp_path):
    out = str(tmp_path / "fresh.npz")
    subprocess.run([sys.executable, make_golden.__file__, "--out", out], check=True,
                   stdout=subprocess.DEVNULL)
    fresh = np.load(out, allow_pickle=False)
    z = np.load(make_golden.OUT, allow_pickle=False)
    assert sorted(fresh.files) == sorted(z.files)
    for k in z.files:
        a, b = fresh[k], z[k]
        assert a.dtype == b.dtype and a.shape == b.shape, k
        assert np.array_equal(a, b, equal_nan=a.dtype.kind == "f"), k


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_device_reproduces_reference_fixtures(case):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import wema_amd
    c = CASES[case]
    w, h, kw = _setting(c["params"])
    sess = wema_amd.Session(w, h, capture=True, **kw)
    d_in = torch.from_numpy(c["in"]).cuda()
    d_out = torch.empty_like(d_in)
    for t in range(c["in"].shape[0]):
        sess.set_thresholds(c["thr"][t][None])
        sess.process_device(d_in[t:t + 1], d_out_u8=d_out[t:t + 1], n=1)
    sess.sync()
    got = d_out.cpu().numpy()
    d = np.abs(got.astype(int) - c["out"].astype(int))
    assert d.max() <= 1, (int(d.max()), np.argwhere(d > 1)[:4])
    for ch in range(c["coef"].shape[0]):
        np.testing.assert_array_equal(sess.read(0, ch=ch), c["coef"][ch], err_msg="coefficients")
        _check_temporal(sess.read(1, ch=ch), c["filt"][ch], kw.get("use_iir", True), "temporal")
    sess.close()
