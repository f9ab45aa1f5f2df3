"""The C ABI library: it loads, exports every function include/*.h declares, and its
host-only entry points behave (no GPU needed). CPU only."""
import ctypes as C
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import wema_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(wema_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_exports_every_declared_function():
    out = subprocess.run(["nm", "-D", "--defined-only", wema_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = declared_functions() - exported
    assert not missing, sorted(missing)
    assert len(declared_functions()) >= 20


def test_header_compiles_as_c():
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-fsyntax-only", "-x", "c",
                    os.path.join(ROOT, "include", "wema_dev.h")], check=True)


def test_params_default_follows_cli_defaults():
    p = wema_amd.Params()
    wema_amd.lib.wema_dev_params_default(C.byref(p), 640, 480, 30.0)
    assert (p.amp, p.f_low, p.window, p.use_iir, p.edge_aware, p.bilateral) == (120.0, 0.5, 32, 1, 1, 0)
    assert abs(p.f_high - (15.0 - 0.0001)) < 1e-5   # main.c:248-250 clamp of the default


def test_query_bytes_and_rejections():
    b = wema_amd.Session.query_bytes(1920, 1080, 30.0, f_high=3.0, color=True, levels=5)
    assert 40e6 < b < 400e6
    b4 = wema_amd.Session.query_bytes(3840, 2160, 60.0, f_low=0.8, f_high=2.0, levels=6, max_batch=8)
    assert b4 > b
    with pytest.raises(wema_amd.WemaError, match="band"):
        wema_amd.Session.query_bytes(640, 480, 30.0, f_low=3.0, f_high=2.0)
    with pytest.raises(wema_amd.WemaError, match="window"):
        wema_amd.Session.query_bytes(640, 480, 30.0, use_iir=0, window=2)
    with pytest.raises(wema_amd.WemaError, match="max_batch"):
        wema_amd.Session.query_bytes(640, 480, 30.0, max_batch=0)


def test_design_matches_oracle():
    from oracle import oracle as O
    for fl, fh, fs in [(0.5, 3.0, 30.0), (0.8, 2.0, 60.0), (0.3, 5.0, 60.0), (1.3, 4.1, 29.97)]:
        c, wu = wema_amd.design_bandpass(fl, fh, fs)
        o = np.zeros(10, np.float32)
        ow = C.c_int()
        O.orc().orc_iir_design(C.c_float(fl), C.c_float(fh), C.c_float(fs), O.ptr(o), C.byref(ow))
        assert np.array_equal(c, o) and wu == ow.value


def test_synth_presets_are_deterministic():
    a = wema_amd.Synth(1, 320, 240, 30.0).frame(7)
    b = wema_amd.Synth(1, 320, 240, 30.0).frame(7)
    assert np.array_equal(a, b)
    c = wema_amd.Synth(1, 320, 240, 30.0, seed_offset=1).frame(7)
    assert not np.array_equal(a, c)
    assert np.array_equal(a[..., 0], a[..., 1])          # grey scene: R = G = B
    col = wema_amd.Synth(2, 192, 108, 30.0).frame(0)
    assert not np.array_equal(col[..., 0], col[..., 2])  # tinted colour scene
    chk = wema_amd.Synth(3, 384, 216, 60.0).frame(0)
    assert 100 < chk.mean() < 156


def test_synth_disc_moves():
    s = wema_amd.Synth(1, 320, 240, 30.0)
    # the disc edge moves by <= 0.2 px; frames differ near its boundary and by noise elsewhere
    f0, f5 = s.frame(0).astype(int), s.frame(5).astype(int)
    assert np.abs(f0 - f5).max() > 0
