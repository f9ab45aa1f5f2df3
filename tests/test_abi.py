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


# header -> the library that implements it
LIBS = {"wema_dev.h": wema_amd.LIB_PATH,
        "wema.h": os.path.join(os.path.dirname(wema_amd.LIB_PATH), "libwema.so")}


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return {m.group(2) for m in re.finditer(r"(^|[\s*])([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
            if not m.group(2).startswith(("if", "sizeof"))}


def test_every_header_has_a_library():
    assert sorted(os.path.basename(h) for h in glob.glob(os.path.join(ROOT, "include", "*.h"))) == sorted(LIBS)


@pytest.mark.parametrize("header", sorted(LIBS))
def test_exports_every_declared_function(header):
    out = subprocess.run(["nm", "-D", "--defined-only", LIBS[header]], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    declared = declared_functions(header)
    missing = declared - exported
    assert not missing, sorted(missing)
    assert len(declared) >= {"wema.h": 14, "wema_dev.h": 40}[header]


@pytest.mark.parametrize("header", sorted(LIBS))
def test_header_compiles_as_c(header):
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Wpedantic", "-Werror", "-fsyntax-only",
                    "-x", "c", os.path.join(ROOT, "include", header)], check=True)


def test_c_host_links_and_rejects_bad_options():
    """The C11 host library loads next to the device library; its option checks run
    without a GPU (the device session opens at the first frame)."""
    host = C.CDLL(LIBS["wema.h"])
    ctx = (C.c_char * 256)()
    assert host.wema_init(ctx, 64, 48, C.c_float(30.0), C.c_float(50.0), C.c_float(0.5),
                          C.c_float(3.0), 8) == 0
    assert host.wema_init_iir(ctx) == 0
    assert host.wema_init(ctx, 64, 48, C.c_float(30.0), C.c_float(50.0), C.c_float(3.0),
                          C.c_float(2.0), 8) == 0
    assert host.wema_init_iir(ctx) == -1          # f_low >= f_high (src/iir_filter.c:74)
    assert host.wema_init(ctx, 64, 48, C.c_float(30.0), C.c_float(50.0), C.c_float(0.5),
                          C.c_float(3.0), 1) == -1  # window too small
    host.wema_last_error.restype = C.c_char_p
    assert b"window" in host.wema_last_error()


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
    # the largest frame is accepted, one more row is not
    big = wema_amd.Session.query_bytes(16384, 16384, 30.0, f_high=3.0)
    assert big > 16384 * 16384 * 16
    with pytest.raises(wema_amd.WemaError, match="exceeds"):
        wema_amd.Session.query_bytes(16384, 16385, 30.0, f_high=3.0)
    with pytest.raises(wema_amd.WemaError, match="exceeds"):
        wema_amd.Session.query_bytes(65536, 65536, 30.0, f_high=3.0)


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
