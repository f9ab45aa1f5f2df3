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
    text = re.sub(r"^\s*typedef[^;]*;", "", text, flags=re.M | re.S)   # (function pointer types)
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


class Frame(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("width", C.c_int), ("height", C.c_int),
                ("channels", C.c_int)]


@pytest.mark.parametrize("cap", ["0", None])
def test_frame_alloc_falls_back_to_pageable_memory(cap, monkeypatch):
    """frame_alloc past its pinned-byte cap (0 here), or where pinning fails (no GPU here):
    64-byte-aligned, zeroed pageable frames that frame_free releases."""
    if cap is None:
        monkeypatch.delenv("WEMA_PINNED_MAX_BYTES", raising=False)
    else:
        monkeypatch.setenv("WEMA_PINNED_MAX_BYTES", cap)
    host = C.CDLL(LIBS["wema.h"])
    host.frame_alloc.argtypes = [C.POINTER(Frame), C.c_int, C.c_int, C.c_int]
    host.frame_free.argtypes = [C.POINTER(Frame)]
    host.wema_frame_pinned.argtypes = [C.POINTER(Frame)]
    probe = wema_amd.lib.wema_dev_host_alloc(64)   # can this process pin memory at all?
    if probe:
        wema_amd.lib.wema_dev_host_free(probe)
    frames = [Frame() for _ in range(3)]
    for i, f in enumerate(frames):
        assert host.frame_alloc(C.byref(f), 33 + i, 17, 3) == 0
        addr = C.cast(f.data, C.c_void_p).value
        assert addr and addr % 64 == 0
        assert (f.width, f.height, f.channels) == (33 + i, 17, 3)
        a = np.ctypeslib.as_array(f.data, ((33 + i) * 17 * 3,))
        assert not a.any()
        a[:] = 0.5
        assert host.wema_frame_pinned(C.byref(f)) == (0 if cap == "0" or not probe else 1)
    for f in frames:
        host.frame_free(C.byref(f))
        assert not f.data
    assert host.frame_alloc(C.byref(frames[0]), 0, 17, 3) == -1


def _closed_form_levels(w, h):
    """dwt_compute_levels (src/dwt.c:588-596) with the shipped cap of 4: the smaller side is
    halved (integer division) while it is >= 16, at most 4 times, and at least one level is
    used: clamp(floor(log2(min(w, h))) - 3, 1, 4)."""
    m = min(w, h)
    return min(max(m.bit_length() - 1 - 3, 1), 4)


def test_auto_levels_closed_form_every_size():
    f = wema_amd.lib.wema_dev_auto_levels
    for m in range(1, 4097):
        for other in (m, m + 1, 4096, 16384):
            for w, h in ((m, other), (other, m)):
                want = _closed_form_levels(w, h)
                assert f(w, h) == want, (w, h, f(w, h), want)
                assert wema_amd.auto_levels(w, h) == want


@pytest.mark.ref
def test_auto_levels_equal_the_compiled_reference():
    """The reference's own dwt_compute_levels, compiled from its sources with the level cap
    raised to 8 (oracle/Makefile): capped at the shipped 4, it gives our count at every
    smaller side from 1 to 4096."""
    from oracle import oracle as O
    ref = O.ref()
    ref.dwt_compute_levels.argtypes = [C.c_int, C.c_int]
    for m in range(1, 4097):
        assert min(ref.dwt_compute_levels(m, 4096), 4) == wema_amd.lib.wema_dev_auto_levels(m, 4096), m
        assert min(ref.dwt_compute_levels(4096, m), 4) == wema_amd.lib.wema_dev_auto_levels(4096, m), m
