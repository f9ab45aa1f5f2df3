"""ctypes bindings of the checkers: oracle/liboracle.so (the C restatement) and
oracle/_ref/libwema_ref.so (the reference pipeline compiled from its own sources).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libwema_ref.so")
REF_SRC = "/root/reference/src"

_vp = C.c_void_p


def build(ref=True):
    """Compile liboracle.so, and the reference library when its sources are present."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref and os.path.isdir(REF_SRC):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def ptr(a):
    return a.ctypes.data_as(_vp)


class OrcParams(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("fps", C.c_float), ("amp", C.c_float),
                ("f_low", C.c_float), ("f_high", C.c_float), ("window", C.c_int),
                ("levels", C.c_int), ("use_iir", C.c_int), ("color", C.c_int),
                ("edge_aware", C.c_int), ("bilateral", C.c_int), ("box_prefix", C.c_int)]


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_LIB):
            build(ref=False)
        L = C.CDLL(ORACLE_LIB)
        L.orc_pipe_open.restype = _vp
        L.orc_pipe_open.argtypes = [C.POINTER(OrcParams)]
        L.orc_pipe_close.argtypes = [_vp]
        L.orc_pipe_frame.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        L.orc_pipe_buffer.restype = C.POINTER(C.c_float)
        L.orc_pipe_buffer.argtypes = [_vp, C.c_int, C.c_int]
        L.orc_pipe_state.restype = C.POINTER(C.c_float)
        L.orc_pipe_state.argtypes = [_vp, C.c_int]
        L.orc_pipe_warmup.argtypes = [_vp]
        L.orc_num_unique.restype = C.c_size_t
        L.orc_abs_fixed.restype = C.c_uint64
        L.orc_abs_fixed.argtypes = [C.c_float]
        L.orc_threshold_from_fixed.restype = C.c_float
        L.orc_threshold_from_fixed.argtypes = [C.c_uint64, C.c_size_t]
        L.orc_fade_amp.restype = C.c_float
        L.orc_fade_amp.argtypes = [C.c_float, C.c_int]
        L.orc_iir_design.argtypes = [C.c_float, C.c_float, C.c_float, _vp, _vp]
        L.orc_iir_design_f64.argtypes = [C.c_double, C.c_double, C.c_double, _vp]
        L.orc_iir_preseed_f32.argtypes = [_vp, _vp, C.c_size_t, C.c_int, _vp]
        L.orc_iir_step_f32.argtypes = [_vp, _vp, C.c_size_t, _vp, _vp]
        L.orc_iir_preseed_f64.argtypes = [_vp, _vp, C.c_size_t, C.c_int, _vp]
        L.orc_iir_step_f64.argtypes = [_vp, _vp, C.c_size_t, _vp, _vp]
        L.orc_ingest_f32.argtypes = [_vp, C.c_size_t, C.c_int, _vp]
        _orc = L
    return _orc


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_LIB)
        L.refh_open.restype = _vp
        L.refh_open.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.refh_close.argtypes = [_vp]
        L.refh_frame.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.refh_buffer.restype = C.POINTER(C.c_float)
        L.refh_buffer.argtypes = [_vp, C.c_int, C.c_int]
        L.refh_num_positions.restype = C.c_long
        L.refh_num_positions.argtypes = [_vp]
        L.refh_warmup.argtypes = [_vp]
        L.refh_batch.argtypes = [_vp, C.c_int, C.c_int, _vp, C.c_int, _vp, _vp]
        L.refh_amplified.argtypes = [_vp, C.c_int, _vp]
        L.refh_iir_state.argtypes = [_vp, C.c_int, _vp]
        L.refh_dwt_forward.restype = C.c_long
        L.refh_dwt_forward.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp]
        L.refh_dwt_inverse.restype = C.c_long
        L.refh_dwt_inverse.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp]
        L.refh_iir_design.argtypes = [C.c_float, C.c_float, C.c_float, _vp, _vp]
        L.refh_iir_run.argtypes = [C.c_float, C.c_float, C.c_float, _vp, C.c_int, C.c_size_t,
                                   C.c_int, C.c_int, _vp, _vp]
        L.refh_threshold.restype = C.c_float
        L.refh_threshold.argtypes = [_vp, C.c_size_t]
        L.refh_u8_to_f32.argtypes = [_vp, _vp, C.c_size_t]
        L.refh_f32_to_u8.argtypes = [_vp, _vp, C.c_size_t]
        L.refh_rgb_to_gray.argtypes = [_vp, C.c_int, C.c_int, _vp]
        L.refh_gray_to_rgb.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp]
        _ref = L
    return _ref


def level_dims(w, h, levels):
    lw, lh = [w], [h]
    for _ in range(levels):
        lw.append((lw[-1] + 1) // 2)
        lh.append((lh[-1] + 1) // 2)
    return lw, lh


def num_unique(w, h, levels):
    lw, lh = level_dims(w, h, levels)
    return 3 * sum(lw[i] * lh[i] for i in range(1, levels + 1))


def planes_to_flat6(planes, w, h, levels):
    """Unique planes -> the reference's 6-slot flat order."""
    lw, lh = level_dims(w, h, levels)
    out, off = [], 0
    for lv in range(levels):
        n = lw[lv + 1] * lh[lv + 1]
        lh_, hh, hl = (planes[off + k * n: off + (k + 1) * n] for k in range(3))
        out += [lh_, hh, hl, hl, hh, lh_]
        off += 3 * n
    return np.concatenate(out).astype(np.float32)


def flat6_to_planes(flat6, w, h, levels):
    lw, lh = level_dims(w, h, levels)
    out, off = [], 0
    for lv in range(levels):
        n = lw[lv + 1] * lh[lv + 1]
        out += [flat6[off: off + 3 * n]]
        off += 6 * n
    return np.concatenate(out).astype(np.float32)


class OraclePipe:
    """The oracle's FP32 pipeline (orc_pipe_*)."""

    def __init__(self, width, height, fps, amp=120.0, f_low=0.5, f_high=3.0, window=32,
                 levels=0, use_iir=True, color=False, edge_aware=True, bilateral=False,
                 box_prefix=False):
        self.p = OrcParams(width, height, fps, amp, f_low, f_high, window, levels, int(use_iir),
                           int(color), int(edge_aware), int(bilateral), int(box_prefix))
        self.h = orc().orc_pipe_open(C.byref(self.p))
        if not self.h:
            raise ValueError("invalid oracle parameters")
        self.width, self.height = width, height
        self.levels = levels if levels > 0 else _auto_levels(width, height)
        self.nu = num_unique(width, height, self.levels)
        self.nch = 3 if color else 1

    def frame(self, rgb, thr_override=None, want_f32=False):
        rgb = np.ascontiguousarray(rgb, np.uint8)
        out = np.empty_like(rgb)
        f32 = np.empty(rgb.shape, np.float32) if want_f32 else None
        thr = np.zeros(3, np.float32)
        ovr = None if thr_override is None else np.ascontiguousarray(thr_override, np.float32)
        rc = orc().orc_pipe_frame(self.h, ptr(rgb), ptr(out), ptr(f32) if want_f32 else None,
                                  ptr(ovr) if ovr is not None else None, ptr(thr))
        assert rc == 0
        return (out, thr, f32) if want_f32 else (out, thr)

    def buffer(self, which, ch=0):
        p = orc().orc_pipe_buffer(self.h, which, ch)
        return np.ctypeslib.as_array(p, (self.nu,)).copy()

    def state(self, ch=0):
        p = orc().orc_pipe_state(self.h, ch)
        return np.ctypeslib.as_array(p, (self.nu * 4,)).copy()

    def close(self):
        if self.h:
            orc().orc_pipe_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _auto_levels(w, h):
    m, lv = min(w, h), 0
    while m >= 16 and lv < 4:
        m //= 2
        lv += 1
    return max(lv, 1)


class RefPipe:
    """The compiled reference pipeline (wema_process_frame / wema_batch_process)."""

    def __init__(self, width, height, fps, amp=120.0, f_low=0.5, f_high=3.0, window=32,
                 levels=0, use_iir=True, color=False, edge_aware=True, bilateral=False):
        self.h = ref().refh_open(width, height, fps, amp, f_low, f_high, window, levels,
                                 int(use_iir), int(color), int(edge_aware), int(bilateral))
        if not self.h:
            raise ValueError("reference init failed")
        self.width, self.height = width, height
        self.n6 = ref().refh_num_positions(self.h)

    def frame(self, rgb, want_f32=False):
        rgb = np.ascontiguousarray(rgb, np.uint8)
        out = np.empty_like(rgb)
        f32 = np.empty(rgb.shape, np.float32)
        thr = np.zeros(3, np.float32)
        assert ref().refh_frame(self.h, ptr(rgb), ptr(out), ptr(f32), ptr(thr)) == 0
        return (out, thr, f32) if want_f32 else (out, thr)

    def batch(self, frames, batch_size=64, threads=0):
        frames = np.ascontiguousarray(frames, np.uint8)
        n = frames.shape[0]
        out = np.empty_like(frames)
        thr = np.zeros(n, np.float32)
        got = ref().refh_batch(self.h, batch_size, threads, ptr(frames), n, ptr(out), ptr(thr))
        assert got == n
        return out, thr

    def buffer(self, which, ch=0):
        p = ref().refh_buffer(self.h, which, ch)
        return np.ctypeslib.as_array(p, (self.n6,)).copy()

    def amplified(self, ch=0):
        out = np.zeros(self.n6, np.float32)
        assert ref().refh_amplified(self.h, ch, ptr(out)) == 0
        return out

    def iir_state(self, ch=0):
        out = np.zeros(4 * (self.n6), np.float32)
        assert ref().refh_iir_state(self.h, ch, ptr(out)) == 0
        return out.reshape(4, self.n6)

    def close(self):
        if self.h:
            ref().refh_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
