"""wema_amd: MI355X (gfx950) implementation of WEMA's per-frame amplification pipeline.

The compute path is libwema_dev.so (HIP kernels behind the C ABI in include/wema_dev.h).
This module binds it with ctypes and mirrors the reference's C API (src/wema.h:253-283):

    ctx = WemaContext(width, height, fps, amp, f_low, f_high, window)   # wema_init
    ctx.color_mode = True; ctx.init_color(); ctx.init_iir()             # as main.c:288-322
    out = ctx.process_frame(frame)                                      # wema_process_frame
    outs = ctx.batch_process(frames)                                    # wema_batch_process

Frames are numpy or torch arrays of shape (H, W, 3): uint8 RGB24, or float32 RGB in
[0, 1] as the reference's Frame (src/wema.h:28-33). torch CUDA tensors stay on the device;
host arrays are staged through the device. There is no CPU fallback: importing this
module without the built library raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwema_dev.so")


class WemaError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "wema_amd: %s is not built (run `python wema_amd/build.py` or "
            "__graft_entry__.build())" % LIB_PATH)
    return C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)


lib = _load()


class Params(C.Structure):
    """wema_dev_params (include/wema_dev.h)."""
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("fps", C.c_float), ("amp", C.c_float),
        ("f_low", C.c_float), ("f_high", C.c_float), ("window", C.c_int), ("levels", C.c_int),
        ("use_iir", C.c_int), ("color", C.c_int), ("edge_aware", C.c_int),
        ("bilateral", C.c_int), ("batch_path", C.c_int), ("max_batch", C.c_int),
        ("device", C.c_int),
    ]


class SynthParams(C.Structure):
    """wema_synth_params (include/wema_dev.h)."""
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("fps", C.c_float), ("seed", C.c_uint32),
        ("kind", C.c_int), ("color", C.c_int), ("ndiscs", C.c_int),
        ("disc_cx", C.c_float * 4), ("disc_cy", C.c_float * 4),
        ("disc_amp", C.c_float * 4), ("disc_freq", C.c_float * 4),
        ("disc_radius", C.c_float), ("disc_contrast", C.c_float),
        ("tint_r", C.c_float), ("tint_b", C.c_float), ("noise_sigma", C.c_float),
        ("shear_amp", C.c_float), ("shear_freq", C.c_float), ("cell", C.c_int),
        ("checker_contrast", C.c_float),
    ]


FRAME_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_long, C.POINTER(C.c_uint8), C.c_int, C.c_int)


class ServeConfig(C.Structure):
    """wema_serve_config (include/wema_dev.h)."""
    _fields_ = [
        ("device", C.c_int), ("duration_s", C.c_double), ("drain_s", C.c_double),
        ("arrival_rate", C.c_double), ("life_min_s", C.c_double), ("life_max_s", C.c_double),
        ("hbm_budget", C.c_size_t), ("seed", C.c_uint64), ("saturate", C.c_int),
        ("sessions", C.c_int), ("batched", C.c_int), ("use_graphs", C.c_int),
        ("max_sessions", C.c_int), ("fps_from_mix", C.c_int), ("warm_start", C.c_int),
        ("frames_per_session", C.c_long), ("warmup_rounds", C.c_int), ("host_io", C.c_int),
        ("on_frame", FRAME_CB), ("cb_user", C.c_void_p),
    ]


class ServeStats(C.Structure):
    """wema_serve_stats (include/wema_dev.h)."""
    _fields_ = [
        ("frames", C.c_long), ("mpix", C.c_double), ("wall_s", C.c_double),
        ("mpix_per_s", C.c_double), ("lat_p50_ms", C.c_double), ("lat_p99_ms", C.c_double),
        ("lat_max_ms", C.c_double), ("sessions_arrived", C.c_int), ("sessions_admitted", C.c_int),
        ("admission_waits", C.c_int), ("peak_sessions", C.c_int), ("peak_bytes", C.c_size_t),
        ("ticks", C.c_long), ("mean_frames_per_tick", C.c_double),
        ("host_ms_per_tick", C.c_double), ("graph_tick_frac", C.c_double),
        ("gpu_busy_frac", C.c_double), ("link_gbs", C.c_double),
    ]


_vp = C.c_void_p
_SIGS = {
    "wema_dev_params_default": (None, [C.POINTER(Params), C.c_int, C.c_int, C.c_float]),
    "wema_dev_query_bytes": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_size_t)]),
    "wema_dev_design_bandpass": (C.c_int, [C.c_float, C.c_float, C.c_float, _vp, _vp]),
    "wema_dev_open": (C.c_int, [C.POINTER(_vp), C.POINTER(Params)]),
    "wema_dev_close": (None, [_vp]),
    "wema_dev_last_error": (C.c_char_p, []),
    "wema_dev_geometry": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "wema_dev_iir_coeffs": (C.c_int, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_int)]),
    "wema_dev_frames": (C.c_long, [_vp]),
    "wema_dev_tile_rows": (C.c_int, [_vp]),
    "wema_dev_set_tile_rows": (C.c_int, [_vp, C.c_int]),
    "wema_dev_autotune": (None, [C.c_int]),
    "wema_dev_strip_rows": (C.c_int, [_vp, C.c_int, C.c_int]),
    "wema_dev_set_strip_rows": (C.c_int, [_vp, C.c_int]),
    "wema_dev_process_u8": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "wema_dev_process_u8_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "wema_dev_process_f32": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    "wema_dev_process_host_u8": (C.c_int, [_vp, _vp, _vp]),
    "wema_dev_sync": (C.c_int, [_vp]),
    "wema_dev_host_alloc": (_vp, [C.c_size_t]),
    "wema_dev_submit_host_u8": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(C.c_long)]),
    "wema_dev_wait_ticket": (C.c_int, [_vp, C.c_long]),
    "wema_dev_submit": (C.c_int, [_vp, _vp, _vp, C.c_int64]),
    "wema_dev_collect": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(_vp)]),
    "wema_dev_submit_f32": (C.c_int, [_vp, _vp, _vp, C.c_int64]),
    "wema_dev_collect_f32": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(_vp)]),
    "wema_dev_host_free": (None, [_vp]),
    "wema_dev_state_bytes": (C.c_int, [_vp, C.POINTER(C.c_size_t)]),
    "wema_dev_export_state": (C.c_int, [_vp, _vp, C.c_size_t]),
    "wema_dev_import_state": (C.c_int, [_vp, _vp, C.c_size_t]),
    "wema_dev_debug_capture": (C.c_int, [_vp, C.c_int]),
    "wema_dev_timing": (C.c_int, [_vp, C.c_int]),
    "wema_dev_timing_report": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_long),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "wema_dev_debug_thresholds": (C.c_int, [_vp, _vp, C.c_int]),
    "wema_dev_debug_set_thresholds": (C.c_int, [_vp, _vp, C.c_int]),
    "wema_dev_debug_read": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_size_t]),
    "wema_dev_debug_div_mid": (C.c_int, [_vp, _vp, _vp, _vp, C.c_long, _vp]),
    "wema_dev_debug_inject_error": (C.c_int, [C.c_int, C.c_int]),
    "wema_pool_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_size_t]),
    "wema_pool_destroy": (None, [_vp]),
    "wema_pool_add": (C.c_int, [_vp, C.POINTER(Params), C.POINTER(C.c_int)]),
    "wema_pool_remove": (C.c_int, [_vp, C.c_int]),
    "wema_pool_session": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "wema_pool_bytes_in_use": (C.c_size_t, [_vp]),
    "wema_pool_graphs": (C.c_int, [_vp, C.c_int]),
    "wema_pool_tick": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
    "wema_pool_sync": (C.c_int, [_vp]),
    "wema_pool_stream": (_vp, [_vp]),
    "wema_pool_timing": (C.c_int, [_vp, C.c_int]),
    "wema_pool_timing_report": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_long),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "wema_serve_run": (C.c_int, [C.POINTER(ServeConfig), C.POINTER(ServeStats)]),
    "wema_serve_plan": (C.c_int, [C.POINTER(ServeConfig), C.c_int, C.POINTER(Params),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "wema_pool_stats": (C.c_int, [_vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "wema_synth_preset": (C.c_int, [C.POINTER(SynthParams), C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint32]),
    "wema_synth_texture_host": (C.c_int, [C.POINTER(SynthParams), _vp]),
    "wema_synth_frame_host": (C.c_int, [C.POINTER(SynthParams), _vp, C.c_int, _vp]),
    "wema_synth_frames_device": (C.c_int, [C.POINTER(SynthParams), _vp, C.c_int, C.c_int, _vp, _vp]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def last_error():
    msg = lib.wema_dev_last_error()
    return msg.decode() if msg else ""


def _check(rc, what):
    if rc != 0:
        raise WemaError("%s: %s" % (what, last_error()))


def _ptr(a):
    """Address of a numpy array or torch tensor."""
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def _stream_handle():
    import torch
    return torch.cuda.current_stream().cuda_stream


class PinnedArray:
    """A numpy array over pinned host memory (wema_dev_host_alloc); close() releases it."""

    def __init__(self, shape, dtype=np.uint8):
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        self._p = lib.wema_dev_host_alloc(n)
        if not self._p:
            raise WemaError(last_error())
        self.array = np.frombuffer((C.c_uint8 * n).from_address(self._p), dtype=dtype).reshape(shape)

    def close(self):
        if self._p:
            self.array = None
            lib.wema_dev_host_free(self._p)
            self._p = None

    def __del__(self):
        self.close()


def inject_error(site, after=1):
    """Arms an injected HIP failure (wema_dev_debug_inject_error): 1 job-table upload, 2 stage
    launch, 3 host-frame upload; 0 disarms."""
    _check(lib.wema_dev_debug_inject_error(int(site), int(after)), "inject_error")


def autotune(enable=True, dwt=None):
    """Turns the autotuners on or off for sessions opened later: the guided filter's segment
    height (enable) and the DWT strip heights (dwt; the same as enable when not given)."""
    dwt = enable if dwt is None else dwt
    lib.wema_dev_autotune((1 if enable else 0) | (2 if dwt else 0))


def design_bandpass(f_low, f_high, fs):
    """iir_design_bandpass + warm-up (src/iir_filter.c:70-101, 150-151), on the host."""
    c = np.zeros(10, np.float32)
    w = C.c_int()
    _check(lib.wema_dev_design_bandpass(f_low, f_high, fs, c.ctypes.data, C.addressof(w)),
           "design_bandpass")
    return c, w.value


def auto_levels(width, height):
    """dwt_compute_levels (src/dwt.c:588-596)."""
    m, lv = min(width, height), 0
    while m >= 16 and lv < 4:
        m //= 2
        lv += 1
    return max(lv, 1)


class Session:
    """One stream session of the device pipeline (a wema_dev_t handle)."""

    def __init__(self, width, height, fps, amp=120.0, f_low=0.5, f_high=None, window=32,
                 levels=0, use_iir=True, color=False, edge_aware=True, bilateral=False,
                 batch_path=False, max_batch=1, device=0, capture=False):
        p = Params()
        lib.wema_dev_params_default(C.byref(p), int(width), int(height), float(fps))
        p.amp = amp
        p.f_low = f_low
        if f_high is not None:
            p.f_high = f_high
        p.window, p.levels, p.use_iir, p.color = window, levels, int(use_iir), int(color)
        p.edge_aware, p.bilateral, p.batch_path = int(edge_aware), int(bilateral), int(batch_path)
        p.max_batch, p.device = max_batch, device
        self.params = p
        self.width, self.height = int(width), int(height)
        h = C.c_void_p()
        _check(lib.wema_dev_open(C.byref(h), C.byref(p)), "wema_dev_open")
        self._h = h
        if capture:
            _check(lib.wema_dev_debug_capture(h, 1), "wema_dev_debug_capture")
        lv, nu, n6 = C.c_int(), C.c_size_t(), C.c_size_t()
        _check(lib.wema_dev_geometry(h, C.byref(lv), C.byref(nu), C.byref(n6)), "geometry")
        self.levels, self.n_unique, self.n6 = lv.value, nu.value, n6.value

    def close(self):
        if getattr(self, "_h", None):
            lib.wema_dev_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def query_bytes(width, height, fps, **kw):
        s = Session.__new__(Session)
        p = Params()
        lib.wema_dev_params_default(C.byref(p), int(width), int(height), float(fps))
        for k, v in kw.items():
            setattr(p, k, int(v) if isinstance(v, bool) else v)
        n = C.c_size_t()
        _check(lib.wema_dev_query_bytes(C.byref(p), C.byref(n)), "wema_dev_query_bytes")
        del s
        return n.value

    @property
    def frames(self):
        return lib.wema_dev_frames(self._h)

    @property
    def tile_rows(self):
        """Guided-filter segment height: rows per workgroup (picked by the autotuner; 0: each
        launch splits its strips itself)."""
        return lib.wema_dev_tile_rows(self._h)

    def set_tile_rows(self, rows):
        """Force the guided-filter segment height (a positive multiple of 4, or 0)."""
        _check(lib.wema_dev_set_tile_rows(self._h, int(rows)), "wema_dev_set_tile_rows")

    def strip_rows(self, inverse=False):
        """DWT strip height (pair rows) of each level, forward or inverse (0: the fixed rule;
        1: LDS tiles, 2: one output pair per thread, levels below full resolution)."""
        return [lib.wema_dev_strip_rows(self._h, l, int(inverse)) for l in range(self.levels)]

    def set_strip_rows(self, rows):
        """Force one DWT strip height (4, 8, 16 or 32 pair rows; 0: the fixed rule) for every level,
        or for the levels below full resolution 1: LDS tiles, 2: one output pair per thread
        (level 0 keeps the rule)."""
        _check(lib.wema_dev_set_strip_rows(self._h, int(rows)), "wema_dev_set_strip_rows")

    def iir_coeffs(self):
        c = (C.c_float * 10)()
        w = C.c_int()
        _check(lib.wema_dev_iir_coeffs(self._h, c, C.byref(w)), "iir_coeffs")
        return np.array(list(c), np.float32), w.value

    def process_device(self, d_in, d_out_u8=None, d_out_f32=None, n=1, stream=None):
        """Enqueue n frames on device buffers (torch CUDA tensors or raw addresses)."""
        st = stream if stream is not None else _stream_handle()
        a = lambda t: 0 if t is None else (t if isinstance(t, int) else t.data_ptr())
        if getattr(d_in, "dtype", None) is not None and "float" in str(d_in.dtype):
            rc = lib.wema_dev_process_f32(self._h, a(d_in), a(d_out_u8), a(d_out_f32), n, st)
        else:
            rc = lib.wema_dev_process_u8(self._h, a(d_in), a(d_out_u8), n, st)
        _check(rc, "wema_dev_process")

    def process_device_noalloc(self, d_in, d_out_u8, n=1, stream=None):
        """wema_dev_process_u8 with the kernels' per-call buffers in the library's static
        device scratch (no device allocation)."""
        st = stream if stream is not None else _stream_handle()
        _check(lib.wema_dev_process_u8_noalloc(self._h, d_in.data_ptr(), d_out_u8.data_ptr(), n, st),
               "wema_dev_process_u8_noalloc")

    def sync(self):
        _check(lib.wema_dev_sync(self._h), "wema_dev_sync")

    def process_host_u8(self, frame):
        frame = np.ascontiguousarray(frame, dtype=np.uint8)
        out = np.empty_like(frame)
        _check(lib.wema_dev_process_host_u8(self._h, _ptr(frame), _ptr(out)), "process_host_u8")
        return out

    def submit_host(self, h_in, h_out, n):
        """Asynchronous host pipeline: uploads n RGB24 frames from h_in, processes them and
        downloads them into h_out (arrays that stay alive until wait(ticket)); returns the
        ticket. Pinned arrays (PinnedArray) copy at full link speed."""
        t = C.c_long()
        _check(lib.wema_dev_submit_host_u8(self._h, _ptr(h_in), _ptr(h_out), int(n), C.byref(t)),
               "submit_host_u8")
        return t.value

    def wait(self, ticket):
        _check(lib.wema_dev_wait_ticket(self._h, int(ticket)), "wait_ticket")

    def submit(self, frame_in, frame_out, pts):
        """Queue one frame (host array: RGB24 uint8, or float32 RGB in [0, 1]) with its
        timestamp; its output goes to frame_out (both stay alive until collected)."""
        f32 = np.asarray(frame_in).dtype == np.float32
        fn = lib.wema_dev_submit_f32 if f32 else lib.wema_dev_submit
        _check(fn(self._h, _ptr(frame_in), _ptr(frame_out), int(pts)), "submit")

    def collect(self, f32=False):
        """Wait for the oldest queued frame: (pts, address of its output buffer)."""
        pts, out = C.c_int64(), _vp()
        fn = lib.wema_dev_collect_f32 if f32 else lib.wema_dev_collect
        _check(fn(self._h, C.byref(pts), C.byref(out)), "collect")
        return pts.value, out.value

    # ---- measurement
    def timing(self, enable):
        """Record HIP events between the launches of every call (on the call's stream)."""
        _check(lib.wema_dev_timing(self._h, int(enable)), "wema_dev_timing")

    def timing_report(self):
        """Per kernel kind: name, launches, total ms, algorithmic bytes (kinds never
        launched are left out)."""
        rows = []
        k = 0
        name = C.create_string_buffer(64)
        while True:
            la, ms, by = C.c_long(), C.c_double(), C.c_double()
            if lib.wema_dev_timing_report(self._h, k, name, 64, C.byref(la), C.byref(ms), C.byref(by)) != 0:
                break
            if la.value:
                rows.append(dict(name=name.value.decode(), launches=la.value, ms=ms.value, bytes=by.value))
            k += 1
        return rows

    # ---- temporal state (move a session to another process or GPU)
    def export_state(self):
        """The session's temporal state as bytes (wema_dev_export_state)."""
        n = C.c_size_t()
        _check(lib.wema_dev_state_bytes(self._h, C.byref(n)), "state_bytes")
        buf = np.zeros(n.value, np.uint8)
        _check(lib.wema_dev_export_state(self._h, _ptr(buf), n.value), "export_state")
        return buf.tobytes()

    def import_state(self, state):
        """Continue from a state exported by a session of the same geometry and filter."""
        buf = np.frombuffer(state, np.uint8).copy()
        _check(lib.wema_dev_import_state(self._h, _ptr(buf), buf.size), "import_state")

    # ---- test hooks
    def thresholds(self, n):
        t = np.zeros(3 * n, np.float32)
        _check(lib.wema_dev_debug_thresholds(self._h, _ptr(t), n), "debug_thresholds")
        return t.reshape(n, 3)

    def set_thresholds(self, thr):
        if thr is None:
            _check(lib.wema_dev_debug_set_thresholds(self._h, None, 0), "set_thresholds")
            return
        t = np.ascontiguousarray(thr, np.float32).reshape(-1, 3)
        _check(lib.wema_dev_debug_set_thresholds(self._h, _ptr(t), t.shape[0]), "set_thresholds")

    def read(self, which, frame=0, ch=0):
        n = self.n_unique * (4 if which == 3 else 1)
        out = np.zeros(n, np.float32)
        _check(lib.wema_dev_debug_read(self._h, which, frame, ch, _ptr(out), n), "debug_read")
        return out


# ---------------------------------------------------------------------------------------
# multi-stream pool and serving loop

def make_params(width, height, fps, amp=120.0, f_low=0.5, f_high=None, window=32, levels=0,
                use_iir=True, color=False, edge_aware=True, bilateral=False, batch_path=False,
                max_batch=1, device=0):
    p = Params()
    lib.wema_dev_params_default(C.byref(p), int(width), int(height), float(fps))
    p.amp, p.f_low = amp, f_low
    if f_high is not None:
        p.f_high = f_high
    p.window, p.levels, p.use_iir, p.color = window, levels, int(use_iir), int(color)
    p.edge_aware, p.bilateral, p.batch_path = int(edge_aware), int(bilateral), int(batch_path)
    p.max_batch, p.device = max_batch, device
    return p


class Pool:
    """Many stream sessions on one GPU, one launch per stage per tick (wema_pool_*)."""

    def __init__(self, hbm_budget, device=0, graphs=False):
        h = C.c_void_p()
        _check(lib.wema_pool_create(C.byref(h), device, int(hbm_budget)), "wema_pool_create")
        self._h = h
        self.device = device
        self.graphs(graphs)

    def graphs(self, enable):
        _check(lib.wema_pool_graphs(self._h, int(enable)), "wema_pool_graphs")

    def add(self, width, height, fps, **kw):
        p = make_params(width, height, fps, device=self.device, **kw)
        i = C.c_int()
        _check(lib.wema_pool_add(self._h, C.byref(p), C.byref(i)), "wema_pool_add")
        return i.value

    def remove(self, sid):
        _check(lib.wema_pool_remove(self._h, sid), "wema_pool_remove")

    @property
    def bytes_in_use(self):
        return lib.wema_pool_bytes_in_use(self._h)

    @property
    def stream(self):
        return lib.wema_pool_stream(self._h)

    def tick(self, ids, ins, outs, stream=None):
        """One frame per entry: device RGB24 buffers ins[i] -> outs[i] of session ids[i]; a
        session listed k times in a row has its next k frames processed in order."""
        n = len(ids)
        a_ids = (C.c_int * n)(*ids)
        a_in = (C.c_void_p * n)(*[t.data_ptr() for t in ins])
        a_out = (C.c_void_p * n)(*[t.data_ptr() for t in outs])
        _check(lib.wema_pool_tick(self._h, n, a_ids, a_in, a_out, stream), "wema_pool_tick")

    def sync(self):
        _check(lib.wema_pool_sync(self._h), "wema_pool_sync")

    def thresholds(self, sid):
        s = C.c_void_p()
        _check(lib.wema_pool_session(self._h, sid, C.byref(s)), "wema_pool_session")
        t = np.zeros(3, np.float32)
        _check(lib.wema_dev_debug_thresholds(s, _ptr(t), 1), "debug_thresholds")
        return t

    def close(self):
        if getattr(self, "_h", None):
            lib.wema_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SERVE_MIX = [(640, 480, 30.0), (1280, 720, 30.0), (1920, 1080, 60.0), (3840, 2160, 30.0)]
SERVE_CLASS_FRAMES = 4


def _serve_config(duration_s, arrival_rate, life, hbm_budget, batched, graphs, saturate, sessions,
                  drain_s, max_sessions, fps_from_mix, seed, device, warm_start, frames_per_session,
                  warmup_rounds, host_io, on_frame=None):
    return ServeConfig(device, duration_s, drain_s, arrival_rate, life[0], life[1], int(hbm_budget),
                       seed, int(saturate), sessions, int(batched), int(graphs), max_sessions,
                       int(fps_from_mix), int(warm_start), int(frames_per_session), int(warmup_rounds),
                       int(host_io), on_frame if on_frame is not None else FRAME_CB(), None)


def serve(duration_s=10.0, arrival_rate=4.0, life=(2.0, 20.0), hbm_budget=200e9, batched=True,
          graphs=True, saturate=False, sessions=0, drain_s=0.0, max_sessions=0,
          fps_from_mix=False, seed=0, device=0, warm_start=False, frames_per_session=0,
          warmup_rounds=0, host_io=False, on_frame=None):
    """Runs the serving loop (wema_serve_run) and returns its statistics as a dict.
    on_frame(session, frame, array) is called for every output frame (a (h, w, 3) uint8 view
    of host memory with host_io, valid during the call; without host_io, the device address
    as an int)."""
    cb = FRAME_CB()
    if on_frame is not None:
        def _cb(_user, session, frame, rgb, w, h):
            if host_io:
                arr = np.ctypeslib.as_array(rgb, (h, w, 3))
            else:
                arr = C.cast(rgb, C.c_void_p).value
            on_frame(session, frame, arr)
        cb = FRAME_CB(_cb)
    c = _serve_config(duration_s, arrival_rate, life, hbm_budget, batched, graphs, saturate, sessions,
                      drain_s, max_sessions, fps_from_mix, seed, device, warm_start, frames_per_session,
                      warmup_rounds, host_io, cb)
    st = ServeStats()
    _check(lib.wema_serve_run(C.byref(c), C.byref(st)), "wema_serve_run")
    return {k: getattr(st, k) for k, _ in ServeStats._fields_}


def serve_plan(duration_s=10.0, arrival_rate=4.0, life=(2.0, 20.0), saturate=False, sessions=0,
               fps_from_mix=False, seed=0, device=0, warm_start=False):
    """The sessions wema_serve_run admits for these settings, in arrival order: a list of
    (Params, arrival time in s, resolution class index into SERVE_MIX)."""
    c = _serve_config(duration_s, arrival_rate, life, 0, True, True, saturate, sessions, 0.0, 0,
                      fps_from_mix, seed, device, warm_start, 0, 0, False)
    n = lib.wema_serve_plan(C.byref(c), 0, None, None, None)
    if n < 0:
        raise WemaError(last_error())
    prm = (Params * max(n, 1))()
    arr = (C.c_double * max(n, 1))()
    cls = (C.c_int * max(n, 1))()
    lib.wema_serve_plan(C.byref(c), n, prm, arr, cls)
    return [(prm[i], arr[i], cls[i]) for i in range(n)]


# ---------------------------------------------------------------------------------------
# synthetic video

class Synth:
    """Deterministic synthetic clip (see wema_amd/csrc/synth.h)."""

    def __init__(self, preset, width, height, fps, seed_offset=0):
        self.p = SynthParams()
        _check(lib.wema_synth_preset(C.byref(self.p), preset, width, height, float(fps), seed_offset),
               "wema_synth_preset")
        self.width, self.height = width, height
        self.tex = np.zeros(width * height, np.uint16)
        _check(lib.wema_synth_texture_host(C.byref(self.p), _ptr(self.tex)), "texture")
        self._dtex = None

    def frame(self, t):
        out = np.empty((self.height, self.width, 3), np.uint8)
        _check(lib.wema_synth_frame_host(C.byref(self.p), _ptr(self.tex), t, _ptr(out)), "frame")
        return out

    def frames_device(self, t0, n, out=None):
        import torch
        if self._dtex is None:
            self._dtex = torch.from_numpy(self.tex.view(np.int16)).cuda()
        if out is None:
            out = torch.empty((n, self.height, self.width, 3), dtype=torch.uint8, device="cuda")
        _check(lib.wema_synth_frames_device(C.byref(self.p), self._dtex.data_ptr(), t0, n,
                                            out.data_ptr(), _stream_handle()), "frames_device")
        return out


# ---------------------------------------------------------------------------------------
# mirror of the reference C API (src/wema.h:253-283)

class WemaContext:
    """Same call sequence as the reference's WemaContext: wema_init, then the field writes
    of main.c:288-290, then init_color / init_iir / init_iir_color, then process_frame or
    batch_process. The device session is created lazily on the first frame, once all
    options are known."""

    def __init__(self, width, height, fps, amp_factor, f_low, f_high, temporal_window,
                 levels=0, device=0):
        self.width, self.height, self.fps = width, height, fps
        self.amp_factor, self.f_low, self.f_high = amp_factor, f_low, f_high
        self.temporal_window = temporal_window
        self.num_levels = levels if levels > 0 else auto_levels(width, height)
        self.edge_aware = True       # src/wema.c:371-373
        self.bilateral_temp = True
        self.color_mode = False
        self.use_iir = False
        self.device = device
        self._s = None
        self._batch = None

    def init_color(self):
        return 0

    def init_iir(self):
        self.use_iir = True
        return 0

    def init_iir_color(self):
        return 0

    def _session(self, batch_path=False, max_batch=1):
        key = (batch_path, max_batch)
        if self._s is None or self._key != key:
            if self._s is not None:
                raise WemaError("a context runs either the per-frame or the batch path")
            self._s = Session(self.width, self.height, self.fps, self.amp_factor, self.f_low,
                              self.f_high, self.temporal_window, self.num_levels, self.use_iir,
                              self.color_mode, self.edge_aware, self.bilateral_temp,
                              batch_path, max_batch, self.device)
            self._key = key
        return self._s

    def _run(self, frames, batch_path, max_batch):
        import torch
        s = self._session(batch_path, max_batch)
        x = frames if isinstance(frames, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(frames))
        on_host = not x.is_cuda
        x = x.cuda()
        n = x.shape[0]
        if x.dtype == torch.uint8:
            out = torch.empty_like(x)
            s.process_device(x, d_out_u8=out, n=n)
        else:
            x = x.float().contiguous()
            out = torch.empty_like(x)
            s.process_device(x, d_out_f32=out, n=n)
        if on_host:
            torch.cuda.current_stream().synchronize()
            return out.cpu().numpy()
        return out

    def process_frame(self, frame):
        """wema_process_frame (src/wema.c:746): one (H, W, 3) frame in, one out."""
        return self._run(frame[None], False, 1)[0]

    def batch_process(self, frames, batch_size=64):
        """wema_batch_process (src/wema_batch.c:352): (n, H, W, 3) frames, n <= batch_size."""
        if len(frames) > batch_size:
            raise WemaError("frame_count exceeds batch size")
        return self._run(frames, True, batch_size)

    def flush(self):
        """wema_flush (src/wema.c:954): no buffered frames remain."""
        return -1

    def close(self):
        if self._s is not None:
            self._s.close()
            self._s = None
