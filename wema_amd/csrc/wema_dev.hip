// Host side of the C ABI (include/wema_dev.h): session state in HBM, per-call job tables,
// stage launches, IIR design, synthetic frames.
#include <hip/hip_runtime.h>

#include <cmath>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/wema_dev.h"
#include "synth.h"
#include "wema_host.h"
#include "wema_jobs.h"

extern "C" {
hipError_t wema_launch_dwt_fwd(const DwtfJob *, int, int, int, hipStream_t);
hipError_t wema_launch_dwt_inv(const DwtiJob *, int, int, int, hipStream_t);
hipError_t wema_launch_dwt_fwd_tile(const DwtfJob *, int, int, hipStream_t);
hipError_t wema_launch_dwt_inv_tile(const DwtiJob *, int, int, hipStream_t);
hipError_t wema_launch_dwt_fwd_point(const DwtfJob *, int, int, hipStream_t);
hipError_t wema_launch_dwt_inv_point(const DwtiJob *, int, int, hipStream_t);
hipError_t wema_launch_dwt_inv_ycc(const DwtiJob *, int, int, hipStream_t);
hipError_t wema_launch_iir(const TemporalJob *, int, int, int, hipStream_t);
hipError_t wema_launch_win(int, const TemporalJob *, int, int, hipStream_t);
hipError_t wema_launch_egress(int, const EgressJob *, int, int, hipStream_t);
hipError_t wema_launch_synth(const SynthFrame *, const unsigned short *, unsigned char *, int,
                             int, int, hipStream_t);
hipError_t wema_launch_zero_u64(unsigned long long *, long long, hipStream_t);
hipError_t wema_launch_debug_div(const float *, const float *, float *, float *, long long, hipStream_t);
}

// ---------------------------------------------------------------------------------------
// errors

static thread_local char g_err[512] = "";

void wema_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *wema_dev_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------------------
// injected failures (test hook): the armed site fails at its `after`-th occurrence

namespace {
std::mutex g_inj_mu;
int g_inj_site = INJ_NONE;
int g_inj_left = 0;
}  // namespace

bool wema_inject(int site) {
    std::lock_guard<std::mutex> lk(g_inj_mu);
    if (g_inj_site != site || g_inj_site == INJ_NONE) return false;
    if (--g_inj_left > 0) return false;
    g_inj_site = INJ_NONE;
    return true;
}

extern "C" int wema_dev_debug_inject_error(int site, int after) {
    if (site < INJ_NONE || site > INJ_FRAME_UPLOAD || after < 1) { wema_set_error("bad injection site"); return -1; }
    std::lock_guard<std::mutex> lk(g_inj_mu);
    g_inj_site = site;
    g_inj_left = after;
    return 0;
}

#define HIP_OK(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            wema_set_error("%s failed: %s", #call, hipGetErrorString(e_));                 \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

// ---------------------------------------------------------------------------------------
// band-pass design, as the reference's -Ofast build evaluates iir_design_bandpass
// (src/iir_filter.c:36-101) and the warm-up count (src/iir_filter.c:150-151). This file is
// compiled with -ffp-contract=off; the fmaf calls are the contractions gcc makes there.

static void design_hp(float *c, float fc, float fs) {
    const float sqrt2 = 1.41421356237f;
    const float wc = tanf(((float)M_PI * fc) / fs);
    const float s2w = sqrt2 * wc;
    const float norm = 1.0f / (fmaf(wc, wc, s2w) + 1.0f);
    c[0] = norm;
    c[1] = -2.0f * norm;
    c[2] = norm;
    const float wm1 = fmaf(wc, wc, -1.0f);
    c[3] = (wm1 + wm1) * norm;
    c[4] = (fmaf(wc, wc, 1.0f) - s2w) * norm;
}

static void design_lp(float *c, float fc, float fs) {
    const float sqrt2 = 1.41421356237f;
    const float wc = tanf(((float)M_PI * fc) / fs);
    const float wc2 = wc * wc;
    const float norm = 1.0f / (fmaf(sqrt2, wc, wc2) + 1.0f);
    c[0] = wc2 * norm;
    c[1] = c[0] + c[0];
    c[2] = c[0];
    const float wm1 = wc2 - 1.0f;
    c[3] = (wm1 + wm1) * norm;
    c[4] = fmaf(-wc, sqrt2, wc2 + 1.0f) * norm;
}

int wema_design_bandpass(float f_low, float f_high, float fs, float *c10, int *warmup) {
    if (!(f_low < f_high) || !(f_low > 0) || !(f_high < fs * 0.5f)) return -1;
    design_hp(c10, f_low, fs);
    design_lp(c10 + 5, f_high, fs);
    const float k = 7.0f / (2.0f * (float)M_PI);
    int wu = (int)((fs * k) / f_low);
    *warmup = wu < 30 ? 30 : wu;
    return 0;
}

extern "C" int wema_dev_design_bandpass(float f_low, float f_high, float fs, float *c10,
                                        int *warmup) {
    int wu = 0;
    if (!c10 || wema_design_bandpass(f_low, f_high, fs, c10, &wu) < 0) {
        wema_set_error("invalid band %g..%g Hz at %g fps", f_low, f_high, fs);
        return -1;
    }
    if (warmup) *warmup = wu;
    return 0;
}

// Fade-in after the preseed (src/wema.c:830-839, src/wema_batch.c:426-435).
float wema_fade_amp(float amp, long global_frame) {
    const int fade_in_frames = 15;
    if (global_frame < 0) return 0.0f;
    if (global_frame < fade_in_frames) {
        // the reference's -Ofast build evaluates amp * t * t, t = (g + 1) / 15, as
        // amp * (t * t) with t = (g + 1) * fl(1/15) (reciprocal math and reassociation)
        const float t = (float)(global_frame + 1) * (1.0f / (float)fade_in_frames);
        const float tt = t * t;
        return amp * tt;
    }
    return amp;
}

extern "C" int wema_dev_auto_levels(int width, int height) { return wema_auto_levels(width, height); }

int wema_auto_levels(int width, int height) {
    // dwt_compute_levels (src/dwt.c:588-596), shipped cap of 4 levels
    int m = width < height ? width : height, levels = 0;
    while (m >= 16 && levels < 4) {
        m /= 2;
        levels++;
    }
    return levels > 0 ? levels : 1;
}

// ---------------------------------------------------------------------------------------
// geometry and memory plan

void wema_geometry_init(WemaGeom *g, const wema_dev_params *p) {
    memset(g, 0, sizeof(*g));
    g->W = p->width;
    g->H = p->height;
    g->L = p->levels > 0 ? p->levels : wema_auto_levels(p->width, p->height);
    g->nch = p->color ? 3 : 1;
    g->P = (size_t)g->W * g->H;
    g->lw[0] = g->W;
    g->lh[0] = g->H;
    g->plane_off[0] = 0;
    g->ll_off[0] = 0;
    for (int l = 0; l < g->L; l++) {
        g->lw[l + 1] = (g->lw[l] + 1) / 2;
        g->lh[l + 1] = (g->lh[l] + 1) / 2;
        const size_t n = (size_t)g->lw[l + 1] * g->lh[l + 1];
        g->plane_off[l + 1] = g->plane_off[l] + 3 * n;
        g->ll_off[l + 1] = g->ll_off[l] + n;
    }
    g->nu = g->plane_off[g->L];
    g->nus = (g->nu + 63) / 64 * 64;
    g->ll_total = g->ll_off[g->L];
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Byte layout of one session's arena.
struct Layout {
    size_t coef, ampc, filt, state, ll, rec, acc, thr, thr_ovr, tw, total;
};

static void plan_layout(const WemaGeom &g, const wema_dev_params &p, int with_filt, Layout *L) {
    const size_t B = (size_t)p.max_batch, nch = g.nch, nu = g.nus;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const int ring_slots = p.use_iir ? 0 : p.window + p.max_batch - 1;
    L->coef = take(sizeof(float) * nu * nch * (p.use_iir ? B : (size_t)ring_slots));
    L->ampc = take(p.use_iir && !with_filt ? 0 : sizeof(float) * nu * nch * B);
    L->filt = take(with_filt ? sizeof(float) * nu * nch * B : 0);
    L->state = take(p.use_iir ? sizeof(float) * 4 * nu * nch : 0);
    L->ll = take(sizeof(float) * g.ll_total * nch * B);
    L->rec = take(sizeof(float) * g.P * nch * B);
    L->acc = take(sizeof(unsigned long long) * wema_acc_entries((int)nch, (int)B));
    L->thr = take(sizeof(float) * 3 * B);
    L->thr_ovr = take(sizeof(float) * 3 * B);
    L->tw = take(sizeof(float) * 256);
    L->total = off;
}

static int validate(const wema_dev_params *p) {
    if (p->width < 1 || p->height < 1) { wema_set_error("invalid frame size %dx%d", p->width, p->height); return -1; }
    if ((long long)p->width * p->height > WEMA_DEV_MAX_PIXELS) {
        wema_set_error("frame size %dx%d exceeds %ld pixels", p->width, p->height, WEMA_DEV_MAX_PIXELS);
        return -1;
    }
    if (!(p->fps > 0)) { wema_set_error("fps must be positive"); return -1; }
    if (!(p->amp > 0)) { wema_set_error("amplification factor must be positive"); return -1; }
    if (p->max_batch < 1 || p->max_batch > WEMA_MAX_BATCH) {
        wema_set_error("max_batch must be in 1..%d", WEMA_MAX_BATCH);
        return -1;
    }
    int L = p->levels > 0 ? p->levels : wema_auto_levels(p->width, p->height);
    if (L > WEMA_MAX_LEVELS) { wema_set_error("levels must be <= %d", WEMA_MAX_LEVELS); return -1; }
    if (!p->use_iir && (p->window < 4 || p->window > 256)) {
        // src/main.c:135-143
        wema_set_error("temporal window must be in 4..256");
        return -1;
    }
    if (p->use_iir) {
        float c[10];
        int wu;
        if (wema_design_bandpass(p->f_low, p->f_high, p->fps, c, &wu) < 0) {
            // iir_design_bandpass rejects these (src/iir_filter.c:74)
            wema_set_error("invalid band %g..%g Hz at %g fps", p->f_low, p->f_high, p->fps);
            return -1;
        }
    }
    return 0;
}

extern "C" void wema_dev_params_default(wema_dev_params *p, int width, int height, float fps) {
    memset(p, 0, sizeof(*p));
    p->width = width;
    p->height = height;
    p->fps = fps;
    p->amp = 120.0f;           // src/main.c:66
    p->f_low = 0.5f;
    p->f_high = fps / 2.0f - 0.0001f;   // default 1048576 clamped to Nyquist (main.c:68, 248-250)
    p->window = 32;            // WEMA_TEMPORAL_WINDOW_DEF
    p->levels = 0;
    p->use_iir = 1;
    p->color = 0;
    p->edge_aware = 1;
    p->bilateral = 0;
    p->batch_path = 0;
    p->max_batch = 1;
    p->device = 0;
}

extern "C" int wema_dev_query_bytes(const wema_dev_params *p, size_t *bytes) {
    if (!p || !bytes) { wema_set_error("null argument"); return -1; }
    if (validate(p) < 0) return -1;
    WemaGeom g;
    wema_geometry_init(&g, p);
    Layout L;
    plan_layout(g, *p, 0, &L);
    *bytes = L.total + sizeof(DwtfJob) * 4096;
    return 0;
}

// ---------------------------------------------------------------------------------------
// static scratch of wema_dev_process_u8_noalloc: the per-call buffers of the kernels
// (coefficient planes, LL pyramid, reconstructions, guided-filter a and b, threshold
// accumulators) for callers that cannot allocate device memory. Each GPU the library runs
// on holds its own copy until the process ends. In host code `scratch` names a host-side
// stand-in: its device address on each GPU comes from hipGetSymbolAddress, taken once per
// device when a session is opened there, and only that address is ever handed to kernels
// or copies.

static __device__ __align__(256) unsigned char scratch[WEMA_DEV_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_done[kMaxDevices];
void *g_scratch_ptr[kMaxDevices];
char g_scratch_err[kMaxDevices][160];

// With `device` current: looks up the device address of `scratch` the first time.
void scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    g_scratch_done[device] = true;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        snprintf(g_scratch_err[device], sizeof(g_scratch_err[device]),
                 "hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
        (void)hipGetLastError();
        return;
    }
    g_scratch_ptr[device] = ptr;
}

// The scratch's device address on `device` (null, with the error set, if there is none).
void *scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) {
        wema_set_error("device %d out of range for the static scratch", device);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device])
        wema_set_error("%s", g_scratch_done[device] ? g_scratch_err[device] : "no session opened on this device");
    return g_scratch_ptr[device];
}
}  // namespace

// ---------------------------------------------------------------------------------------
// session

int wema_session_alloc(wema_dev *s, int with_filt);

int wema_dev_open_ex(wema_dev_t **out, const wema_dev_params *p, bool pooled) {
    if (!out || !p) { wema_set_error("null argument"); return -1; }
    *out = nullptr;
    if (validate(p) < 0) return -1;
    wema_dev *s = new wema_dev();
    s->pooled = pooled;
    s->p = *p;
    wema_geometry_init(&s->g, p);
    if (p->use_iir) wema_design_bandpass(p->f_low, p->f_high, p->fps, s->c10, &s->warmup);
    s->ring_slots = p->use_iir ? 0 : p->window + p->max_batch - 1;
    if (wema_session_alloc(s, 0) < 0) {
        wema_dev_close(s);
        return -1;
    }
    scratch_setup(p->device);   // (the session's device is current)
    if (wema_dwt_tune(s) < 0 || wema_gf_tune(s) < 0) {
        wema_dev_close(s);
        return -1;
    }
    *out = s;
    return 0;
}

extern "C" int wema_dev_open(wema_dev_t **out, const wema_dev_params *p) {
    return wema_dev_open_ex(out, p, false);
}

int wema_session_alloc(wema_dev *s, int with_filt) {
    if (hipSetDevice(s->p.device) != hipSuccess) {
        wema_set_error("hipSetDevice(%d) failed", s->p.device);
        return -1;
    }
    if (!s->stream && !s->pooled) HIP_OK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    Layout L;
    plan_layout(s->g, s->p, with_filt, &L);
    if (s->arena) {
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipFree(s->arena));
        s->arena = nullptr;
    }
    HIP_OK(hipMalloc(&s->arena, L.total));
    HIP_OK(hipMemsetAsync(s->arena, 0, L.total, s->stream));
    // the clear runs on the session's non-blocking stream, which the synchronous copies
    // below (null stream) do not wait for: let it finish first, or it may zero them again
    HIP_OK(hipStreamSynchronize(s->stream));
    char *a = (char *)s->arena;
    s->coef = (float *)(a + L.coef);
    s->ampc = (s->p.use_iir && !with_filt) ? nullptr : (float *)(a + L.ampc);
    s->filt = with_filt ? (float *)(a + L.filt) : nullptr;
    s->state = s->p.use_iir ? (float4 *)(a + L.state) : nullptr;
    s->ll = (float *)(a + L.ll);
    s->rec = (float *)(a + L.rec);
    s->acc = (unsigned long long *)(a + L.acc);
    s->thr_d = (float *)(a + L.thr);
    s->thr_ovr_d = (float *)(a + L.thr_ovr);
    s->tw_d = (float *)(a + L.tw);
    s->arena_bytes = L.total;
    // time weights of the bilateral window (src/temporal_filter.c:210-218), as their base-2
    // logarithms: the kernels fold them into the exponent of the value weight
    if (!s->p.use_iir) {
        const int ws = s->p.window;
        std::vector<float> ltw(256, 0.0f);
        const float ts = (float)ws * 0.15f;
        const float inv = 1.0f / (2.0f * ts * ts);
        for (int t = 0; t < ws; t++) {
            const float d = (float)(t - ws / 2);
            ltw[t] = (float)std::log2((double)expf(-(d * d) * inv));
        }
        HIP_OK(hipMemcpy(s->tw_d, ltw.data(), 256 * sizeof(float), hipMemcpyHostToDevice));
    }
    s->with_filt = with_filt;
    for (int k = 0; k < 2 && !s->pooled; k++) {
        // pooled sessions run through the pool's job tables
        if (!s->stage_h[k]) HIP_OK(hipHostMalloc(&s->stage_h[k], WEMA_STAGE_BYTES, hipHostMallocDefault));
        if (!s->stage_d[k]) HIP_OK(hipMalloc(&s->stage_d[k], WEMA_STAGE_BYTES));
        if (!s->stage_ev[k]) HIP_OK(hipEventCreateWithFlags(&s->stage_ev[k], hipEventDisableTiming));
    }
    HIP_OK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" void wema_dev_close(wema_dev_t *s) {
    if (!s) return;
    (void)hipSetDevice(s->p.device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (int k = 0; k < 2; k++) {
        if (s->stage_h[k]) (void)hipHostFree(s->stage_h[k]);
        if (s->stage_d[k]) (void)hipFree(s->stage_d[k]);
        if (s->stage_ev[k]) (void)hipEventDestroy(s->stage_ev[k]);
    }
    if (s->arena) (void)hipFree(s->arena);
    if (s->io_in) (void)hipFree(s->io_in);
    if (s->io_out) (void)hipFree(s->io_out);
    if (s->io_in_f32) (void)hipFree(s->io_in_f32);
    if (s->io_out_f32) (void)hipFree(s->io_out_f32);
    for (auto &a : s->aslot) {
        if (a.ev_out) (void)hipEventSynchronize(a.ev_out);
        if (a.d_in) (void)hipFree(a.d_in);
        if (a.d_out) (void)hipFree(a.d_out);
        if (a.d_in_f) (void)hipFree(a.d_in_f);
        if (a.d_out_f) (void)hipFree(a.d_out_f);
        if (a.ev_in) (void)hipEventDestroy(a.ev_in);
        if (a.ev_comp) (void)hipEventDestroy(a.ev_comp);
        if (a.ev_out) (void)hipEventDestroy(a.ev_out);
    }
    if (s->s_up) (void)hipStreamDestroy(s->s_up);
    if (s->s_down) (void)hipStreamDestroy(s->s_down);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

extern "C" int wema_dev_geometry(const wema_dev_t *s, int *levels, size_t *n_unique, size_t *n6) {
    if (!s) { wema_set_error("null session"); return -1; }
    if (levels) *levels = s->g.L;
    if (n_unique) *n_unique = s->g.nu;
    if (n6) *n6 = 2 * s->g.nu;
    return 0;
}

extern "C" int wema_dev_iir_coeffs(const wema_dev_t *s, float *c10, int *warmup) {
    if (!s || !s->p.use_iir) { wema_set_error("not an IIR session"); return -1; }
    if (c10) memcpy(c10, s->c10, sizeof(s->c10));
    if (warmup) *warmup = s->warmup;
    return 0;
}

extern "C" long wema_dev_frames(const wema_dev_t *s) { return s ? s->frames : -1; }

// ---------------------------------------------------------------------------------------
// job tables for n frames of one session, appended to a JobBuilder (also used by the pool)

void wema_build_jobs(wema_dev *s, JobBuilder &jb, const void *in, int in_f32, uint8_t *out_u8,
                     float *out_f32, int n, const CallBufs *bufs, const void *const *in_list,
                     uint8_t *const *out_list) {
    const WemaGeom &g = s->g;
    const wema_dev_params &p = s->p;
    const CallBufs b = bufs ? *bufs : CallBufs{s->coef, s->ampc, s->ll, s->rec, s->acc};
    const int nch = g.nch, L = g.L;
    const int acc_s = wema_acc_slots(n);   // Σ|c| slots per frame and channel
    const size_t P = g.P, nu = g.nu, nus = g.nus;
    const size_t in_stride = P * 3 * (in_f32 ? sizeof(float) : 1);
    const long frame0 = s->frames;
    // frame f's source and RGB24 output: consecutive frames from in / out_u8, or one pointer
    // per frame (in_list / out_list: a pool tick's frames of one session)
    auto in_of = [&](int f) -> const char * {
        return in_list ? (const char *)in_list[f] : (const char *)in + (size_t)f * in_stride;
    };
    auto out_u8_of = [&](int f) -> uint8_t * {
        return out_list ? out_list[f] : (out_u8 ? out_u8 + (size_t)f * P * 3 : nullptr);
    };

    // per-frame effective amplification (channel-major) and threshold overrides
    for (int c = 0; c < nch; c++)
        for (int f = 0; f < n; f++) {
            float a = p.use_iir ? wema_fade_amp(p.amp, frame0 + f) : p.amp;
            if (c > 0) a = a * 0.5f;   // chroma (src/wema.c:796-799, 898-901)
            jb.host_amp.push_back(a);
        }

    auto coef_ptr = [&](int f, int c) -> float * {
        if (p.use_iir) return b.coef + ((size_t)f * nch + c) * nus;
        const long gidx = frame0 + f;
        return s->coef + ((size_t)c * s->ring_slots + (size_t)(gidx % s->ring_slots)) * nus;
    };
    auto ll_ptr = [&](int f, int c, int l) -> float * {
        return b.ll + ((size_t)f * nch + c) * g.ll_total + g.ll_off[l];
    };
    auto rec_ptr = [&](int f, int c) -> float * { return b.rec + ((size_t)f * nch + c) * P; };

    // forward DWT levels
    for (int l = 0; l < L; l++) {
        const int w = g.lw[l], h = g.lh[l];
        const int tiles_x = wema_ceil_div(g.lw[l + 1], WEMA_TILE_PAIRS_X);
        // level 0: one job per frame for all channels (they share the RGB source; each wave
        // computes its strip of every channel); strip heights and block ranges are set when
        // the launch's table is packed
        const int jch = l == 0 ? nch : 1;
        for (int f = 0; f < n; f++)
            for (int c = 0; c < nch; c += jch) {
                DwtfJob j{};
                if (l == 0) {
                    j.src = in_of(f);
                    j.src_kind = in_f32 ? SRC_F32_Y : SRC_U8_Y;
                    j.prev = nullptr;
                } else {
                    j.src = ll_ptr(f, c, l - 1);
                    j.src_kind = SRC_F32_PLANE;
                    j.prev = coef_ptr(f, c) + g.plane_off[l - 1];
                }
                j.ll = ll_ptr(f, c, l);
                j.planes = coef_ptr(f, c) + g.plane_off[l];
                j.acc = b.acc + ((size_t)f * nch + c) * acc_s;
                j.acc_slots = acc_s;
                j.nch = jch;
                j.ll_cstride = jch > 1 ? (long long)(ll_ptr(f, 1, l) - ll_ptr(f, 0, l)) : 0;
                j.planes_cstride = jch > 1 ? (long long)(coef_ptr(f, 1) - coef_ptr(f, 0)) : 0;
                j.w = w;
                j.h = h;
                j.tiles_x = tiles_x;
                j.sp = s->sp_fwd[l];
                if ((j.sp == WEMA_SP_TILE || j.sp == WEMA_SP_POINT) && (l == 0 || w < 2 || h < 2))
                    j.sp = 0;   // (tiles and pairs: coarser levels of 2 x 2 samples or more)
                // input read once (RGB24 shared by the channels at level 0) + LL and three
                // subbands written per channel
                const double in_b = l == 0 ? 3.0 * w * h * (in_f32 ? 4 : 1) : 4.0 * w * h;
                const double out_b = 16.0 * g.lw[l + 1] * g.lh[l + 1] * jch;
                jb.add_dwtf(l, j, 0, in_b + out_b);
            }
    }

    // temporal stage: one job per channel over all n frames (an IIR job of one frame takes
    // WEMA_IIR_PPT positions per thread)
    const int ppt = p.use_iir && n == 1 ? WEMA_IIR_PPT : 1;
    const int tiles_t = (int)((nu + 256 * ppt - 1) / (256 * ppt));
    for (int c = 0; c < nch; c++) {
        TemporalJob j{};
        j.n = (long long)nu;
        j.level0_end = (long long)g.plane_off[1];
        j.acc = b.acc + (size_t)c * acc_s;
        j.acc_stride = nch * acc_s;
        j.acc_slots = acc_s;
        j.nframes = n;
        j.n6f = (float)(2 * nu);
        j.thr_out = s->thr_d + c;
        j.thr_override = s->thr_ovr_active ? s->thr_ovr_d + c : nullptr;
        j.amp_index = (int)(jb.host_amp.size() - (size_t)nch * n + (size_t)c * n);
        if (p.use_iir) {
            j.coef = coef_ptr(0, c);
            j.coef_stride = (long long)(nch * nus);
            j.out_stride = (long long)(nch * nus);
            j.amp_out = b.ampc ? b.ampc + (size_t)c * nus : nullptr;
            j.filt_out = s->filt ? s->filt + (size_t)c * nus : nullptr;
            j.state = s->state + (size_t)c * nus;
            memcpy(j.c, s->c10, sizeof(j.c));
            j.preseed = frame0 == 0;
            j.warmup = s->warmup;
            j.ppt = ppt;
        } else {
            j.coef = nullptr;
            j.ring = s->coef + (size_t)c * s->ring_slots * nus;
            j.amp_out = b.ampc + (size_t)c * nus;
            j.coef_stride = (long long)nus;          // ring plane stride
            j.out_stride = (long long)(nch * nus);   // amp_out / filt_out frame stride
            j.filt_out = s->filt ? s->filt + (size_t)c * nus : nullptr;
            j.window = p.window;
            j.ring_slots = s->ring_slots;
            long first = frame0 - p.window + 1;
            first %= s->ring_slots;
            if (first < 0) first += s->ring_slots;
            j.ring_first = (int)first;
            j.bilateral = p.bilateral;
            j.ltw = s->tw_d;
        }
        // IIR: coefficients read + written per frame, state read + written once; window:
        // the ws - 1 + n ring samples the call's windows cover, read once, and one amplified
        // coefficient written per frame
        const double tb = p.use_iir ? (double)nu * (8.0 * n + 32.0)
                                    : (double)nu * (4.0 * (p.window - 1 + n) + 4.0 * n);
        jb.add_temporal(j, tiles_t, tb, !p.use_iir);
    }

    // frames whose output is the input itself (window mode while the window fills)
    auto passthrough = [&](int f) { return !p.use_iir && frame0 + f + 1 < p.window; };
    auto planes_ptr = [&](int f, int c, int l) -> const float * {
        if (!b.ampc) return coef_ptr(f, c) + g.plane_off[l];
        return b.ampc + ((size_t)f * nch + c) * nus + g.plane_off[l];
    };
    const size_t out_b = (out_u8 || out_list ? 3 * P : 0) + (out_f32 ? 12 * P : 0);

    // inverse DWT levels, coarsest first
    for (int l = L - 1; l >= 0; l--) {
        const int w = g.lw[l], h = g.lh[l];
        const int tiles_x = wema_ceil_div(g.lw[l + 1], WEMA_TILE_PAIRS_X);
        const int tiles = 0;   // (strip heights and block ranges are set when the table is packed)
        const double in_b = 16.0 * g.lw[l + 1] * g.lh[l + 1];
        for (int f = 0; f < n; f++) {
            if (l == 0 && p.color && !passthrough(f)) {
                // colour: level 0 of Y, Cb, Cr and the RGB egress in one job
                DwtiJob j{};
                j.ll = ll_ptr(f, 0, 0);
                j.planes = planes_ptr(f, 0, 0);
                j.ll_cstride = (long long)(ll_ptr(f, 1, 0) - ll_ptr(f, 0, 0));
                j.planes_cstride = (long long)(planes_ptr(f, 1, 0) - planes_ptr(f, 0, 0));
                j.nch = nch;
                j.out_u8 = out_u8_of(f);
                j.out_f32 = out_f32 ? out_f32 + (size_t)f * P * 3 : nullptr;
                j.acc_clear = L == 1 ? b.acc + (size_t)f * nch * acc_s : nullptr;
                j.acc_n = nch * acc_s;
                j.w = w;
                j.h = h;
                j.tiles_x = tiles_x;
                j.sp = s->sp_inv[0];
                jb.add_dwti_ycc(j, tiles, in_b * nch + (double)out_b);
                continue;
            }
            for (int c = 0; c < nch; c++) {
                DwtiJob j{};
                j.ll = ll_ptr(f, c, l);
                j.planes = planes_ptr(f, c, l);
                j.out = l == 0 ? rec_ptr(f, c) : ll_ptr(f, c, l - 1);
                j.acc_clear = l == L - 1 ? b.acc + ((size_t)f * nch + c) * acc_s : nullptr;
                j.acc_n = acc_s;
                j.nch = 1;
                j.w = w;
                j.h = h;
                j.tiles_x = tiles_x;
                j.sp = s->sp_inv[l];
                if ((j.sp == WEMA_SP_TILE || j.sp == WEMA_SP_POINT) && (l == 0 || w < 2 || h < 2)) j.sp = 0;
                jb.add_dwti(WEMA_MAX_LEVELS - 1 - l, j, tiles, in_b + 4.0 * w * h);
            }
        }
    }

    // output stage
    const int gf_strips = wema_ceil_div(g.W, WEMA_GF_OUT);            // guided filter
    const int gs_tx = wema_ceil_div(g.W, WEMA_GF_TX);
    const int gs_tiles = gs_tx * wema_ceil_div(g.H, WEMA_GF_TY);      // 5x5 path
    const int byte_tiles = (int)((3 * P + 255) / 256);
    for (int f = 0; f < n; f++) {
        EgressJob j{};
        j.gray = rec_ptr(f, 0);
        j.cb = p.color ? rec_ptr(f, 1) : nullptr;
        j.cr = p.color ? rec_ptr(f, 2) : nullptr;
        if (in_f32) j.src_f32 = (const float *)in_of(f);
        else j.src_u8 = (const uint8_t *)in_of(f);
        j.out_u8 = out_u8_of(f);
        j.out_f32 = out_f32 ? out_f32 + (size_t)f * P * 3 : nullptr;
        j.w = g.W;
        j.h = g.H;
        const double src_b = 3.0 * P * (in_f32 ? 4 : 1);
        if (passthrough(f)) {
            jb.add_egress(EG_PASSTHRU, j, byte_tiles, src_b + out_b);
        } else if (p.color) {
            continue;   // written by the colour output job of the inverse stage
        } else if (p.edge_aware || p.batch_path) {
            j.tiles_x = gf_strips;
            j.seg = s->gf_seg;   // 0: one height for the launch (JobBuilder::assign_segments)
            jb.add_egress(wema_gf_fast(j) ? EG_GF : EG_GF_ANY, j, 0, 4.0 * P + src_b + out_b);
        } else {
            j.tiles_x = gs_tx;
            jb.add_egress(EG_GAUSS, j, gs_tiles, 4.0 * P + src_b + out_b);
        }
    }
}

const char *const wema_timer_names[TK_COUNT] = {
    "k_dwt_fwd0", "k_dwt_fwd", "k_iir", "k_win", "k_dwt_inv",
    "k_dwt_inv0", "k_dwt_inv0_ycc", "k_gf", "k_gauss", "k_passthru"};

static int timer_mark(LaunchTimer *t, int kind, double bytes, hipStream_t st) {
    if (!t || !t->enabled) return 0;
    hipEvent_t ev;
    if (!t->spare.empty()) {
        ev = t->spare.back();
        t->spare.pop_back();
    } else {
        HIP_OK(hipEventCreate(&ev));
    }
    HIP_OK(hipEventRecord(ev, st));
    t->recs.push_back({ev, kind, bytes});
    return 0;
}

// Sums the recorded launches: the time between a launch's start marker and the next marker
// is that launch's duration on the stream.
int wema_timer_collect(LaunchTimer *t) {
    for (size_t i = 0; i < t->recs.size(); i++) {
        if (t->recs[i].kind >= 0 && i > 0) {
            HIP_OK(hipEventSynchronize(t->recs[i].ev));
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, t->recs[i - 1].ev, t->recs[i].ev));
            const int k = t->recs[i].kind;
            t->ms[k] += ms;
            t->bytes[k] += t->recs[i].bytes;
            t->launches[k] += 1;
        }
    }
    for (auto &r : t->recs) t->spare.push_back(r.ev);
    t->recs.clear();
    return 0;
}

// Uploads the tables of a builder and launches all stages on stream st.
int wema_run_jobs(JobBuilder &jb, void *stage_h, void *stage_d, hipEvent_t stage_ev, hipStream_t st,
                  LaunchTimer *timer) {
    size_t bytes = 0;
    if (jb.pack(stage_h, stage_d, WEMA_STAGE_BYTES, &bytes) < 0) {
        wema_set_error("job tables exceed %d bytes", WEMA_STAGE_BYTES);
        return -1;
    }
    if (wema_inject(INJ_TABLE_UPLOAD)) {
        wema_set_error("hipMemcpyAsync of the job tables failed: injected failure");
        return -1;
    }
    HIP_OK(hipMemcpyAsync(stage_d, stage_h, bytes, hipMemcpyHostToDevice, st));
    if (stage_ev) HIP_OK(hipEventRecord(stage_ev, st));
    char *base = (char *)stage_d;
    // from here on part of the call is enqueued: a failure returns -2
#define LAUNCH_OK(call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = wema_inject(INJ_LAUNCH) ? hipErrorLaunchFailure : (call);           \
        if (e_ != hipSuccess) {                                                            \
            wema_set_error("%s failed: %s%s", #call, hipGetErrorString(e_),                \
                           " (the session's state is now unknown)");                     \
            return -2;                                                                     \
        }                                                                                  \
    } while (0)
    if (timer_mark(timer, -1, 0, st) < 0) return -2;
    for (int l = 0; l < WEMA_MAX_LEVELS; l++) {
        const JobList &jl = jb.dwtf[l];
        if (jl.count) {
            TraceRange r_(l == 0 ? "k_dwt_fwd0" : "k_dwt_fwd");
            LAUNCH_OK(wema_launch_dwt_fwd((const DwtfJob *)(base + jl.offset), jl.count, jl.blocks, l == 0, st));
            if (timer_mark(timer, l == 0 ? TK_DWT_FWD0 : TK_DWT_FWD, jl.bytes, st) < 0) return -2;
        }
        const JobList &tl = jb.dwtf_t[l];
        if (tl.count) {
            TraceRange r_("k_dwt_fwd_tile");
            LAUNCH_OK(wema_launch_dwt_fwd_tile((const DwtfJob *)(base + tl.offset), tl.count, tl.blocks, st));
            if (timer_mark(timer, TK_DWT_FWD, tl.bytes, st) < 0) return -2;
        }
        const JobList &pl = jb.dwtf_p[l];
        if (pl.count) {
            TraceRange r_("k_dwt_fwd_point");
            LAUNCH_OK(wema_launch_dwt_fwd_point((const DwtfJob *)(base + pl.offset), pl.count, pl.blocks, st));
            if (timer_mark(timer, TK_DWT_FWD, pl.bytes, st) < 0) return -2;
        }
    }
    const JobList &tl = jb.temporal;
    if (tl.count) {
        TraceRange r_("k_iir");
        LAUNCH_OK(wema_launch_iir((const TemporalJob *)(base + tl.offset), tl.count, tl.blocks, 0, st));
        if (timer_mark(timer, TK_IIR, tl.bytes, st) < 0) return -2;
    }
    const JobList &t1 = jb.temporal1;
    if (t1.count) {
        TraceRange r_("k_iir1");
        LAUNCH_OK(wema_launch_iir((const TemporalJob *)(base + t1.offset), t1.count, t1.blocks, 1, st));
        if (timer_mark(timer, TK_IIR, t1.bytes, st) < 0) return -2;
    }
    for (int v = 0; v < WEMA_WIN_KINDS; v++) {
        const JobList &wl = jb.window[v];
        if (!wl.count) continue;
        TraceRange r_("k_win");
        LAUNCH_OK(wema_launch_win(v, (const TemporalJob *)(base + wl.offset), wl.count, wl.blocks, st));
        if (timer_mark(timer, TK_WIN, wl.bytes, st) < 0) return -2;
    }
    // inverse stages, coarsest first: level l of every stream is stage WEMA_MAX_LEVELS-1-l,
    // so the last stage holds exactly the level-0 (full-resolution) jobs
    const int last_inv = WEMA_MAX_LEVELS - 1;
    for (int l = 0; l < WEMA_MAX_LEVELS; l++) {
        const JobList &jl = jb.dwti[l];
        if (jl.count) {
            TraceRange r_(l == last_inv ? "k_dwt_inv0" : "k_dwt_inv");
            LAUNCH_OK(wema_launch_dwt_inv((const DwtiJob *)(base + jl.offset), jl.count, jl.blocks, l == last_inv, st));
            if (timer_mark(timer, l == last_inv ? TK_DWT_INV0 : TK_DWT_INV, jl.bytes, st) < 0) return -2;
        }
        const JobList &tl = jb.dwti_t[l];
        if (tl.count) {
            TraceRange r_("k_dwt_inv_tile");
            LAUNCH_OK(wema_launch_dwt_inv_tile((const DwtiJob *)(base + tl.offset), tl.count, tl.blocks, st));
            if (timer_mark(timer, TK_DWT_INV, tl.bytes, st) < 0) return -2;
        }
        const JobList &pl = jb.dwti_p[l];
        if (pl.count) {
            TraceRange r_("k_dwt_inv_point");
            LAUNCH_OK(wema_launch_dwt_inv_point((const DwtiJob *)(base + pl.offset), pl.count, pl.blocks, st));
            if (timer_mark(timer, TK_DWT_INV, pl.bytes, st) < 0) return -2;
        }
    }
    if (jb.dwti_ycc.count) {
        const JobList &jl = jb.dwti_ycc;
        TraceRange r_("k_dwt_inv0_ycc");
        LAUNCH_OK(wema_launch_dwt_inv_ycc((const DwtiJob *)(base + jl.offset), jl.count, jl.blocks, st));
        if (timer_mark(timer, TK_DWT_INV0_YCC, jl.bytes, st) < 0) return -2;
    }
    static const int eg_kind[EG_KINDS] = {TK_GF, TK_GF, TK_GAUSS, TK_PASSTHRU};
    for (int k = 0; k < EG_KINDS; k++) {
        const JobList &jl = jb.egress[k];
        if (!jl.count) continue;
        TraceRange r_(wema_timer_names[eg_kind[k]]);
        LAUNCH_OK(wema_launch_egress(k, (const EgressJob *)(base + jl.offset), jl.count, jl.blocks, st));
        if (timer_mark(timer, eg_kind[k], jl.bytes, st) < 0) return -2;
    }
    return 0;
#undef LAUNCH_OK
}

static int process(wema_dev *s, const void *in, int in_f32, uint8_t *out_u8, float *out_f32,
                   int n, void *stream, const CallBufs *bufs = nullptr) {
    if (!s) { wema_set_error("null session"); return -1; }
    if (s->pooled) { wema_set_error("a pooled session is processed by wema_pool_tick"); return -1; }
    if (s->failed) { wema_set_error("the session failed in an earlier call; close it"); return -1; }
    if (n < 1 || n > s->p.max_batch) {
        wema_set_error("frame count %d outside 1..%d", n, s->p.max_batch);
        return -1;
    }
    if (!in || (!out_u8 && !out_f32)) { wema_set_error("null frame buffer"); return -1; }
    HIP_OK(hipSetDevice(s->p.device));
    TraceRange trace("wema_dev_process");
    hipStream_t st = (hipStream_t)stream;    // NULL: the default (null) stream
    s->last_stream = st;
    const int k = s->stage_cur;
    s->stage_cur ^= 1;
    HIP_OK(hipEventSynchronize(s->stage_ev[k]));   // staging buffer k free again
    if (bufs)   // the call's accumulators start at zero (the scratch keeps the last call's contents)
        HIP_OK(wema_launch_zero_u64(bufs->acc, (long long)n * s->g.nch * wema_acc_slots(n), st));
    JobBuilder jb;
    wema_build_jobs(s, jb, in, in_f32, out_u8, out_f32, n, bufs);
    const int rc = wema_run_jobs(jb, s->stage_h[k], s->stage_d[k], s->stage_ev[k], st, &s->timer);
    if (rc < 0) {
        s->failed = rc == -2;
        return -1;
    }
    s->frames += n;
    s->last_n = n;
    s->thr_ovr_active = 0;
    return 0;
}

extern "C" int wema_dev_process_u8(wema_dev_t *s, const uint8_t *d_in, uint8_t *d_out, int n,
                                   void *hip_stream) {
    return process(s, d_in, 0, d_out, nullptr, n, hip_stream);
}

// Bytes of static scratch a call of n frames needs, and the buffers carved from base.
static size_t scratch_plan(const wema_dev *s, int n, char *base, CallBufs *b) {
    const WemaGeom &g = s->g;
    const size_t nf = (size_t)n, nch = (size_t)g.nch;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + off : nullptr;
        off = align_up(off + bytes, 256);
        return p;
    };
    // IIR: the coefficient planes are per call (window mode keeps them in the session's ring);
    // window mode: the amplified planes; IIR with the test capture keeps its arena copies
    b->coef = s->p.use_iir ? (float *)take(sizeof(float) * g.nus * nch * nf) : s->coef;
    b->ampc = s->p.use_iir ? s->ampc : (float *)take(sizeof(float) * g.nus * nch * nf);
    b->ll = (float *)take(sizeof(float) * g.ll_total * nch * nf);
    b->rec = (float *)take(sizeof(float) * g.P * nch * nf);
    b->acc = (unsigned long long *)take(sizeof(unsigned long long) * wema_acc_entries((int)nch, n));
    return off;
}

extern "C" int wema_dev_process_u8_noalloc(wema_dev_t *s, const uint8_t *d_in, uint8_t *d_out, int n,
                                           void *hip_stream) {
    if (!s) { wema_set_error("null session"); return -1; }
    if (n < 1 || n > s->p.max_batch) {
        wema_set_error("frame count %d outside 1..%d", n, s->p.max_batch);
        return -1;
    }
    CallBufs b;
    const size_t need = scratch_plan(s, n, nullptr, &b);
    if (need > (size_t)WEMA_DEV_SCRATCH_BYTES) {
        wema_set_error("%d frame(s) of %dx%d need %zu bytes of scratch; the static scratch holds %zu",
                       n, s->p.width, s->p.height, need, (size_t)WEMA_DEV_SCRATCH_BYTES);
        return -1;
    }
    char *base = (char *)scratch_ptr(s->p.device);
    if (!base) return -1;
    scratch_plan(s, n, base, &b);
    return process(s, d_in, 0, d_out, nullptr, n, hip_stream, &b);
}

extern "C" int wema_dev_process_f32(wema_dev_t *s, const float *d_in, uint8_t *d_out_u8,
                                    float *d_out_f32, int n, void *hip_stream) {
    return process(s, d_in, 1, d_out_u8, d_out_f32, n, hip_stream);
}

extern "C" int wema_dev_sync(wema_dev_t *s) {
    if (!s) { wema_set_error("null session"); return -1; }
    HIP_OK(hipSetDevice(s->p.device));
    HIP_OK(hipStreamSynchronize(s->last_stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return 0;
}

// ---------------------------------------------------------------------------------------
// temporal state export / import (the IIR sections' state, or the window ring's last
// window - 1 frames, with the frame count): a session continues elsewhere, in another
// process or on another GPU, with the outputs it would have given in place

namespace {
struct StateHeader {
    char magic[8];              // "WEMASTT1"
    int32_t width, height, levels, nch, use_iir, window;
    float c10[10];              // IIR sections (must match for the state to mean the same)
    uint64_t n_unique;
    int64_t frames;             // frames processed
    int32_t ring_frames;        // window mode: frames of coefficients that follow
    int32_t pad_;
    uint64_t payload;           // bytes after the header
};
const char kStateMagic[8] = {'W', 'E', 'M', 'A', 'S', 'T', 'T', '1'};

int ring_frames_of(const wema_dev *s) {
    if (s->p.use_iir) return 0;
    const long k = s->p.window - 1;
    return (int)(s->frames < k ? s->frames : k);
}

size_t payload_bytes(const wema_dev *s, int ring_frames) {
    const size_t nu = s->g.nu, nch = s->g.nch;
    return s->p.use_iir ? nch * nu * sizeof(float4) : nch * (size_t)ring_frames * nu * sizeof(float);
}
}  // namespace

extern "C" int wema_dev_state_bytes(const wema_dev_t *s, size_t *bytes) {
    if (!s || !bytes) { wema_set_error("null argument"); return -1; }
    *bytes = sizeof(StateHeader) + payload_bytes(s, ring_frames_of(s));
    return 0;
}

extern "C" int wema_dev_export_state(wema_dev_t *s, void *host, size_t bytes) {
    if (!s || !host) { wema_set_error("null argument"); return -1; }
    const int rf = ring_frames_of(s);
    const size_t need = sizeof(StateHeader) + payload_bytes(s, rf);
    if (bytes < need) { wema_set_error("state needs %zu bytes, buffer has %zu", need, bytes); return -1; }
    HIP_OK(hipSetDevice(s->p.device));
    if (s->last_stream) HIP_OK(hipStreamSynchronize(s->last_stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    StateHeader hd;
    memset(&hd, 0, sizeof(hd));
    memcpy(hd.magic, kStateMagic, 8);
    hd.width = s->p.width;
    hd.height = s->p.height;
    hd.levels = s->g.L;
    hd.nch = s->g.nch;
    hd.use_iir = s->p.use_iir;
    hd.window = s->p.use_iir ? 0 : s->p.window;
    memcpy(hd.c10, s->c10, sizeof(hd.c10));
    hd.n_unique = s->g.nu;
    hd.frames = s->frames;
    hd.ring_frames = rf;
    hd.payload = need - sizeof(StateHeader);
    memcpy(host, &hd, sizeof(hd));
    char *dst = (char *)host + sizeof(hd);
    const size_t nu = s->g.nu, nus = s->g.nus;
    for (int c = 0; c < s->g.nch; c++) {
        if (s->p.use_iir) {
            HIP_OK(hipMemcpy(dst, s->state + (size_t)c * nus, nu * sizeof(float4), hipMemcpyDeviceToHost));
            dst += nu * sizeof(float4);
        } else {
            // oldest first: frames frames - rf .. frames - 1, ring slot g % ring_slots
            for (long g = s->frames - rf; g < s->frames; g++) {
                const float *src = s->coef + ((size_t)c * s->ring_slots + (size_t)(g % s->ring_slots)) * nus;
                HIP_OK(hipMemcpy(dst, src, nu * sizeof(float), hipMemcpyDeviceToHost));
                dst += nu * sizeof(float);
            }
        }
    }
    return 0;
}

extern "C" int wema_dev_import_state(wema_dev_t *s, const void *host, size_t bytes) {
    if (!s || !host) { wema_set_error("null argument"); return -1; }
    StateHeader hd;
    if (bytes < sizeof(hd)) { wema_set_error("state buffer too short"); return -1; }
    memcpy(&hd, host, sizeof(hd));
    if (memcmp(hd.magic, kStateMagic, 8) != 0) { wema_set_error("not a WEMA session state"); return -1; }
    if (hd.width != s->p.width || hd.height != s->p.height || hd.levels != s->g.L || hd.nch != s->g.nch ||
        hd.n_unique != s->g.nu) {
        wema_set_error("state of a %dx%d, %d-level, %d-channel session", hd.width, hd.height, hd.levels, hd.nch);
        return -1;
    }
    if (hd.use_iir != s->p.use_iir || (!hd.use_iir && hd.window != s->p.window)) {
        wema_set_error("state of another temporal filter (%s, window %d)", hd.use_iir ? "IIR" : "window", hd.window);
        return -1;
    }
    if (hd.use_iir && memcmp(hd.c10, s->c10, sizeof(hd.c10)) != 0) {
        wema_set_error("state of another band-pass (frame rate or band edges differ)");
        return -1;
    }
    const long want_rf = hd.use_iir ? 0 : (hd.frames < s->p.window - 1 ? hd.frames : s->p.window - 1);
    if (hd.frames < 0 || hd.ring_frames != want_rf) {
        wema_set_error("inconsistent state header");
        return -1;
    }
    const int rf = hd.ring_frames;
    const size_t payload = payload_bytes(s, rf);
    if (hd.payload != payload || bytes < sizeof(hd) + payload) {
        wema_set_error("state payload of %llu bytes, expected %zu", (unsigned long long)hd.payload, payload);
        return -1;
    }
    HIP_OK(hipSetDevice(s->p.device));
    if (s->last_stream) HIP_OK(hipStreamSynchronize(s->last_stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    const char *src = (const char *)host + sizeof(hd);
    const size_t nu = s->g.nu, nus = s->g.nus;
    for (int c = 0; c < s->g.nch; c++) {
        if (s->p.use_iir) {
            HIP_OK(hipMemcpy(s->state + (size_t)c * nus, src, nu * sizeof(float4), hipMemcpyHostToDevice));
            src += nu * sizeof(float4);
        } else {
            for (long g = hd.frames - rf; g < hd.frames; g++) {
                float *dst = s->coef + ((size_t)c * s->ring_slots + (size_t)(g % s->ring_slots)) * nus;
                HIP_OK(hipMemcpy(dst, src, nu * sizeof(float), hipMemcpyHostToDevice));
                src += nu * sizeof(float);
            }
        }
    }
    s->frames = (long)hd.frames;
    s->last_n = 0;
    return 0;
}

extern "C" int wema_dev_process_host_u8(wema_dev_t *s, const uint8_t *h_in, uint8_t *h_out) {
    if (!s || !h_in || !h_out) { wema_set_error("null argument"); return -1; }
    HIP_OK(hipSetDevice(s->p.device));
    const size_t bytes = s->g.P * 3;
    if (!s->io_in) {
        HIP_OK(hipMalloc(&s->io_in, bytes));
        HIP_OK(hipMalloc(&s->io_out, bytes));
    }
    HIP_OK(hipMemcpyAsync(s->io_in, h_in, bytes, hipMemcpyHostToDevice, s->stream));
    if (process(s, s->io_in, 0, (uint8_t *)s->io_out, nullptr, 1, s->stream) < 0) return -1;
    s->last_stream = s->stream;
    HIP_OK(hipMemcpyAsync(h_out, s->io_out, bytes, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int wema_dev_process_host_f32(wema_dev_t *s, const float *const *h_in, float *const *h_out,
                                         int n) {
    if (!s || !h_in || !h_out || n < 1) { wema_set_error("bad argument"); return -1; }
    if (n > s->p.max_batch) { wema_set_error("%d frames exceed max_batch %d", n, s->p.max_batch); return -1; }
    for (int i = 0; i < n; i++)
        if (!h_in[i] || !h_out[i]) { wema_set_error("null frame %d", i); return -1; }
    HIP_OK(hipSetDevice(s->p.device));
    const size_t frame = s->g.P * 3;
    if (!s->io_in_f32) {
        HIP_OK(hipMalloc(&s->io_in_f32, sizeof(float) * frame * s->p.max_batch));
        HIP_OK(hipMalloc(&s->io_out_f32, sizeof(float) * frame * s->p.max_batch));
    }
    for (int i = 0; i < n; i++)
        HIP_OK(hipMemcpyAsync(s->io_in_f32 + frame * i, h_in[i], sizeof(float) * frame,
                              hipMemcpyHostToDevice, s->stream));
    if (process(s, s->io_in_f32, 1, nullptr, s->io_out_f32, n, s->stream) < 0) return -1;
    s->last_stream = s->stream;
    for (int i = 0; i < n; i++)
        HIP_OK(hipMemcpyAsync(h_out[i], s->io_out_f32 + frame * i, sizeof(float) * frame,
                              hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return 0;
}

// One batch of frames from host memory through a device slot: upload on s_up, process on
// the session's stream, download on s_down. Runs of consecutive host frames go in one copy.
// Frames are RGB24, or float RGB (f32) through float slots allocated on first use.
static int submit_frames(wema_dev *s, const void *const *h_in, void *const *h_out, int n, bool f32,
                         long *ticket) {
    HIP_OK(hipSetDevice(s->p.device));
    const size_t fb = s->g.P * 3 * (f32 ? sizeof(float) : 1);
    if (!s->s_up) {
        const size_t cap = s->g.P * 3 * (size_t)s->p.max_batch;
        HIP_OK(hipStreamCreateWithFlags(&s->s_up, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&s->s_down, hipStreamNonBlocking));
        for (auto &a : s->aslot) {
            HIP_OK(hipMalloc(&a.d_in, cap));
            HIP_OK(hipMalloc(&a.d_out, cap));
            HIP_OK(hipEventCreateWithFlags(&a.ev_in, hipEventDisableTiming));
            HIP_OK(hipEventCreateWithFlags(&a.ev_comp, hipEventDisableTiming));
            HIP_OK(hipEventCreateWithFlags(&a.ev_out, hipEventDisableTiming));
        }
    }
    if (f32 && !s->aslot[0].d_in_f) {
        const size_t cap = fb * (size_t)s->p.max_batch;
        for (auto &a : s->aslot) {
            HIP_OK(hipMalloc(&a.d_in_f, cap));
            HIP_OK(hipMalloc(&a.d_out_f, cap));
        }
    }
    if (s->failed) { wema_set_error("the session failed in an earlier call; close it"); return -1; }
    if (wema_inject(INJ_FRAME_UPLOAD)) {
        wema_set_error("hipMemcpyAsync of the frames failed: injected failure");
        return -1;
    }
    const long t = s->next_ticket++;
    auto &a = s->aslot[t % WEMA_ASYNC_SLOTS];
    char *d_in = f32 ? (char *)a.d_in_f : (char *)a.d_in;
    char *d_out = f32 ? (char *)a.d_out_f : (char *)a.d_out;
    // the slot's previous batch has left the device before its buffers are overwritten
    if (a.ticket >= 0) HIP_OK(hipStreamWaitEvent(s->s_up, a.ev_out, 0));
    for (int i = 0, k; i < n; i = k) {
        for (k = i + 1; k < n && (const char *)h_in[k] == (const char *)h_in[k - 1] + fb; k++) {}
        HIP_OK(hipMemcpyAsync(d_in + fb * i, h_in[i], fb * (k - i), hipMemcpyHostToDevice, s->s_up));
    }
    HIP_OK(hipEventRecord(a.ev_in, s->s_up));
    HIP_OK(hipStreamWaitEvent(s->stream, a.ev_in, 0));
    if (process(s, d_in, f32 ? 1 : 0, f32 ? nullptr : (uint8_t *)d_out, f32 ? (float *)d_out : nullptr, n,
                s->stream) < 0)
        return -1;
    HIP_OK(hipEventRecord(a.ev_comp, s->stream));
    HIP_OK(hipStreamWaitEvent(s->s_down, a.ev_comp, 0));
    for (int i = 0, k; i < n; i = k) {
        for (k = i + 1; k < n && (char *)h_out[k] == (char *)h_out[k - 1] + fb; k++) {}
        HIP_OK(hipMemcpyAsync(h_out[i], d_out + fb * i, fb * (k - i), hipMemcpyDeviceToHost, s->s_down));
    }
    HIP_OK(hipEventRecord(a.ev_out, s->s_down));
    a.ticket = t;
    if (ticket) *ticket = t;
    return 0;
}

extern "C" int wema_dev_submit_host_u8(wema_dev_t *s, const uint8_t *h_in, uint8_t *h_out, int n,
                                       long *ticket) {
    if (!s || !h_in || !h_out || n < 1) { wema_set_error("bad argument"); return -1; }
    if (s->pooled) { wema_set_error("a pooled session is processed by wema_pool_tick"); return -1; }
    if (n > s->p.max_batch) { wema_set_error("%d frames exceed max_batch %d", n, s->p.max_batch); return -1; }
    if (s->fq_pending) { wema_set_error("frames submitted one by one are still pending"); return -1; }
    const size_t fb = s->g.P * 3;
    std::vector<const void *> in((size_t)n);
    std::vector<void *> out((size_t)n);
    for (int i = 0; i < n; i++) {
        in[(size_t)i] = h_in + fb * i;
        out[(size_t)i] = h_out + fb * i;
    }
    return submit_frames(s, in.data(), out.data(), n, false, ticket);
}

// The undispatched tail of the per-frame queue as one batch.
static int dispatch_pending(wema_dev *s) {
    const int n = s->fq_pending;
    if (!n) return 0;
    std::vector<const void *> in((size_t)n);
    std::vector<void *> out((size_t)n);
    const size_t first = s->fq.size() - (size_t)n;
    for (int i = 0; i < n; i++) {
        in[(size_t)i] = s->fq[first + i].in;
        out[(size_t)i] = s->fq[first + i].out;
    }
    long t;
    if (submit_frames(s, in.data(), out.data(), n, s->fq[first].f32, &t) < 0) return -1;
    for (int i = 0; i < n; i++) s->fq[first + i].ticket = t;
    s->fq_pending = 0;
    return 0;
}

static int submit_one(wema_dev *s, const void *in, void *out, int64_t pts, bool f32) {
    if (!s || !in || !out) { wema_set_error("bad argument"); return -1; }
    if (s->pooled) { wema_set_error("a pooled session is processed by wema_pool_tick"); return -1; }
    // a batch holds frames of one kind: a change of kind dispatches the frames before it
    if (s->fq_pending && s->fq.back().f32 != f32 && dispatch_pending(s) < 0) return -1;
    s->fq.push_back({in, out, pts, -1, f32});
    if (++s->fq_pending == s->p.max_batch) return dispatch_pending(s);
    return 0;
}

static int collect_one(wema_dev *s, int64_t *pts, void **out, bool f32) {
    if (!s) { wema_set_error("null session"); return -1; }
    if (s->fq.empty()) { wema_set_error("no frame submitted"); return -1; }
    if (s->fq.front().f32 != f32) {
        wema_set_error("the oldest frame is %s: collect it with wema_dev_collect%s",
                       f32 ? "RGB24" : "float RGB", f32 ? "" : "_f32");
        return -1;
    }
    if (s->fq.front().ticket < 0 && dispatch_pending(s) < 0) return -1;
    const auto f = s->fq.front();
    HIP_OK(hipEventSynchronize(s->aslot[f.ticket % WEMA_ASYNC_SLOTS].ev_out));
    s->fq.pop_front();
    if (pts) *pts = f.pts;
    if (out) *out = f.out;
    return 0;
}

extern "C" int wema_dev_submit(wema_dev_t *s, const uint8_t *rgb24_in, uint8_t *rgb24_out, int64_t pts) {
    return submit_one(s, rgb24_in, rgb24_out, pts, false);
}

extern "C" int wema_dev_collect(wema_dev_t *s, int64_t *pts, uint8_t **rgb24_out) {
    void *o = nullptr;
    const int rc = collect_one(s, pts, &o, false);
    if (rc == 0 && rgb24_out) *rgb24_out = (uint8_t *)o;
    return rc;
}

extern "C" int wema_dev_submit_f32(wema_dev_t *s, const float *rgb_in, float *rgb_out, int64_t pts) {
    return submit_one(s, rgb_in, rgb_out, pts, true);
}

extern "C" int wema_dev_collect_f32(wema_dev_t *s, int64_t *pts, float **rgb_out) {
    void *o = nullptr;
    const int rc = collect_one(s, pts, &o, true);
    if (rc == 0 && rgb_out) *rgb_out = (float *)o;
    return rc;
}

extern "C" int wema_dev_wait_ticket(wema_dev_t *s, long ticket) {
    if (!s || ticket < 0 || ticket >= s->next_ticket) { wema_set_error("bad ticket"); return -1; }
    // the slot's event marks this batch's download or a later one, and downloads run in
    // order on one stream
    HIP_OK(hipEventSynchronize(s->aslot[ticket % WEMA_ASYNC_SLOTS].ev_out));
    return 0;
}

extern "C" void *wema_dev_host_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        wema_set_error("hipHostMalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}

extern "C" void wema_dev_host_free(void *p) {
    if (p) hipHostFree(p);
}

// ---------------------------------------------------------------------------------------
// test hooks

extern "C" int wema_dev_timing(wema_dev_t *s, int enable) {
    if (!s) { wema_set_error("null session"); return -1; }
    if (!enable && s->timer.enabled && wema_timer_collect(&s->timer) < 0) return -1;
    s->timer.enabled = enable != 0;
    if (enable) {
        memset(s->timer.ms, 0, sizeof(s->timer.ms));
        memset(s->timer.bytes, 0, sizeof(s->timer.bytes));
        memset(s->timer.launches, 0, sizeof(s->timer.launches));
    }
    return 0;
}

extern "C" int wema_dev_timing_report(wema_dev_t *s, int kind, char *name, int name_len,
                                      long *launches, double *total_ms, double *bytes) {
    if (!s) { wema_set_error("null session"); return -1; }
    if (kind < 0 || kind >= TK_COUNT) return -1;
    if (wema_timer_collect(&s->timer) < 0) return -1;
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", wema_timer_names[kind]);
    if (launches) *launches = s->timer.launches[kind];
    if (total_ms) *total_ms = s->timer.ms[kind];
    if (bytes) *bytes = s->timer.bytes[kind];
    return 0;
}

extern "C" int wema_dev_debug_capture(wema_dev_t *s, int enable) {
    if (!s) { wema_set_error("null session"); return -1; }
    if ((enable != 0) == (s->with_filt != 0)) return 0;
    if (s->frames > 0) { wema_set_error("capture must be enabled before the first frame"); return -1; }
    return wema_session_alloc(s, enable != 0);
}

extern "C" int wema_dev_debug_thresholds(wema_dev_t *s, float *thr, int n) {
    if (!s || !thr) { wema_set_error("null argument"); return -1; }
    if (n > s->last_n) n = s->last_n;
    HIP_OK(hipStreamSynchronize(s->last_stream));
    std::vector<float> t((size_t)3 * s->p.max_batch);
    HIP_OK(hipMemcpy(t.data(), s->thr_d, t.size() * sizeof(float), hipMemcpyDeviceToHost));
    const long frame0 = s->frames - s->last_n;
    for (int f = 0; f < n; f++)
        for (int c = 0; c < 3; c++) {
            float v = c < s->g.nch ? t[(size_t)f * 3 + c] : NAN;
            if (!s->p.use_iir && frame0 + f + 1 < s->p.window) v = NAN;
            thr[f * 3 + c] = v;
        }
    return 0;
}

extern "C" int wema_dev_debug_set_thresholds(wema_dev_t *s, const float *thr, int n) {
    if (!s) { wema_set_error("null session"); return -1; }
    if (!thr) { s->thr_ovr_active = 0; return 0; }
    if (n > s->p.max_batch) { wema_set_error("too many frames"); return -1; }
    std::vector<float> t((size_t)3 * s->p.max_batch, NAN);
    memcpy(t.data(), thr, sizeof(float) * 3 * n);
    HIP_OK(hipStreamSynchronize(s->last_stream));
    HIP_OK(hipMemcpy(s->thr_ovr_d, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    s->thr_ovr_active = 1;
    return 0;
}

extern "C" int wema_dev_debug_div_mid(const float *d_a, const float *d_b, float *d_q, float *d_rcp, long n,
                                      void *hip_stream) {
    if (!d_a || !d_b || !d_q || !d_rcp || n < 0 || (long long)n > (1ll << 31)) {
        wema_set_error("bad argument");
        return -1;
    }
    HIP_OK(wema_launch_debug_div(d_a, d_b, d_q, d_rcp, (long long)n, (hipStream_t)hip_stream));
    return 0;
}

extern "C" int wema_dev_debug_read(wema_dev_t *s, int which, int frame, int ch, float *host,
                                   size_t count) {
    if (!s || !host) { wema_set_error("null argument"); return -1; }
    if (ch < 0 || ch >= s->g.nch || frame < 0 || frame >= s->last_n) {
        wema_set_error("frame/channel out of range");
        return -1;
    }
    const size_t nu = s->g.nu, nus = s->g.nus, nch = s->g.nch;
    const long gidx = s->frames - s->last_n + frame;
    const float *src = nullptr;
    size_t avail = nu;
    switch (which) {
    case 0:
    case 2:
        if (which == 2 && s->ampc) {
            src = s->ampc + ((size_t)frame * nch + ch) * nus;
        } else if (which == 2 || s->p.use_iir) {
            if (which == 0 && !s->ampc) {
                wema_set_error("original coefficients are amplified in place; enable wema_dev_debug_capture");
                return -1;
            }
            src = s->coef + ((size_t)frame * nch + ch) * nus;
        } else {
            src = s->coef + ((size_t)ch * s->ring_slots + (size_t)(gidx % s->ring_slots)) * nus;
        }
        break;
    case 1:
        if (!s->filt) { wema_set_error("enable wema_dev_debug_capture first"); return -1; }
        src = s->filt + ((size_t)frame * nch + ch) * nus;
        break;
    case 3:
        if (!s->state) { wema_set_error("no IIR state"); return -1; }
        src = (const float *)(s->state + (size_t)ch * nus);
        avail = 4 * nu;
        break;
    default:
        wema_set_error("unknown buffer %d", which);
        return -1;
    }
    if (count > avail) count = avail;
    HIP_OK(hipStreamSynchronize(s->last_stream));
    HIP_OK(hipMemcpy(host, src, count * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}
