// Host-side session state and job-table builder shared by wema_dev.hip (one stream) and
// wema_pool.hip (many streams per tick).
#pragma once
#include <rocprofiler-sdk-roctx/roctx.h>
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <deque>
#include <vector>

#include "../../include/wema_dev.h"
#include "wema_jobs.h"

// Batches of the asynchronous host pipeline held on the device: the upload of one overlaps
// the processing and download of the other. (Three or four slots, which let uploads run
// back to back, measured 3-5% slower at 1080p: the link, not the gaps, is the limit.)
#ifndef WEMA_ASYNC_SLOTS
#define WEMA_ASYNC_SLOTS 2
#endif

#define WEMA_MAX_LEVELS 12
// Guided filter: workgroups that run at a time (256 CUs x 4: each takes 4 waves, and the
// kernel's registers allow 4 waves per SIMD), for the launch's segment height.
#ifndef WEMA_GF_CONCURRENT
#define WEMA_GF_CONCURRENT 1024
#endif
#define WEMA_STAGE_BYTES (8 << 20)

struct WemaGeom {
    int W, H, L, nch;
    size_t P;
    int lw[WEMA_MAX_LEVELS + 1], lh[WEMA_MAX_LEVELS + 1];   // input size of each level
    size_t plane_off[WEMA_MAX_LEVELS + 1];  // start of level l in the unique-plane layout
    size_t ll_off[WEMA_MAX_LEVELS + 1];     // start of level l's LL in the LL buffer
    size_t nu;                               // unique positions (3 per subband position)
    size_t nus;                              // their stride in the coefficient buffers (nu
                                             // rounded up to 64: planes start on 256 bytes)
    size_t ll_total;
};



// Device buffers that a call's kernels write and read within the call: by default the
// session's arena; wema_dev_process_u8_noalloc points them into the static scratch.
struct CallBufs {
    float *coef;                // IIR: coefficient planes [B][nch][nus] (window mode: the
                                // session's ring, never replaced)
    float *ampc;                // amplified coefficients when kept apart (or null)
    float *ll;                  // [B][nch][ll_total]
    float *rec;                 // [B][nch][P]
    unsigned long long *acc;    // [B][nch][acc slots], zero at the start of the call
};

struct JobList {
    size_t offset = 0;   // byte offset in the packed table
    int count = 0;
    int blocks = 0;
    double bytes = 0;    // algorithmic HBM bytes of the launch (compulsory traffic)
};

// Per-launch timing with HIP events recorded between the launches of a call on its
// stream; read back (and summed per kernel) by wema_dev_timing_report.
struct LaunchTimer {
    bool enabled = false;
    struct Rec {
        hipEvent_t ev;
        int kind;        // -1: start marker; otherwise TimerKind of the launch that ended here
        double bytes;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> spare;   // events of collected records, reused by the next marks
    double ms[16] = {0};
    double bytes[16] = {0};
    long launches[16] = {0};
    LaunchTimer() = default;
    LaunchTimer(const LaunchTimer &) = delete;
    LaunchTimer &operator=(const LaunchTimer &) = delete;
    ~LaunchTimer() {
        for (auto &r : recs) (void)hipEventDestroy(r.ev);
        for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    }
};
enum TimerKind { TK_DWT_FWD0 = 0, TK_DWT_FWD = 1, TK_IIR = 2, TK_WIN = 3, TK_DWT_INV = 4,
                 TK_DWT_INV0 = 5, TK_DWT_INV0_YCC = 6, TK_GF = 7, TK_GAUSS = 8,
                 TK_PASSTHRU = 9, TK_COUNT = 10 };

struct wema_dev {
    wema_dev_params p;
    WemaGeom g;
    float c10[10];
    int warmup = 0;
    int ring_slots = 0;
    int gf_seg = 0;         // guided-filter segment height (rows per workgroup; autotuned at
                            // open; 0: chosen per launch by JobBuilder::assign_segments)
    int sp_fwd[WEMA_MAX_LEVELS] = {0};   // DWT strip height (pair rows) of each forward level
    int sp_inv[WEMA_MAX_LEVELS] = {0};   // and inverse level (autotuned at open; 0: the rule)
    long frames = 0;        // frames processed since open
    int last_n = 0;         // frames of the last call
    hipStream_t stream = nullptr;       // the session's own stream (host-buffer calls)
    hipStream_t last_stream = nullptr;  // stream of the last process call
    void *arena = nullptr;
    size_t arena_bytes = 0;
    float *coef = nullptr;      // IIR: [B][nch][nus] (amplified in place); window: ring
    float *ampc = nullptr;      // amplified coefficients when kept apart: [B][nch][nus]
    float *filt = nullptr;      // captured temporal output [B][nch][nus] (test hook)
    float4 *state = nullptr;    // IIR state [nch][nus]
    float *ll = nullptr;        // [B][nch][ll_total]
    float *rec = nullptr;       // level-0 reconstructions [B][nch][P]
    unsigned long long *acc = nullptr;   // [B][nch][acc slots] (wema_acc_entries)
    float *thr_d = nullptr;     // [B][3] thresholds used
    float *thr_ovr_d = nullptr; // [B][3] overrides (NaN = none)
    float *tw_d = nullptr;      // log2 of the bilateral time weights [256]
    int thr_ovr_active = 0;
    int with_filt = 0;
    void *stage_h[2] = {nullptr, nullptr};
    void *stage_d[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    int stage_cur = 0;
    LaunchTimer timer;
    bool pooled = false;        // owned by a wema_pool (no staging buffers of its own)
    bool failed = false;        // a call failed after enqueuing part of its stages: the
                                // session's state is unknown and every later call fails
    void *io_in = nullptr, *io_out = nullptr;           // host-buffer staging (RGB24)
    float *io_in_f32 = nullptr, *io_out_f32 = nullptr;  // host-buffer staging (float RGB)
    // asynchronous host pipeline: WEMA_ASYNC_SLOTS device slots of max_batch RGB24 frames in
    // and out (batch t uses slot t % WEMA_ASYNC_SLOTS), uploads and downloads on their own
    // streams
    struct AsyncSlot {
        uint8_t *d_in = nullptr, *d_out = nullptr;
        float *d_in_f = nullptr, *d_out_f = nullptr;   // float RGB frames (allocated on first use)
        hipEvent_t ev_in = nullptr, ev_comp = nullptr, ev_out = nullptr;
        long ticket = -1;
    } aslot[WEMA_ASYNC_SLOTS];
    hipStream_t s_up = nullptr, s_down = nullptr;
    long next_ticket = 0;
    // per-frame submit/collect: frames in submission order; the trailing fq_pending ones
    // wait to be dispatched as one batch (ticket -1)
    struct QueuedFrame {
        const void *in;
        void *out;
        int64_t pts;
        long ticket;
        bool f32;       // float RGB frame (wema_dev_submit_f32)
    };
    std::deque<QueuedFrame> fq;
    int fq_pending = 0;
};

// A named host-side range for tracers (rocprofv3 --marker-trace): a call, a pool tick, and
// each stage launch within them.
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

struct JobBuilder {
    std::vector<DwtfJob> dwtf_v[WEMA_MAX_LEVELS];
    std::vector<TemporalJob> temporal_v;      // IIR jobs of several frames
    std::vector<TemporalJob> temporal1_v;     // IIR jobs of one frame (k_iir1)
    std::vector<TemporalJob> window_v[WEMA_WIN_KINDS];   // sliding-window jobs by kernel (wema_win_kind)
    std::vector<DwtiJob> dwti_v[WEMA_MAX_LEVELS];
    std::vector<DwtiJob> dwti_ycc_v;          // colour output jobs (level 0 + YCbCr -> RGB)
    std::vector<DwtfJob> dwtf_t_v[WEMA_MAX_LEVELS];   // coarser levels as LDS tiles
    std::vector<DwtiJob> dwti_t_v[WEMA_MAX_LEVELS];
    std::vector<DwtfJob> dwtf_p_v[WEMA_MAX_LEVELS];   // coarser levels, one output pair per thread
    std::vector<DwtiJob> dwti_p_v[WEMA_MAX_LEVELS];
    std::vector<EgressJob> egress_v[EG_KINDS];
    JobList dwtf[WEMA_MAX_LEVELS], temporal, temporal1, window[WEMA_WIN_KINDS], dwti[WEMA_MAX_LEVELS], dwti_ycc, egress[EG_KINDS];
    JobList dwtf_t[WEMA_MAX_LEVELS], dwti_t[WEMA_MAX_LEVELS];
    JobList dwtf_p[WEMA_MAX_LEVELS], dwti_p[WEMA_MAX_LEVELS];
    std::vector<float> host_amp;
    size_t amp_offset = 0;

    void add_dwtf(int l, DwtfJob j, int tiles, double bytes) {
        if (j.sp == WEMA_SP_POINT) {   // 256 output pairs per workgroup
            j.tile_begin = dwtf_p[l].blocks;
            dwtf_p[l].blocks += wema_pt_blocks(j.w, j.h);
            dwtf_p[l].bytes += bytes;
            dwtf_p_v[l].push_back(j);
            return;
        }
        if (j.sp == WEMA_SP_TILE) {   // one workgroup per tile
            j.tile_begin = dwtf_t[l].blocks;
            dwtf_t[l].blocks += wema_dt_tiles(j.w, j.h);
            dwtf_t[l].bytes += bytes;
            dwtf_t_v[l].push_back(j);
            return;
        }
        j.tile_begin = dwtf[l].blocks;
        dwtf[l].blocks += tiles;
        dwtf[l].bytes += bytes;
        dwtf_v[l].push_back(j);
    }
    // (IIR jobs of one frame go to a launch of their own: a tick of sessions that are
    // behind mixes them with jobs of two frames)
    void add_temporal(TemporalJob j, int tiles, double bytes, bool window_mode) {
        const int wv = !window_mode ? 0 : wema_win_kind(j.window);
        const bool one = !window_mode && j.nframes == 1 && j.ppt == WEMA_IIR_PPT;
        JobList &jl = window_mode ? window[wv] : (one ? temporal1 : temporal);
        j.tile_begin = jl.blocks;
        j.tiles = tiles;
        jl.blocks += tiles;
        jl.bytes += bytes;
        (window_mode ? window_v[wv] : (one ? temporal1_v : temporal_v)).push_back(j);
    }
    void add_dwti(int stage, DwtiJob j, int tiles, double bytes) {
        if (j.sp == WEMA_SP_POINT) {
            j.tile_begin = dwti_p[stage].blocks;
            dwti_p[stage].blocks += wema_pt_blocks(j.w, j.h);
            dwti_p[stage].bytes += bytes;
            dwti_p_v[stage].push_back(j);
            return;
        }
        if (j.sp == WEMA_SP_TILE) {
            j.tile_begin = dwti_t[stage].blocks;
            dwti_t[stage].blocks += wema_dt_tiles(j.w, j.h);
            dwti_t[stage].bytes += bytes;
            dwti_t_v[stage].push_back(j);
            return;
        }
        j.tile_begin = dwti[stage].blocks;
        dwti[stage].blocks += tiles;
        dwti[stage].bytes += bytes;
        dwti_v[stage].push_back(j);
    }
    void add_dwti_ycc(DwtiJob j, int tiles, double bytes) {
        j.tile_begin = dwti_ycc.blocks;
        dwti_ycc.blocks += tiles;
        dwti_ycc.bytes += bytes;
        dwti_ycc_v.push_back(j);
    }
    void add_egress(int kind, EgressJob j, int tiles, double bytes) {
        j.tile_begin = egress[kind].blocks;
        egress[kind].blocks += tiles;
        egress[kind].bytes += bytes;
        egress_v[kind].push_back(j);
    }
    // The temporal jobs read their frames' amplification from the table's own copy of
    // host_amp (at amp_offset of the device table at dev_base).
    void fix_amp(const void *dev_base) {
        const float *amp_d = (const float *)((const char *)dev_base + amp_offset);
        for (auto &j : temporal_v) j.amp = amp_d + j.amp_index;
        for (auto &j : temporal1_v) j.amp = amp_d + j.amp_index;
        for (auto &v : window_v)
            for (auto &j : v) j.amp = amp_d + j.amp_index;
    }
    template <typename T>
    static bool put(char *dst, size_t cap, size_t &off, const std::vector<T> &v, JobList &jl) {
        off = (off + 63) & ~(size_t)63;
        jl.offset = off;
        jl.count = (int)v.size();
        const size_t bytes = sizeof(T) * v.size();
        if (off + bytes > cap) return false;
        if (bytes) memcpy(dst + off, v.data(), bytes);
        off += bytes;
        return true;
    }
    // Strip height of each DWT job: the height its session's autotuner picked for that level
    // (sp set), else one height for the rest of the launch, the tallest of 32, 16 and 8 pair
    // rows that still gives the target count of channel-strips (shorter strips re-read more
    // halo rows, taller ones leave the chip under-filled); block ranges follow. Results
    // never depend on the strip height.
    template <typename Job>
    static void strips_of(std::vector<Job> &v, JobList &jl, long target) {
        if (v.empty()) return;
        auto strips = [&](const Job &j, int sp) { return (long)j.tiles_x * wema_ceil_div((j.h + 1) >> 1, sp); };
        int sp = WEMA_SP_MAX;
        for (; sp > WEMA_SP_MIN; sp /= 2) {
            long work = 0;   // channel-strips: a wave of a colour job computes 3 channels
            for (const auto &j : v) work += strips(j, j.sp ? j.sp : sp) * (j.nch > 1 ? j.nch : 1);
            if (work >= target) break;
        }
        int blocks = 0;
        for (auto &j : v) {
            if (!j.sp) j.sp = sp;
            j.tile_begin = blocks;
            blocks += wema_ceil_div((int)strips(j, j.sp), WEMA_STRIPS_PER_BLOCK);
        }
        jl.blocks = blocks;
    }
    // Segment height of each guided-filter job: its session's tuned height (seg set), else
    // one height for the rest of the launch, picked by a cost model: a launch with R rows to
    // process over all its workgroups (each segment recomputes the 16 rows above it) and C
    // workgroups running at a time takes about max(R / C, the rows of its longest
    // workgroup). Frames of different sizes get the same segment height, so a small frame is
    // not cut into small segments to make up a large one's count. Results never depend on
    // the split.
    static int gf_seg_rows(int h, int nseg) {
        const int seg = 4 * wema_ceil_div(h, 4 * nseg), whole = 4 * wema_ceil_div(h, 4);
        return std::max(seg, std::min(32, whole));
    }
    static void segments_of(std::vector<EgressJob> &v, JobList &jl, long concurrent) {
        if (v.empty()) return;
        int hmax = 0;
        for (const auto &j : v) hmax = std::max(hmax, j.h);
        const int top = 4 * wema_ceil_div(hmax, 4);
        int best = top;
        double best_t = 0.0;
        for (int seg = 32; ; seg = seg < 64 ? seg + 16 : seg + seg / 2) {
            const int sg = std::min(4 * wema_ceil_div(seg, 4), top);
            double rows = 0.0;
            int longest = 0;
            for (const auto &j : v) {
                const int js = j.seg ? j.seg : sg, ns = wema_ceil_div(j.h, js);
                rows += (double)j.tiles_x * (j.h + 16.0 * ns);
                longest = std::max(longest, std::min(js, j.h) + 16);
            }
            const double t = std::max(rows / (double)concurrent, (double)longest);
            if (best_t == 0.0 || t < best_t) { best_t = t; best = sg; }
            if (sg >= top) break;
        }
        int blocks = 0;
        for (auto &j : v) {
            if (!j.seg) j.seg = best;
            j.tile_begin = blocks;
            blocks += j.tiles_x * wema_ceil_div(j.h, j.seg);
        }
        jl.blocks = blocks;
    }
    void assign_strips() {
        for (int l = 0; l < WEMA_MAX_LEVELS; l++)
            strips_of(dwtf_v[l], dwtf[l], l ? WEMA_STRIP_TARGET_COARSE : WEMA_STRIP_TARGET);
        for (int st = 0; st < WEMA_MAX_LEVELS; st++)   // inverse stage st holds level MAX - 1 - st
            strips_of(dwti_v[st], dwti[st],
                      st < WEMA_MAX_LEVELS - 1 ? WEMA_STRIP_TARGET_INV_COARSE : WEMA_STRIP_TARGET_INV);
        strips_of(dwti_ycc_v, dwti_ycc, WEMA_STRIP_TARGET_INV);
        segments_of(egress_v[EG_GF], egress[EG_GF], WEMA_GF_CONCURRENT);
        segments_of(egress_v[EG_GF_ANY], egress[EG_GF_ANY], WEMA_GF_CONCURRENT);
    }
    // Packs every list into dst (pinned host memory) for one upload to dev_base: the
    // amplification array first, so that its device address is known when the temporal
    // jobs that point into it are written.
    int pack(void *dst_v, const void *dev_base, size_t cap, size_t *used) {
        assign_strips();
        char *dst = (char *)dst_v;
        size_t off = 0;
        bool ok = true;
        JobList amp_list;
        ok = ok && put(dst, cap, off, host_amp, amp_list);
        amp_offset = amp_list.offset;
        fix_amp(dev_base);
        for (int l = 0; l < WEMA_MAX_LEVELS; l++) ok = ok && put(dst, cap, off, dwtf_v[l], dwtf[l]);
        ok = ok && put(dst, cap, off, temporal_v, temporal);
        ok = ok && put(dst, cap, off, temporal1_v, temporal1);
        for (int k = 0; k < WEMA_WIN_KINDS; k++) ok = ok && put(dst, cap, off, window_v[k], window[k]);
        for (int l = 0; l < WEMA_MAX_LEVELS; l++) ok = ok && put(dst, cap, off, dwti_v[l], dwti[l]);
        ok = ok && put(dst, cap, off, dwti_ycc_v, dwti_ycc);
        for (int l = 0; l < WEMA_MAX_LEVELS; l++) ok = ok && put(dst, cap, off, dwtf_t_v[l], dwtf_t[l]);
        for (int l = 0; l < WEMA_MAX_LEVELS; l++) ok = ok && put(dst, cap, off, dwti_t_v[l], dwti_t[l]);
        for (int l = 0; l < WEMA_MAX_LEVELS; l++) ok = ok && put(dst, cap, off, dwtf_p_v[l], dwtf_p[l]);
        for (int l = 0; l < WEMA_MAX_LEVELS; l++) ok = ok && put(dst, cap, off, dwti_p_v[l], dwti_p[l]);
        for (int k = 0; k < EG_KINDS; k++) ok = ok && put(dst, cap, off, egress_v[k], egress[k]);
        *used = off;
        return ok ? 0 : -1;
    }
};

void wema_set_error(const char *fmt, ...);
int wema_design_bandpass(float f_low, float f_high, float fs, float *c10, int *warmup);
float wema_fade_amp(float amp, long global_frame);
int wema_auto_levels(int width, int height);
int wema_gf_tune(wema_dev *s);
int wema_dev_open_ex(wema_dev_t **out, const wema_dev_params *p, bool pooled);
int wema_dwt_tune(wema_dev *s);
void wema_geometry_init(WemaGeom *g, const wema_dev_params *p);
void wema_build_jobs(wema_dev *s, JobBuilder &jb, const void *in, int in_f32, uint8_t *out_u8,
                     float *out_f32, int n, const CallBufs *bufs = nullptr,
                     const void *const *in_list = nullptr, uint8_t *const *out_list = nullptr);
// Injected failures (test hook, wema_dev_debug_inject_error): true when the armed site is
// reached for the armed time.
enum InjectSite { INJ_NONE = 0, INJ_TABLE_UPLOAD = 1, INJ_LAUNCH = 2, INJ_FRAME_UPLOAD = 3 };
bool wema_inject(int site);
// Launches all stages of a builder; returns 0, -1 if it failed before enqueuing anything,
// or -2 if it failed after enqueuing some stages.
int wema_run_jobs(JobBuilder &jb, void *stage_h, void *stage_d, hipEvent_t stage_ev, hipStream_t st,
                  LaunchTimer *timer);
int wema_timer_collect(LaunchTimer *t);
extern const char *const wema_timer_names[TK_COUNT];
