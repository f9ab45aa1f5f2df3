// Tile-shape autotuners.
//
// Guided filter: its kernels come in tile heights of 8, 16 and 32 rows (64 columns each);
// which is fastest depends on the frame size and the pass (edge tiles, blocks per XCD, LDS
// per block). The first session of a frame size on a device times the three variants of
// each pass on frames of that size and keeps the fastest of each for the process. Tile
// shapes never change results (the box sums are anchored to the image), so the two passes
// may use different heights.
//
// DWT strips: every forward and inverse level runs as strips 60 pair columns wide and 8, 16
// or 32 pair rows tall; taller strips re-read fewer halo rows, shorter ones give more waves
// in flight, and which wins depends on the level's size, its channels and the frames per
// launch. The first session of a shape (level size, channels, frames per call) on a device
// times the three heights on launches of that shape, using the session's own per-call
// buffers before its first frame, and keeps the fastest. The strip height never changes a
// result: each output is computed from the same operands whatever strip it is in.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/wema_dev.h"
#include "wema_host.h"

extern "C" hipError_t wema_launch_egress(int kind, int rows, const EgressJob *jobs, int njobs, int nblocks,
                                         hipStream_t s);
extern "C" hipError_t wema_launch_dwt_fwd(const DwtfJob *, int, int, int, hipStream_t);
extern "C" hipError_t wema_launch_dwt_inv(const DwtiJob *, int, int, int, hipStream_t);
extern "C" hipError_t wema_launch_dwt_inv_ycc(const DwtiJob *, int, int, hipStream_t);

namespace {

std::mutex g_mu;
std::map<std::tuple<int, int, int>, std::pair<int, int>> g_rows;   // (device, w, h) -> rows of each pass
// (device, kind, w, h, nch, frames per call) -> strip height; kind: 0 forward level 0,
// 1 forward coarser level, 2 inverse coarser level, 3 inverse level 0, 4 colour level-0 inverse
std::map<std::tuple<int, int, int, int, int, int>, int> g_strips;
int g_mask = WEMA_TUNE_GF | WEMA_TUNE_DWT;
const int kRows[3] = {8, 16, 32};

struct Scratch {
    void *p[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Scratch() {
        if (st) hipStreamSynchronize(st);
        for (void *q : p)
            if (q) hipFree(q);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (st) hipStreamDestroy(st);
    }
};

// Mean time of k_gf1 (ms1) and of k_gf2 (ms2) over F frames at each tile height; -1 on
// error.
int time_variants(int w, int h, double *ms1, double *ms2) {
    const size_t P = (size_t)w * h;
    const int F = (int)std::max<size_t>(1, std::min<size_t>(32, ((size_t)4 << 20) / P));
    Scratch sc;
    float *gray = nullptr;
    float2 *ab = nullptr;
    uint8_t *src = nullptr, *out = nullptr;
    EgressJob *jobs_d = nullptr;
    if (hipMalloc((void **)&gray, sizeof(float) * P * F) != hipSuccess) return -1;
    sc.p[0] = gray;
    if (hipMalloc((void **)&ab, sizeof(float2) * P * F) != hipSuccess) return -1;
    sc.p[1] = ab;
    if (hipMalloc((void **)&src, 3 * P * F) != hipSuccess) return -1;
    sc.p[2] = src;
    if (hipMalloc((void **)&out, 3 * P * F) != hipSuccess) return -1;
    sc.p[3] = out;
    if (hipMalloc((void **)&jobs_d, sizeof(EgressJob) * 2 * F) != hipSuccess) return -1;
    sc.p[4] = jobs_d;
    if (hipStreamCreateWithFlags(&sc.st, hipStreamNonBlocking) != hipSuccess) return -1;
    if (hipEventCreate(&sc.e0) != hipSuccess || hipEventCreate(&sc.e1) != hipSuccess) return -1;
    if (hipMemsetAsync(gray, 0, sizeof(float) * P * F, sc.st) != hipSuccess ||
        hipMemsetAsync(src, 0x80, 3 * P * F, sc.st) != hipSuccess)
        return -1;
    for (int v = 0; v < 3; v++) {
        const int rows = kRows[v];
        const int tiles_x = wema_ceil_div(w, WEMA_GF_TX), tiles = tiles_x * wema_ceil_div(h, rows);
        std::vector<EgressJob> jobs((size_t)F);
        for (int f = 0; f < F; f++) {
            EgressJob j{};
            j.gray = gray + (size_t)f * P;
            j.src_u8 = src + (size_t)f * P * 3;
            j.ab = ab + (size_t)f * P;
            j.out_u8 = out + (size_t)f * P * 3;
            j.w = w;
            j.h = h;
            j.tiles_x = tiles_x;
            j.tile_begin = f * tiles;
            jobs[(size_t)f] = j;
        }
        if (hipMemcpyAsync(jobs_d, jobs.data(), sizeof(EgressJob) * F, hipMemcpyHostToDevice, sc.st) != hipSuccess)
            return -1;
        const int nb = F * tiles;
        for (int k = 0; k < 2; k++) {
            const int kind = k == 0 ? EG_GF1 : EG_GF2;
            if (wema_launch_egress(kind, rows, jobs_d, F, nb, sc.st) != hipSuccess) return -1;   // warm-up
            const int reps = 3;
            if (hipEventRecord(sc.e0, sc.st) != hipSuccess) return -1;
            for (int r = 0; r < reps; r++)
                if (wema_launch_egress(kind, rows, jobs_d, F, nb, sc.st) != hipSuccess) return -1;
            if (hipEventRecord(sc.e1, sc.st) != hipSuccess || hipEventSynchronize(sc.e1) != hipSuccess)
                return -1;
            float t = 0.0f;
            if (hipEventElapsedTime(&t, sc.e0, sc.e1) != hipSuccess) return -1;
            (k == 0 ? ms1 : ms2)[v] = t / reps;
        }
    }
    return 0;
}

}  // namespace

// Tile rows of each guided-filter pass at this frame size (timed once per size and device).
void wema_gf_rows_for(int device, int w, int h, int *rows1, int *rows2) {
    *rows1 = *rows2 = 16;
    if (!(g_mask & WEMA_TUNE_GF) || w < 4 || h < 4) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_tuple(device, w, h);
    auto it = g_rows.find(key);
    if (it == g_rows.end()) {
        double ms1[3] = {0, 0, 0}, ms2[3] = {0, 0, 0};
        std::pair<int, int> rows(16, 16);
        if (time_variants(w, h, ms1, ms2) == 0) {
            int b1 = 1, b2 = 1;
            for (int v = 0; v < 3; v++) {
                if (ms1[v] < ms1[b1]) b1 = v;
                if (ms2[v] < ms2[b2]) b2 = v;
            }
            rows = std::make_pair(kRows[b1], kRows[b2]);
        }
        it = g_rows.emplace(key, rows).first;
    }
    *rows1 = it->second.first;
    *rows2 = it->second.second;
}

extern "C" void wema_dev_autotune(int mask) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_mask = mask & (WEMA_TUNE_GF | WEMA_TUNE_DWT);
}

namespace {

const int kStrips[3] = {8, 16, 32};

// Mean launch time of one job list at each strip height (the jobs' buffers are a session's
// own per-call buffers; their contents do not matter to the timing).
template <typename Job, typename Launch>
int time_strips(std::vector<Job> v, void *tab_h, void *tab_d, size_t cap, Scratch &sc, double *ms, Launch launch) {
    for (int c = 0; c < 3; c++) {
        for (auto &j : v) j.sp = kStrips[c];
        JobList jl;
        JobBuilder::strips_of(v, jl, 0);
        size_t off = 0;
        if (!JobBuilder::put((char *)tab_h, cap, off, v, jl)) return -1;
        if (hipMemcpyAsync(tab_d, tab_h, off, hipMemcpyHostToDevice, sc.st) != hipSuccess) return -1;
        const Job *jd = (const Job *)((char *)tab_d + jl.offset);
        if (launch(jd, jl.count, jl.blocks) != hipSuccess) return -1;   // warm-up
        const int reps = 3;
        if (hipEventRecord(sc.e0, sc.st) != hipSuccess) return -1;
        for (int r = 0; r < reps; r++)
            if (launch(jd, jl.count, jl.blocks) != hipSuccess) return -1;
        if (hipEventRecord(sc.e1, sc.st) != hipSuccess || hipEventSynchronize(sc.e1) != hipSuccess) return -1;
        float t = 0.0f;
        if (hipEventElapsedTime(&t, sc.e0, sc.e1) != hipSuccess) return -1;
        ms[c] = t / reps;
    }
    return 0;
}

}  // namespace

// Strip heights of a session's DWT levels, from the shapes tuned so far or by timing the
// new ones on the session's own per-call buffers (before its first frame); the threshold
// accumulators the forward launches add into are cleared afterwards. Sessions keep the rule
// (sp 0) while DWT tuning is off.
int wema_dwt_tune(wema_dev *s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!(g_mask & WEMA_TUNE_DWT)) return 0;
    const int n = s->p.max_batch, dev = s->p.device, L = s->g.L;
    JobBuilder jb;
    // a session's per-call buffers as every job's source and destination: the rec planes
    // hold at least 3 bytes per pixel per frame, enough for RGB24 in and out
    wema_build_jobs(s, jb, s->rec, 0, (uint8_t *)s->rec, nullptr, n);
    Scratch sc;
    void *tab_h = nullptr, *tab_d = nullptr;
    const size_t cap = 1 << 20;
    bool timed = false;
    auto ensure = [&]() {
        if (sc.st) return 0;
        if (hipStreamCreateWithFlags(&sc.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&sc.e0) != hipSuccess || hipEventCreate(&sc.e1) != hipSuccess ||
            hipMalloc(&tab_d, cap) != hipSuccess || hipHostMalloc(&tab_h, cap, hipHostMallocDefault) != hipSuccess) {
            wema_set_error("strip autotuner set-up failed");
            return -1;
        }
        sc.p[0] = tab_d;
        return 0;
    };
    auto pick = [](const double *ms) {
        int b = 0;
        for (int c = 1; c < 3; c++)
            if (ms[c] < ms[b]) b = c;
        return kStrips[b];
    };
    int rc = 0;
    for (int l = 0; l < L && rc == 0; l++) {
        // forward level l
        const auto &fv = jb.dwtf_v[l];
        if (!fv.empty()) {
            const auto key = std::make_tuple(dev, l == 0 ? 0 : 1, fv[0].w, fv[0].h, fv[0].nch, n);
            auto it = g_strips.find(key);
            if (it == g_strips.end()) {
                double ms[3];
                const bool lev0 = l == 0;
                if (ensure() < 0 ||
                    time_strips(fv, tab_h, tab_d, cap, sc, ms, [&](const DwtfJob *j, int cnt, int nb) {
                        return wema_launch_dwt_fwd(j, cnt, nb, lev0, sc.st);
                    }) < 0) { rc = -1; break; }
                timed = true;
                it = g_strips.emplace(key, pick(ms)).first;
            }
            s->sp_fwd[l] = it->second;
        }
        // inverse level l (colour level 0: the Y, Cb, Cr + RGB job)
        const bool ycc = l == 0 && !jb.dwti_ycc_v.empty();
        const auto &iv = ycc ? jb.dwti_ycc_v : jb.dwti_v[WEMA_MAX_LEVELS - 1 - l];
        if (!iv.empty()) {
            const int kind = ycc ? 4 : (l == 0 ? 3 : 2);
            const auto key = std::make_tuple(dev, kind, iv[0].w, iv[0].h, iv[0].nch, n);
            auto it = g_strips.find(key);
            if (it == g_strips.end()) {
                double ms[3];
                const bool lev0 = l == 0;
                if (ensure() < 0) { rc = -1; break; }
                const int r = ycc ? time_strips(iv, tab_h, tab_d, cap, sc, ms, [&](const DwtiJob *j, int cnt, int nb) {
                                        return wema_launch_dwt_inv_ycc(j, cnt, nb, sc.st);
                                    })
                                  : time_strips(iv, tab_h, tab_d, cap, sc, ms, [&](const DwtiJob *j, int cnt, int nb) {
                                        return wema_launch_dwt_inv(j, cnt, nb, lev0, sc.st);
                                    });
                if (r < 0) { rc = -1; break; }
                timed = true;
                it = g_strips.emplace(key, pick(ms)).first;
            }
            s->sp_inv[l] = it->second;
        }
    }
    if (timed && rc == 0) {
        if (hipMemsetAsync(s->acc, 0, sizeof(unsigned long long) * (size_t)s->g.nch * (size_t)n, sc.st) != hipSuccess ||
            hipStreamSynchronize(sc.st) != hipSuccess) {
            wema_set_error("strip autotuner clean-up failed");
            rc = -1;
        }
    }
    if (tab_h) hipHostFree(tab_h);
    if (rc < 0 && !wema_dev_last_error()[0]) wema_set_error("strip autotuner failed");
    return rc;
}

extern "C" int wema_dev_strip_rows(const wema_dev_t *s, int level, int inverse) {
    if (!s || level < 0 || level >= s->g.L) return -1;
    return inverse ? s->sp_inv[level] : s->sp_fwd[level];
}

extern "C" int wema_dev_set_strip_rows(wema_dev_t *s, int rows) {
    if (!s) { wema_set_error("null session"); return -1; }
    if (rows != 0 && rows != 8 && rows != 16 && rows != 32) {
        wema_set_error("strip rows must be 0 (the rule), 8, 16 or 32");
        return -1;
    }
    for (int l = 0; l < WEMA_MAX_LEVELS; l++) s->sp_fwd[l] = s->sp_inv[l] = rows;
    return 0;
}

extern "C" int wema_dev_tile_rows(const wema_dev_t *s) { return s ? s->gf_rows : -1; }

extern "C" int wema_dev_tile_rows_pass(const wema_dev_t *s, int pass) {
    if (!s || (pass != 1 && pass != 2)) return -1;
    return pass == 1 ? s->gf_rows : s->gf_rows2;
}

extern "C" int wema_dev_set_tile_rows_pass(wema_dev_t *s, int pass, int rows) {
    if (!s) { wema_set_error("null session"); return -1; }
    if (pass < 0 || pass > 2) { wema_set_error("pass must be 0 (both), 1 or 2"); return -1; }
    if (rows != 8 && rows != 16 && rows != 32) { wema_set_error("tile rows must be 8, 16 or 32"); return -1; }
    if (pass != 2) s->gf_rows = rows;
    if (pass != 1) s->gf_rows2 = rows;
    return 0;
}

extern "C" int wema_dev_set_tile_rows(wema_dev_t *s, int rows) { return wema_dev_set_tile_rows_pass(s, 0, rows); }
