// Tile-shape autotuners.
//
// Guided filter: its kernel streams strips of 240 columns down segments of rows; every
// segment recomputes the 16 rows above it, so tall segments waste the least work and short
// ones give more workgroups. The first plain session of a shape (frame size, frames per
// call) on a device times the strips split into 1, 2, 4, .. 64 equal segments (each at
// least 32 rows) on its own per-call buffers when it is opened, and keeps the fastest.
// Pooled sessions leave the split to each launch (JobBuilder::segments_of), whose size
// depends on the other sessions of the tick. The split never changes a result: the window
// sums are anchored to the image.
//
// DWT strips: every forward and inverse level runs as strips 60 pair columns wide and 4, 8,
// 16 or 32 pair rows tall; taller strips re-read fewer halo rows, shorter ones give more waves
// in flight, and which wins depends on the level's size, its channels and the frames per
// launch. The levels below full resolution can also run as LDS tiles or with one output
// pair per thread. The first session of a shape (level size, channels, frames per call) on
// a device times the candidates on launches of that shape, using the session's own per-call
// buffers before its first frame, and keeps the fastest. The shape never changes a result:
// each output is computed from the same operands whatever strip, tile or thread it is in.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/wema_dev.h"
#include "wema_host.h"

extern "C" hipError_t wema_launch_egress(int kind, const EgressJob *jobs, int njobs, int nblocks, hipStream_t s);
extern "C" hipError_t wema_launch_dwt_fwd(const DwtfJob *, int, int, int, hipStream_t);
extern "C" hipError_t wema_launch_dwt_inv(const DwtiJob *, int, int, int, hipStream_t);
extern "C" hipError_t wema_launch_dwt_inv_ycc(const DwtiJob *, int, int, hipStream_t);
extern "C" hipError_t wema_launch_dwt_fwd_tile(const DwtfJob *, int, int, hipStream_t);
extern "C" hipError_t wema_launch_dwt_inv_tile(const DwtiJob *, int, int, hipStream_t);
extern "C" hipError_t wema_launch_dwt_fwd_point(const DwtfJob *, int, int, hipStream_t);
extern "C" hipError_t wema_launch_dwt_inv_point(const DwtiJob *, int, int, hipStream_t);

namespace {

std::mutex g_mu;
std::map<std::tuple<int, int, int, int>, int> g_segs;   // (device, w, h, frames per call) -> segment rows
// (device, kind, w, h, nch, frames per call) -> strip height; kind: 0 forward level 0,
// 1 forward coarser level, 2 inverse coarser level, 3 inverse level 0, 4 colour level-0 inverse
std::map<std::tuple<int, int, int, int, int, int>, int> g_strips;
int g_mask = WEMA_TUNE_GF | WEMA_TUNE_DWT;

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

}  // namespace

extern "C" void wema_dev_autotune(int mask) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_mask = mask & (WEMA_TUNE_GF | WEMA_TUNE_DWT);
}

namespace {

// Candidate shapes: strips 8, 16, 32 and 4 pair rows tall, and (levels below full
// resolution) LDS tiles and one output pair per thread.
const int kStrips[6] = {8, 16, 32, 4, WEMA_SP_TILE, WEMA_SP_POINT};
// Launch time of one job list at each candidate shape (the jobs' buffers are a session's own
// per-call buffers; their contents do not matter to the timing). The candidates' tables are
// uploaded together and timed in alternation over several rounds, keeping each one's best
// round: launches of a few microseconds differ by less than one round's noise.
template <typename Job, typename Launch, typename LaunchTile, typename LaunchPoint>
int time_strips(std::vector<Job> v, void *tab_h, void *tab_d, size_t cap, Scratch &sc, int ncand, double *ms,
                Launch launch, LaunchTile launch_tile, LaunchPoint launch_point) {
    JobList jl[6];
    size_t off = 0;
    for (int c = 0; c < ncand; c++) {
        const bool tile = kStrips[c] == WEMA_SP_TILE, point = kStrips[c] == WEMA_SP_POINT;
        for (auto &j : v) j.sp = kStrips[c];
        if (tile || point) {
            for (auto &j : v) {
                j.tile_begin = jl[c].blocks;
                jl[c].blocks += tile ? wema_dt_tiles(j.w, j.h) : wema_pt_blocks(j.w, j.h);
            }
        } else {
            JobBuilder::strips_of(v, jl[c], 0);
        }
        if (!JobBuilder::put((char *)tab_h, cap, off, v, jl[c])) return -1;
    }
    if (hipMemcpyAsync(tab_d, tab_h, off, hipMemcpyHostToDevice, sc.st) != hipSuccess) return -1;
    auto run = [&](int c) {
        const Job *jd = (const Job *)((char *)tab_d + jl[c].offset);
        if (kStrips[c] == WEMA_SP_TILE) return launch_tile(jd, jl[c].count, jl[c].blocks);
        if (kStrips[c] == WEMA_SP_POINT) return launch_point(jd, jl[c].count, jl[c].blocks);
        return launch(jd, jl[c].count, jl[c].blocks);
    };
    auto timed = [&](int c, int reps, double &t_ms) {
        if (hipEventRecord(sc.e0, sc.st) != hipSuccess) return -1;
        for (int r = 0; r < reps; r++)
            if (run(c) != hipSuccess) return -1;
        if (hipEventRecord(sc.e1, sc.st) != hipSuccess || hipEventSynchronize(sc.e1) != hipSuccess) return -1;
        float t = 0.0f;
        if (hipEventElapsedTime(&t, sc.e0, sc.e1) != hipSuccess) return -1;
        t_ms = t / reps;
        return 0;
    };
    // warm-up round (one launch each), which also sizes the rounds: launches under 0.2 ms
    // get 4 rounds of 8, longer ones 2 rounds of 2
    double longest = 0.0;
    for (int c = 0; c < ncand; c++) {
        if (timed(c, 1, ms[c]) < 0) return -1;
        longest = std::max(longest, ms[c]);
    }
    const bool small = longest < 0.2;
    const int rounds = small ? 4 : 2, reps = small ? 8 : 2;
    for (int c = 0; c < ncand; c++) ms[c] = 1e30;
    for (int r = 0; r < rounds; r++)
        for (int c = 0; c < ncand; c++) {
            double t = 0.0;
            if (timed(c, reps, t) < 0) return -1;
            ms[c] = std::min(ms[c], t);
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
    auto pick = [](const double *ms, int ncand) {
        int b = 0;
        for (int c = 1; c < ncand; c++)
            if (ms[c] < ms[b]) b = c;
        return kStrips[b];
    };
    // LDS tiles and pairs are candidates for the levels below full resolution of 2 x 2
    // samples or more
    auto cands = [](int l, int w, int h) { return l > 0 && w >= 2 && h >= 2 ? 6 : 4; };
    int rc = 0;
    for (int l = 0; l < L && rc == 0; l++) {
        // forward level l
        const auto &fv = jb.dwtf_v[l];
        if (!fv.empty()) {
            const auto key = std::make_tuple(dev, l == 0 ? 0 : 1, fv[0].w, fv[0].h, fv[0].nch, n);
            auto it = g_strips.find(key);
            if (it == g_strips.end()) {
                double ms[6];
                const bool lev0 = l == 0;
                const int nc = cands(l, fv[0].w, fv[0].h);
                if (ensure() < 0 ||
                    time_strips(fv, tab_h, tab_d, cap, sc, nc, ms,
                                [&](const DwtfJob *j, int cnt, int nb) { return wema_launch_dwt_fwd(j, cnt, nb, lev0, sc.st); },
                                [&](const DwtfJob *j, int cnt, int nb) { return wema_launch_dwt_fwd_tile(j, cnt, nb, sc.st); },
                                [&](const DwtfJob *j, int cnt, int nb) { return wema_launch_dwt_fwd_point(j, cnt, nb, sc.st); }) < 0) {
                    rc = -1;
                    break;
                }
                timed = true;
                it = g_strips.emplace(key, pick(ms, nc)).first;
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
                double ms[6];
                const bool lev0 = l == 0;
                const int nc = ycc ? 4 : cands(l, iv[0].w, iv[0].h);
                if (ensure() < 0) { rc = -1; break; }
                auto tile = [&](const DwtiJob *j, int cnt, int nb) { return wema_launch_dwt_inv_tile(j, cnt, nb, sc.st); };
                auto point = [&](const DwtiJob *j, int cnt, int nb) { return wema_launch_dwt_inv_point(j, cnt, nb, sc.st); };
                const int r = ycc ? time_strips(iv, tab_h, tab_d, cap, sc, nc, ms,
                                                [&](const DwtiJob *j, int cnt, int nb) { return wema_launch_dwt_inv_ycc(j, cnt, nb, sc.st); },
                                                tile, point)
                                  : time_strips(iv, tab_h, tab_d, cap, sc, nc, ms,
                                                [&](const DwtiJob *j, int cnt, int nb) { return wema_launch_dwt_inv(j, cnt, nb, lev0, sc.st); },
                                                tile, point);
                if (r < 0) { rc = -1; break; }
                timed = true;
                it = g_strips.emplace(key, pick(ms, nc)).first;
            }
            s->sp_inv[l] = it->second;
        }
    }
    if (timed && rc == 0) {
        if (hipMemsetAsync(s->acc, 0, sizeof(unsigned long long) * (size_t)wema_acc_entries(s->g.nch, n), sc.st) != hipSuccess ||
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
    if (rows != 0 && rows != 4 && rows != 8 && rows != 16 && rows != 32 && rows != WEMA_SP_TILE && rows != WEMA_SP_POINT) {
        wema_set_error("strip rows must be 0 (the rule), 4, 8, 16, 32, 1 (LDS tiles below full resolution) "
                       "or 2 (one output pair per thread below full resolution)");
        return -1;
    }
    for (int l = 0; l < WEMA_MAX_LEVELS; l++)
        s->sp_fwd[l] = s->sp_inv[l] = ((rows == WEMA_SP_TILE || rows == WEMA_SP_POINT) && l == 0) ? 0 : rows;
    return 0;
}

// Guided filter segments of a session: times the candidate splits on the session's own
// per-call buffers (the reconstruction planes as the guide input, the coefficient planes as
// source and output frames; their contents do not matter to the timing).
int wema_gf_tune(wema_dev *s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!(g_mask & WEMA_TUNE_GF) || s->pooled || s->p.color || !(s->p.edge_aware || s->p.batch_path)) return 0;
    const int n = s->p.max_batch, dev = s->p.device, w = s->g.W, h = s->g.H;
    const auto key = std::make_tuple(dev, w, h, n);
    auto it = g_segs.find(key);
    if (it == g_segs.end()) {
        const size_t P = s->g.P;
        std::vector<int> cand;
        for (int nseg = 1; nseg <= 64; nseg *= 2) {
            const int seg = JobBuilder::gf_seg_rows(h, nseg);
            if (cand.empty() || seg < cand.back()) cand.push_back(seg);
        }
        Scratch sc;
        EgressJob *jobs_d = nullptr;
        if (hipStreamCreateWithFlags(&sc.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&sc.e0) != hipSuccess || hipEventCreate(&sc.e1) != hipSuccess ||
            hipMalloc((void **)&jobs_d, sizeof(EgressJob) * (size_t)n) != hipSuccess) {
            wema_set_error("guided-filter autotuner set-up failed");
            return -1;
        }
        sc.p[0] = jobs_d;
        std::vector<EgressJob> jobs((size_t)n);
        double best = 0.0;
        int best_seg = cand[0];
        for (int seg : cand) {
            const int strips = wema_ceil_div(w, WEMA_GF_OUT), tiles = strips * wema_ceil_div(h, seg);
            for (int f = 0; f < n; f++) {
                EgressJob j{};
                j.gray = s->rec + (size_t)f * P;
                j.src_u8 = (const uint8_t *)s->coef + (size_t)f * P * 3;
                j.out_u8 = (uint8_t *)s->coef + (size_t)f * P * 3;
                j.w = w;
                j.h = h;
                j.tiles_x = strips;
                j.tile_begin = f * tiles;
                j.seg = seg;
                jobs[(size_t)f] = j;
            }
            const int kind = wema_gf_fast(jobs[0]) ? EG_GF : EG_GF_ANY;
            if (hipMemcpyAsync(jobs_d, jobs.data(), sizeof(EgressJob) * (size_t)n, hipMemcpyHostToDevice, sc.st) != hipSuccess ||
                wema_launch_egress(kind, jobs_d, n, n * tiles, sc.st) != hipSuccess ||   // warm-up
                hipEventRecord(sc.e0, sc.st) != hipSuccess) {
                wema_set_error("guided-filter autotuner launch failed");
                return -1;
            }
            const int reps = 3;
            for (int r = 0; r < reps; r++)
                if (wema_launch_egress(kind, jobs_d, n, n * tiles, sc.st) != hipSuccess) {
                    wema_set_error("guided-filter autotuner launch failed");
                    return -1;
                }
            float t = 0.0f;
            if (hipEventRecord(sc.e1, sc.st) != hipSuccess || hipEventSynchronize(sc.e1) != hipSuccess ||
                hipEventElapsedTime(&t, sc.e0, sc.e1) != hipSuccess) {
                wema_set_error("guided-filter autotuner timing failed");
                return -1;
            }
            if (seg == cand[0] || t < best) {
                best = t;
                best_seg = seg;
            }
        }
        it = g_segs.emplace(key, best_seg).first;
    }
    s->gf_seg = it->second;
    return 0;
}

extern "C" int wema_dev_tile_rows(const wema_dev_t *s) { return s ? s->gf_seg : -1; }

extern "C" int wema_dev_set_tile_rows(wema_dev_t *s, int rows) {
    if (!s) { wema_set_error("null session"); return -1; }
    if (rows < 0 || rows % 4 != 0) { wema_set_error("segment rows must be 0 (the launch's split) or a positive multiple of 4"); return -1; }
    s->gf_seg = rows;
    return 0;
}
