// Serving loop over a wema_pool: stream sessions arrive as a Poisson process, are admitted
// as HBM (budget) and slots allow, deliver frames at their own frame rate, and leave after
// their lifetime. Every iteration gathers the frames that are due and runs them as one
// pool tick (continuous batching: sessions that arrive while others are on the GPU join
// the next tick); the plain mode processes the same frames one session at a time.
// Frames stay in HBM, or (host_io) travel from pinned host memory and back on copy streams
// that overlap the ticks. Reports per-frame latency (due time -> output complete) and
// aggregate Mpixel/s, with a per-tick breakdown (host time, graph replays, GPU busy time).
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include "../../include/wema_dev.h"
#include "wema_host.h"

// Frames of one session a batched tick takes at most: a session that has fallen behind its
// frame rate catches up that many frames at a time, as one call of that many frames (its
// IIR state is then read and written once for them).
#ifndef WEMA_SERVE_RUN
#define WEMA_SERVE_RUN 2
#endif

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

struct ResClass {
    int w, h;
    float fps;
    int weight;
};

// The multi-stream mix: 8 x 640x480@30, 8 x 1280x720@30, 6 x 1920x1080@60, 2 x 3840x2160@30.
const ResClass kMix[4] = {{640, 480, 30.f, 8}, {1280, 720, 30.f, 8}, {1920, 1080, 60.f, 6}, {3840, 2160, 30.f, 2}};
constexpr int kFrames = WEMA_SERVE_CLASS_FRAMES;

struct Planned {
    double arrive, life, phase;
    wema_dev_params p;
    int cls;
};

// The sessions of a run, in arrival order (deterministic for a seed).
std::vector<Planned> make_plan(const wema_serve_config *cfg) {
    Rng rng{cfg->seed ? cfg->seed : 0x57454D41ull};
    std::vector<Planned> plan;
    int wsum = 0;
    for (const ResClass &c : kMix) wsum += c.weight;
    const int fixed = cfg->saturate ? cfg->sessions : 0;
    const bool warm = !fixed && cfg->warm_start;
    double t = warm ? -cfg->life_max_s : 0.0;
    while (true) {
        if (fixed) {
            if ((int)plan.size() >= fixed) break;
        } else {
            t += -log(1.0 - rng.uniform()) / cfg->arrival_rate;
            if (t >= cfg->duration_s) break;
        }
        Planned pl;
        int cls = 0;
        if (fixed) {
            // a fixed set of sessions takes the mix in its exact proportions
            int r = (int)(plan.size() % (size_t)wsum);
            while (r >= kMix[cls].weight) r -= kMix[cls++].weight;
        } else {
            int r = (int)(rng.uniform() * wsum);
            while (r >= kMix[cls].weight) r -= kMix[cls++].weight;
        }
        pl.cls = cls;
        const float fps = cfg->fps_from_mix ? kMix[cls].fps : (rng.uniform() < 0.5 ? 30.f : 60.f);
        wema_dev_params_default(&pl.p, kMix[cls].w, kMix[cls].h, fps);
        double a = rng.uniform(0.3, 5.0), b = rng.uniform(0.3, 5.0);
        if (a > b) std::swap(a, b);
        if (b - a < 0.2) b = a + 0.2;
        pl.p.f_low = (float)a;
        pl.p.f_high = (float)b;
        pl.p.amp = (float)rng.uniform(20.0, 150.0);
        const int wsz[3] = {16, 32, 64};
        pl.p.window = wsz[rng.next() % 3];
        pl.p.device = cfg->device;
        pl.p.max_batch = cfg->batched ? WEMA_SERVE_RUN : 1;
        pl.arrive = fixed ? 0.0 : t;
        pl.life = fixed ? 1e30 : rng.uniform(cfg->life_min_s, cfg->life_max_s);
        pl.phase = rng.uniform();
        if (pl.arrive < 0.0) {
            // before the start: alive at t = 0 only if it has not left yet
            if (pl.arrive + pl.life <= 0.0) continue;
            pl.life += pl.arrive;
            pl.arrive = 0.0;
        }
        plan.push_back(pl);
    }
    return plan;
}

constexpr int kRun = WEMA_SERVE_RUN;
// Per-session frame buffers: 2 kRun device inputs (host_io), device outputs and pinned host
// outputs (host_io), frame f in slot f mod 2 kRun, so that the frames of the next tick
// (launched before this one is waited for) never share a slot with this one's. Buffers of
// departed sessions are kept per resolution class and handed to the next session of that
// class.
constexpr int kSlots = 2 * kRun;
struct Slots {
    uint8_t *din[kSlots] = {};
    uint8_t *dout[kSlots] = {};
    uint8_t *hout[kSlots] = {};
};

struct SlotPool {
    bool host_io = false;
    std::vector<Slots> free[4];
    std::vector<Slots> all;
    int take(int c, Slots &out) {
        if (!free[c].empty()) {
            out = free[c].back();
            free[c].pop_back();
            return 0;
        }
        const size_t bytes = (size_t)kMix[c].w * kMix[c].h * 3;
        Slots s;
        for (int k = 0; k < kSlots; k++) {
            if (hipMalloc(&s.dout[k], bytes) != hipSuccess ||
                (host_io && (hipMalloc(&s.din[k], bytes) != hipSuccess ||
                             hipHostMalloc(&s.hout[k], bytes, hipHostMallocDefault) != hipSuccess))) {
                wema_set_error("frame slot allocation failed");
                all.push_back(s);
                return -1;
            }
        }
        all.push_back(s);
        out = s;
        return 0;
    }
    void give(int c, const Slots &s) { free[c].push_back(s); }
    void release() {
        for (Slots &s : all)
            for (int k = 0; k < kSlots; k++) {
                if (s.din[k]) hipFree(s.din[k]);
                if (s.dout[k]) hipFree(s.dout[k]);
                if (s.hout[k]) hipHostFree(s.hout[k]);
            }
        all.clear();
        for (auto &f : free) f.clear();
    }
};

struct Live {
    int id;
    int cls;
    int plan_idx;
    double start, end;      // seconds since run start
    double fps;
    long next;              // index of the next frame to deliver
    Slots sl;
};

struct Done {                // a frame whose output the callback sees
    int plan_idx, cls;
    long frame;
    const uint8_t *out;
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" int wema_serve_plan(const wema_serve_config *cfg, int max, wema_dev_params *params,
                               double *arrive_s, int *cls) {
    if (!cfg) { wema_set_error("null argument"); return -1; }
    if (!cfg->saturate && !(cfg->arrival_rate > 0)) { wema_set_error("arrival rate must be positive"); return -1; }
    const std::vector<Planned> plan = make_plan(cfg);
    for (int i = 0; i < max && i < (int)plan.size(); i++) {
        if (params) params[i] = plan[(size_t)i].p;
        if (arrive_s) arrive_s[i] = plan[(size_t)i].arrive;
        if (cls) cls[i] = plan[(size_t)i].cls;
    }
    return (int)plan.size();
}

extern "C" int wema_serve_run(const wema_serve_config *cfg, wema_serve_stats *out) {
    if (!cfg || !out) { wema_set_error("null argument"); return -1; }
    memset(out, 0, sizeof(*out));
    if (!cfg->saturate && !(cfg->arrival_rate > 0)) { wema_set_error("arrival rate must be positive"); return -1; }
    if (hipSetDevice(cfg->device) != hipSuccess) { wema_set_error("hipSetDevice failed"); return -1; }
    const std::vector<Planned> plan = make_plan(cfg);
    out->sessions_arrived = (int)plan.size();
    const bool fixed = cfg->saturate != 0;
    const bool warm = !fixed && cfg->warm_start;
    const bool host_io = cfg->host_io != 0;

    // ---- input frames: a few per resolution class, cycled; in HBM, or pinned host memory
    std::vector<uint8_t *> in_frames[4];
    for (int c = 0; c < 4; c++) {
        wema_synth_params sp;
        wema_synth_preset(&sp, 4, kMix[c].w, kMix[c].h, kMix[c].fps, (uint32_t)c);
        const size_t P = (size_t)kMix[c].w * kMix[c].h;
        std::vector<uint16_t> tex(P);
        wema_synth_texture_host(&sp, tex.data());
        uint8_t *buf = nullptr;
        if (host_io) {
            if (hipHostMalloc(&buf, P * 3 * kFrames, hipHostMallocDefault) != hipSuccess) {
                wema_set_error("pinned frame allocation failed");
                return -1;
            }
            for (int k = 0; k < kFrames; k++) wema_synth_frame_host(&sp, tex.data(), k, buf + P * 3 * k);
        } else {
            uint16_t *dtex = nullptr;
            if (hipMalloc(&dtex, P * 2) != hipSuccess || hipMalloc(&buf, P * 3 * kFrames) != hipSuccess) {
                wema_set_error("frame allocation failed");
                return -1;
            }
            hipMemcpy(dtex, tex.data(), P * 2, hipMemcpyHostToDevice);
            if (wema_synth_frames_device(&sp, dtex, 0, kFrames, buf, nullptr) < 0) return -1;
            hipDeviceSynchronize();
            hipFree(dtex);
        }
        for (int k = 0; k < kFrames; k++) in_frames[c].push_back(buf + P * 3 * k);
    }
    auto free_frames = [&]() {
        for (int c = 0; c < 4; c++)
            if (!in_frames[c].empty()) {
                if (host_io) hipHostFree(in_frames[c][0]);
                else hipFree(in_frames[c][0]);
                in_frames[c].clear();
            }
    };

    // DWT strip heights of the mix's sizes, tuned before the clock starts (the first session
    // of a size would otherwise time the candidate strips while ticks run): one session of
    // each size is opened and closed
    for (const ResClass &c : kMix) {
        wema_dev_params tp;
        wema_dev_params_default(&tp, c.w, c.h, 30.0f);
        tp.device = cfg->device;
        tp.max_batch = cfg->batched ? WEMA_SERVE_RUN : 1;
        wema_dev_t *ts = nullptr;
        if (wema_dev_open_ex(&ts, &tp, true) < 0) { free_frames(); return -1; }
        wema_dev_close(ts);
    }

    wema_pool_t *pool = nullptr;
    if (wema_pool_create(&pool, cfg->device, cfg->hbm_budget) < 0) { free_frames(); return -1; }
    wema_pool_graphs(pool, cfg->use_graphs);
    hipStream_t st = (hipStream_t)wema_pool_stream(pool);
    hipStream_t s_up = nullptr, s_down = nullptr;
    if (host_io && (hipStreamCreateWithFlags(&s_up, hipStreamNonBlocking) != hipSuccess ||
                    hipStreamCreateWithFlags(&s_down, hipStreamNonBlocking) != hipSuccess)) {
        wema_set_error("copy stream creation failed");
        wema_pool_destroy(pool);
        free_frames();
        return -1;
    }
    SlotPool slots_pool;
    slots_pool.host_io = host_io;

    std::vector<double> lat;
    lat.reserve(1 << 20);
    std::deque<int> waiting;      // planned sessions not yet admitted
    std::vector<Live> live;
    size_t next_plan = 0;
    long frames = 0, ticks = 0, tick_frames = 0, rounds = 0, graph_ticks = 0;
    double mpix = 0.0, host_ms = 0.0, gpu_ms = 0.0, link_bytes = 0.0;
    int rc = 0;
    auto t0 = std::chrono::steady_clock::now();
    double t_meas = 0.0;          // start of the measured part (after the warm-up rounds)
    double t_last = 0.0;          // end of the last round
    const double horizon = fixed ? cfg->duration_s : cfg->duration_s + cfg->drain_s;
    const long quota = cfg->frames_per_session > 0 ? cfg->frames_per_session + cfg->warmup_rounds : 0;
    std::vector<int> ids;
    std::vector<const uint8_t *> ins;
    std::vector<uint8_t *> outs;
    std::vector<double> due;
    std::vector<size_t> who;
    std::vector<long> fidx;       // frame index of each entry

    // Ticks are pipelined: a tick is launched before the previous one is waited for, so the
    // host gathers the next tick (and, with host_io, uploads its frames) while the GPU runs
    // this one. ev_begin / ev_end bracket the tick's kernels on the pool stream (GPU busy
    // time); ev_done marks its outputs complete (downloaded, with host_io).
    struct Pending {
        hipEvent_t ev_up = nullptr, ev_begin = nullptr, ev_end = nullptr, ev_done = nullptr;
        std::vector<double> due;
        std::vector<Done> done;
        bool measured = false;
    };
    Pending slots[2];
    for (Pending &q : slots)
        if (hipEventCreateWithFlags(&q.ev_up, hipEventDisableTiming) != hipSuccess ||
            hipEventCreate(&q.ev_begin) != hipSuccess || hipEventCreate(&q.ev_end) != hipSuccess ||
            hipEventCreateWithFlags(&q.ev_done, hipEventDisableTiming) != hipSuccess) {
            wema_set_error("event creation failed");
            rc = -1;
        }
    std::deque<Pending *> inflight;
    int slot = 0;
    auto complete = [&](Pending *q) {
        if (hipEventSynchronize(q->ev_done) != hipSuccess) { wema_set_error("tick failed"); return -1; }
        const double done = now_s(t0);
        if (q->measured) {
            for (double d : q->due) lat.push_back(done - d);
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, q->ev_begin, q->ev_end) == hipSuccess) gpu_ms += ms;
        }
        if (cfg->on_frame)
            for (const Done &d : q->done)
                cfg->on_frame(cfg->cb_user, d.plan_idx, d.frame, d.out, kMix[d.cls].w, kMix[d.cls].h);
        t_last = done;
        return 0;
    };
    auto drain_one = [&]() {
        Pending *q = inflight.front();
        inflight.pop_front();
        return complete(q);
    };
    auto drain_all = [&]() {
        while (!inflight.empty())
            if (drain_one() < 0) return -1;
        return 0;
    };

    // admission (first come, first served; the head waits until it fits)
    auto admit = [&](double now) {
        while (!waiting.empty()) {
            const int pi = waiting.front();
            const Planned &pl = plan[(size_t)pi];
            if (cfg->max_sessions > 0 && (int)live.size() >= cfg->max_sessions) break;
            int id = -1;
            if (wema_pool_add(pool, &pl.p, &id) < 0) {
                out->admission_waits++;
                break;
            }
            Live L;
            L.id = id;
            L.cls = pl.cls;
            L.plan_idx = pi;
            // streams that join in steady state run at a random phase of their frame period
            L.start = now + (warm && pl.arrive == 0.0 ? pl.phase / pl.p.fps : 0.0);
            L.end = now + pl.life;
            L.fps = pl.p.fps;
            L.next = 0;
            if (slots_pool.take(pl.cls, L.sl) < 0) {
                wema_pool_remove(pool, id);
                return -1;
            }
            live.push_back(L);
            out->sessions_admitted++;
            waiting.pop_front();
        }
        return 0;
    };

    if (warm && rc == 0) {
        // the steady-state population joins before the clock starts
        while (next_plan < plan.size() && plan[next_plan].arrive <= 0.0) waiting.push_back((int)next_plan++);
        if (admit(0.0) < 0) rc = -1;
        t0 = std::chrono::steady_clock::now();
    }

    // uploads of this round's due frames (host_io), then a wait of the pool stream on them
    auto upload = [&](Pending &q, size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; k++) {
            const Live &L = live[who[k]];
            const size_t bytes = (size_t)kMix[L.cls].w * kMix[L.cls].h * 3;
            if (hipMemcpyAsync((void *)ins[k], in_frames[L.cls][(size_t)(fidx[k] % kFrames)], bytes,
                               hipMemcpyHostToDevice, s_up) != hipSuccess)
                return -1;
        }
        if (hipEventRecord(q.ev_up, s_up) != hipSuccess || hipStreamWaitEvent(st, q.ev_up, 0) != hipSuccess) return -1;
        return 0;
    };
    auto download = [&](Pending &q, size_t k0, size_t k1) {
        if (hipStreamWaitEvent(s_down, q.ev_end, 0) != hipSuccess) return -1;
        for (size_t k = k0; k < k1; k++) {
            const Live &L = live[who[k]];
            const size_t bytes = (size_t)kMix[L.cls].w * kMix[L.cls].h * 3;
            const int b = (int)(fidx[k] % kSlots);
            if (hipMemcpyAsync(L.sl.hout[b], outs[k], bytes, hipMemcpyDeviceToHost, s_down) != hipSuccess) return -1;
        }
        return hipEventRecord(q.ev_done, s_down) == hipSuccess ? 0 : -1;
    };
    auto record_done = [&](Pending &q, size_t k0, size_t k1) {
        q.done.clear();
        if (!cfg->on_frame) return;
        for (size_t k = k0; k < k1; k++) {
            const Live &L = live[who[k]];
            const int b = (int)(fidx[k] % kSlots);
            q.done.push_back({L.plan_idx, L.cls, fidx[k], host_io ? L.sl.hout[b] : L.sl.dout[b]});
        }
    };

    while (rc == 0) {
        const double now = now_s(t0);
        if (now >= horizon) break;
        // arrivals
        while (next_plan < plan.size() && plan[next_plan].arrive <= now) waiting.push_back((int)next_plan++);
        // departures: every session leaving now, after one wait for the ticks in flight
        bool leaving = false;
        for (const Live &L : live) leaving = leaving || L.end <= now || (quota && L.next >= quota);
        if (leaving) {
            if (drain_all() < 0) { rc = -1; break; }
            for (size_t i = 0; i < live.size();) {
                Live &L = live[i];
                if (L.end <= now || (quota && L.next >= quota)) {
                    wema_pool_remove(pool, L.id);
                    slots_pool.give(L.cls, L.sl);
                    live.erase(live.begin() + (long)i);
                } else {
                    ++i;
                }
            }
        }
        if (admit(now) < 0) { rc = -1; break; }
        if (live.empty() && waiting.empty() && next_plan == plan.size() && (quota || fixed)) break;
        out->peak_sessions = std::max(out->peak_sessions, (int)live.size());
        out->peak_bytes = std::max(out->peak_bytes, wema_pool_bytes_in_use(pool));
        // frames due now: batched, up to kRun of a session (a backlog is taken oldest first);
        // plain, one per session per round
        ids.clear(); ins.clear(); outs.clear(); due.clear(); who.clear(); fidx.clear();
        double next_due = 1e30;
        const int run = cfg->batched ? kRun : 1;
        for (size_t i = 0; i < live.size(); i++) {
            Live &L = live[i];
            for (int k = 0; k < run; k++) {
                const long f = L.next + k;
                if (quota && f >= quota) break;
                const double d = fixed ? now : L.start + (double)f / L.fps;
                if (d > now || (fixed && k > 0)) {
                    next_due = std::min(next_due, d);
                    break;
                }
                const int b = (int)(f % kSlots);
                ids.push_back(L.id);
                ins.push_back(host_io ? L.sl.din[b] : in_frames[L.cls][(size_t)(f % kFrames)]);
                outs.push_back(L.sl.dout[b]);
                due.push_back(d);
                who.push_back(i);
                fidx.push_back(f);
            }
        }
        if (ids.empty()) {
            if (!inflight.empty()) {
                // nothing due: see the running tick finish
                if (drain_one() < 0) { rc = -1; break; }
                continue;
            }
            double wake = next_due;
            if (next_plan < plan.size()) wake = std::min(wake, plan[next_plan].arrive);
            const double dt = std::min(wake, horizon) - now_s(t0);
            if (dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(dt, 0.002)));
            continue;
        }
        const bool measured = rounds >= cfg->warmup_rounds;
        long g0 = 0, g1 = 0;
        wema_pool_stats(pool, nullptr, &g0);
        if (cfg->batched) {
            Pending &q = slots[slot];
            slot ^= 1;
            if (!inflight.empty() && inflight.front() == &q && drain_one() < 0) { rc = -1; break; }
            const auto h0 = std::chrono::steady_clock::now();
            if (host_io && upload(q, 0, ids.size()) < 0) { wema_set_error("frame upload failed"); rc = -1; break; }
            if (hipEventRecord(q.ev_begin, st) != hipSuccess) { wema_set_error("event record failed"); rc = -1; break; }
            if (wema_pool_tick(pool, (int)ids.size(), ids.data(), ins.data(), outs.data(), nullptr) < 0) { rc = -1; break; }
            if (hipEventRecord(q.ev_end, st) != hipSuccess) { wema_set_error("event record failed"); rc = -1; break; }
            if (host_io) {
                if (download(q, 0, ids.size()) < 0) { wema_set_error("frame download failed"); rc = -1; break; }
            } else if (hipEventRecord(q.ev_done, st) != hipSuccess) {
                wema_set_error("event record failed");
                rc = -1;
                break;
            }
            const double hms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
            q.due = due;
            q.measured = measured;
            record_done(q, 0, ids.size());
            inflight.push_back(&q);
            if (inflight.size() > 1 && drain_one() < 0) { rc = -1; break; }
            if (measured) {
                ticks++;
                host_ms += hms;
            }
        } else {
            // the plain path: each frame through the device alone, one call per frame, two
            // frames in flight as the batched ticks are (the next frame's launches are
            // issued while the previous one runs)
            for (size_t k = 0; k < ids.size(); k++) {
                Pending &q = slots[slot];
                slot ^= 1;
                if (!inflight.empty() && inflight.front() == &q && drain_one() < 0) { rc = -1; break; }
                const auto h0 = std::chrono::steady_clock::now();
                if (host_io && upload(q, k, k + 1) < 0) { wema_set_error("frame upload failed"); rc = -1; break; }
                if (hipEventRecord(q.ev_begin, st) != hipSuccess ||
                    wema_pool_tick(pool, 1, &ids[k], &ins[k], &outs[k], nullptr) < 0 ||
                    hipEventRecord(q.ev_end, st) != hipSuccess) {
                    rc = -1;
                    break;
                }
                if (host_io) {
                    if (download(q, k, k + 1) < 0) { wema_set_error("frame download failed"); rc = -1; break; }
                } else if (hipEventRecord(q.ev_done, st) != hipSuccess) {
                    rc = -1;
                    break;
                }
                const double hms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
                q.due.assign(1, due[k]);
                q.measured = measured;
                record_done(q, k, k + 1);
                inflight.push_back(&q);
                if (inflight.size() > 1 && drain_one() < 0) { rc = -1; break; }
                if (measured) {
                    ticks++;
                    host_ms += hms;
                }
            }
            if (rc < 0) break;
        }
        wema_pool_stats(pool, nullptr, &g1);
        if (measured) graph_ticks += g1 - g0;
        for (size_t k = 0; k < who.size(); k++) {
            Live &L = live[who[k]];
            if (measured) {
                frames++;
                const double px = (double)kMix[L.cls].w * kMix[L.cls].h;
                mpix += px / 1e6;
                if (host_io) link_bytes += 6.0 * px;
            }
            L.next++;
        }
        if (measured) tick_frames += (long)ids.size();
        rounds++;
        if (!cfg->batched) t_last = now_s(t0);
        if (rounds == cfg->warmup_rounds) t_meas = now_s(t0);
    }
    if (drain_all() < 0) rc = -1;
    // a saturate run is measured up to its last round, a paced one over its whole span
    const double wall = (fixed ? t_last : now_s(t0)) - t_meas;
    hipStreamSynchronize(st);
    if (s_up) hipStreamSynchronize(s_up);
    if (s_down) hipStreamSynchronize(s_down);
    for (Pending &q : slots) {
        if (q.ev_up) hipEventDestroy(q.ev_up);
        if (q.ev_begin) hipEventDestroy(q.ev_begin);
        if (q.ev_end) hipEventDestroy(q.ev_end);
        if (q.ev_done) hipEventDestroy(q.ev_done);
    }
    for (Live &L : live) wema_pool_remove(pool, L.id);
    wema_pool_destroy(pool);
    slots_pool.release();
    if (s_up) hipStreamDestroy(s_up);
    if (s_down) hipStreamDestroy(s_down);
    free_frames();
    if (rc < 0) return -1;

    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
        if (lat.empty()) return 0.0;
        size_t i = (size_t)std::min((double)lat.size() - 1, floor(q * (double)(lat.size() - 1) + 0.5));
        return lat[i] * 1e3;
    };
    out->frames = frames;
    out->mpix = mpix;
    out->wall_s = wall;
    out->mpix_per_s = wall > 0 ? mpix / wall : 0.0;
    out->lat_p50_ms = pct(0.50);
    out->lat_p99_ms = pct(0.99);
    out->lat_max_ms = lat.empty() ? 0.0 : lat.back() * 1e3;
    out->ticks = ticks;
    out->mean_frames_per_tick = ticks ? (double)tick_frames / (double)ticks : 0.0;
    out->host_ms_per_tick = ticks ? host_ms / (double)ticks : 0.0;
    out->graph_tick_frac = ticks ? (double)graph_ticks / (double)ticks : 0.0;
    out->gpu_busy_frac = wall > 0 ? gpu_ms / 1e3 / wall : 0.0;
    out->link_gbs = wall > 0 ? link_bytes / wall / 1e9 : 0.0;
    return 0;
}
