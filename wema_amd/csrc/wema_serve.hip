// Serving loop over a wema_pool: stream sessions arrive as a Poisson process, are admitted
// as HBM (budget) and slots allow, deliver frames at their own frame rate, and leave after
// their lifetime. Every iteration gathers the frames that are due and runs them as one
// pool tick (continuous batching: sessions that arrive while others are on the GPU join
// the next tick); the plain mode processes the same frames one session at a time.
// Reports per-frame latency (due time -> output complete) and aggregate Mpixel/s.
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

struct Planned {
    double arrive, life, phase;
    wema_dev_params p;
    int cls;
};

struct Live {
    int id;
    int cls;
    double start, end;      // seconds since run start
    double fps;
    long next;              // index of the next frame to deliver
    uint8_t *out;           // device output slot
    size_t out_bytes;
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" int wema_serve_run(const wema_serve_config *cfg, wema_serve_stats *out) {
    if (!cfg || !out) { wema_set_error("null argument"); return -1; }
    memset(out, 0, sizeof(*out));
    if (hipSetDevice(cfg->device) != hipSuccess) { wema_set_error("hipSetDevice failed"); return -1; }
    Rng rng{cfg->seed ? cfg->seed : 0x57454D41ull};

    // ---- arrival plan (deterministic for a seed)
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
    out->sessions_arrived = (int)plan.size();

    // ---- input frames: a few per resolution class, resident in HBM, cycled
    const int kFrames = 4;
    std::vector<uint8_t *> in_frames[4];
    for (int c = 0; c < 4; c++) {
        wema_synth_params sp;
        wema_synth_preset(&sp, 4, kMix[c].w, kMix[c].h, kMix[c].fps, (uint32_t)c);
        const size_t P = (size_t)kMix[c].w * kMix[c].h;
        std::vector<uint16_t> tex(P);
        wema_synth_texture_host(&sp, tex.data());
        uint16_t *dtex = nullptr;
        uint8_t *buf = nullptr;
        if (hipMalloc(&dtex, P * 2) != hipSuccess || hipMalloc(&buf, P * 3 * kFrames) != hipSuccess) {
            wema_set_error("frame allocation failed");
            return -1;
        }
        hipMemcpy(dtex, tex.data(), P * 2, hipMemcpyHostToDevice);
        if (wema_synth_frames_device(&sp, dtex, 0, kFrames, buf, nullptr) < 0) return -1;
        hipFree(dtex);
        for (int k = 0; k < kFrames; k++) in_frames[c].push_back(buf + P * 3 * k);
    }

    wema_pool_t *pool = nullptr;
    if (wema_pool_create(&pool, cfg->device, cfg->hbm_budget) < 0) return -1;
    wema_pool_graphs(pool, cfg->use_graphs);
    hipStream_t st = (hipStream_t)wema_pool_stream(pool);

    std::vector<double> lat;
    lat.reserve(1 << 20);
    std::deque<int> waiting;      // planned sessions not yet admitted
    std::vector<Live> live;
    size_t next_plan = 0;
    long frames = 0, ticks = 0, tick_frames = 0, rounds = 0;
    double mpix = 0.0;
    int rc = 0;
    auto t0 = std::chrono::steady_clock::now();
    double t_meas = 0.0;          // start of the measured part (after the warm-up rounds)
    double t_last = 0.0;          // end of the last round
    const double horizon = cfg->saturate ? cfg->duration_s : cfg->duration_s + cfg->drain_s;
    const long quota = cfg->frames_per_session > 0 ? cfg->frames_per_session + cfg->warmup_rounds : 0;
    std::vector<int> ids;
    std::vector<const uint8_t *> ins;
    std::vector<uint8_t *> outs;
    std::vector<double> due;
    std::vector<size_t> who;

    // batched ticks are pipelined: a tick is launched before the previous one is waited
    // for, so the host gathers and builds the next tick while the GPU runs this one
    struct Pending {
        hipEvent_t ev = nullptr;
        std::vector<double> due;
        bool measured = false;
    };
    Pending slots[2];
    for (Pending &q : slots)
        if (hipEventCreateWithFlags(&q.ev, hipEventDisableTiming) != hipSuccess) {
            wema_set_error("event creation failed");
            return -1;
        }
    std::deque<Pending *> inflight;
    int slot = 0;
    auto drain_one = [&]() {
        Pending *q = inflight.front();
        inflight.pop_front();
        if (hipEventSynchronize(q->ev) != hipSuccess) { wema_set_error("tick failed"); return -1; }
        const double done = now_s(t0);
        if (q->measured)
            for (double d : q->due) lat.push_back(done - d);
        t_last = done;
        return 0;
    };
    auto drain_all = [&]() {
        while (!inflight.empty())
            if (drain_one() < 0) return -1;
        return 0;
    };

    // admission (first come, first served; the head waits until it fits)
    auto admit = [&](double now) {
        while (!waiting.empty()) {
            const Planned &pl = plan[(size_t)waiting.front()];
            if (cfg->max_sessions > 0 && (int)live.size() >= cfg->max_sessions) break;
            int id = -1;
            if (wema_pool_add(pool, &pl.p, &id) < 0) {
                out->admission_waits++;
                break;
            }
            Live L;
            L.id = id;
            L.cls = pl.cls;
            // streams that join in steady state run at a random phase of their frame period
            L.start = now + (warm && pl.arrive == 0.0 ? pl.phase / pl.p.fps : 0.0);
            L.end = now + pl.life;
            L.fps = pl.p.fps;
            L.next = 0;
            L.out_bytes = (size_t)pl.p.width * pl.p.height * 3;
            if (hipMalloc(&L.out, L.out_bytes) != hipSuccess) {
                wema_pool_remove(pool, id);
                wema_set_error("output slot allocation failed");
                return -1;
            }
            live.push_back(L);
            out->sessions_admitted++;
            waiting.pop_front();
        }
        return 0;
    };

    if (warm) {
        // the steady-state population joins before the clock starts
        while (next_plan < plan.size() && plan[next_plan].arrive <= 0.0) waiting.push_back((int)next_plan++);
        if (admit(0.0) < 0) rc = -1;
        t0 = std::chrono::steady_clock::now();
    }

    while (rc == 0) {
        const double now = now_s(t0);
        if (now >= horizon) break;
        // arrivals
        while (next_plan < plan.size() && plan[next_plan].arrive <= now) waiting.push_back((int)next_plan++);
        // departures
        for (size_t i = 0; i < live.size();) {
            if (live[i].end <= now || (quota && live[i].next >= quota)) {
                if (drain_all() < 0) { rc = -1; break; }
                hipStreamSynchronize(st);
                wema_pool_remove(pool, live[i].id);
                hipFree(live[i].out);
                live.erase(live.begin() + (long)i);
            } else {
                ++i;
            }
        }
        if (admit(now) < 0) { rc = -1; break; }
        if (live.empty() && waiting.empty() && next_plan == plan.size() && (quota || fixed)) break;
        out->peak_sessions = std::max(out->peak_sessions, (int)live.size());
        out->peak_bytes = std::max(out->peak_bytes, wema_pool_bytes_in_use(pool));
        // frames due now (one per session per tick; a backlog keeps its oldest frame)
        ids.clear(); ins.clear(); outs.clear(); due.clear(); who.clear();
        double next_due = 1e30;
        for (size_t i = 0; i < live.size(); i++) {
            Live &L = live[i];
            const double d = cfg->saturate ? now : L.start + (double)L.next / L.fps;
            if (d <= now) {
                ids.push_back(L.id);
                ins.push_back(in_frames[L.cls][(size_t)(L.next % kFrames)]);
                outs.push_back(L.out);
                due.push_back(d);
                who.push_back(i);
            } else {
                next_due = std::min(next_due, d);
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
        if (cfg->batched) {
            Pending &q = slots[slot];
            slot ^= 1;
            if (!inflight.empty() && inflight.front() == &q && drain_one() < 0) { rc = -1; break; }
            if (wema_pool_tick(pool, (int)ids.size(), ids.data(), ins.data(), outs.data(), nullptr) < 0) { rc = -1; break; }
            if (hipEventRecord(q.ev, st) != hipSuccess) { wema_set_error("event record failed"); rc = -1; break; }
            q.due = due;
            q.measured = measured;
            inflight.push_back(&q);
            if (inflight.size() > 1 && drain_one() < 0) { rc = -1; break; }
            if (measured) ticks++;
        } else {
            for (size_t k = 0; k < ids.size(); k++) {
                if (wema_pool_tick(pool, 1, &ids[k], &ins[k], &outs[k], nullptr) < 0) { rc = -1; break; }
                if (hipStreamSynchronize(st) != hipSuccess) { wema_set_error("tick failed"); rc = -1; break; }
                if (measured) {
                    lat.push_back(now_s(t0) - due[k]);
                    ticks++;
                }
            }
            if (rc < 0) break;
        }
        for (size_t k = 0; k < who.size(); k++) {
            Live &L = live[who[k]];
            L.next++;
            if (measured) {
                frames++;
                mpix += (double)kMix[L.cls].w * kMix[L.cls].h / 1e6;
            }
        }
        if (measured) tick_frames += (long)ids.size();
        rounds++;
        if (!cfg->batched) t_last = now_s(t0);
        if (rounds == cfg->warmup_rounds) t_meas = now_s(t0);
    }
    if (drain_all() < 0) rc = -1;
    // a saturate run is measured up to its last round, a paced one over its whole span
    const double wall = (cfg->saturate ? t_last : now_s(t0)) - t_meas;
    hipStreamSynchronize(st);
    for (Pending &q : slots) hipEventDestroy(q.ev);
    for (Live &L : live) {
        wema_pool_remove(pool, L.id);
        hipFree(L.out);
    }
    wema_pool_destroy(pool);
    for (int c = 0; c < 4; c++)
        if (!in_frames[c].empty()) hipFree(in_frames[c][0]);
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
    out->mpix_per_s = mpix / wall;
    out->lat_p50_ms = pct(0.50);
    out->lat_p99_ms = pct(0.99);
    out->lat_max_ms = lat.empty() ? 0.0 : lat.back() * 1e3;
    out->ticks = ticks;
    out->mean_frames_per_tick = ticks ? (double)tick_frames / (double)ticks : 0.0;
    return 0;
}
