// Multi-stream pool: many stream sessions on one GPU, admitted under an HBM budget, whose
// frames are processed together, one launch per pipeline stage per tick over the job
// tables of all streams in the tick. A tick's launch sequence can be captured once in a
// hipGraph per tick shape (which sessions, which output stage) and replayed; the job
// tables (pointers, frame-dependent amplification, ring slots) are uploaded from pinned
// memory by the graph's first node on every replay.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/wema_dev.h"
#include "wema_host.h"

int wema_session_setup(wema_dev *s, int with_filt, bool pooled);

#define HIP_OK(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            wema_set_error("%s failed: %s", #call, hipGetErrorString(e_));                 \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

struct StageBuf {
    void *h = nullptr;       // pinned job tables
    void *d = nullptr;       // device copy
    hipEvent_t ev = nullptr; // set when the last upload from h finished
};

struct GraphEntry {
    StageBuf buf;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

struct wema_pool {
    int device = 0;
    size_t budget = 0, in_use = 0;
    std::vector<wema_dev *> sess;      // slot -> session (null = free)
    std::vector<size_t> sess_bytes;
    hipStream_t stream = nullptr;
    StageBuf direct[2];
    int direct_cur = 0;
    bool graphs = false;
    std::map<std::string, GraphEntry> cache;
    std::map<std::string, int> seen;   // tick shapes met so far (captured from the 2nd time)
    LaunchTimer timer;
    long ticks = 0;
    long graph_ticks = 0;   // ticks replayed from a captured hipGraph
};

static int stage_alloc(StageBuf &b) {
    HIP_OK(hipHostMalloc(&b.h, WEMA_STAGE_BYTES, hipHostMallocDefault));
    HIP_OK(hipMalloc(&b.d, WEMA_STAGE_BYTES));
    HIP_OK(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
    return 0;
}

static void stage_free(StageBuf &b) {
    if (b.h) hipHostFree(b.h);
    if (b.d) hipFree(b.d);
    if (b.ev) hipEventDestroy(b.ev);
    b = StageBuf();
}

extern "C" int wema_pool_create(wema_pool_t **out, int device, size_t hbm_budget) {
    if (!out) { wema_set_error("null argument"); return -1; }
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) { wema_set_error("hipSetDevice(%d) failed", device); return -1; }
    wema_pool *p = new wema_pool();
    p->device = device;
    p->budget = hbm_budget;
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
        stage_alloc(p->direct[0]) < 0 || stage_alloc(p->direct[1]) < 0) {
        wema_set_error("pool allocation failed");
        wema_pool_destroy(p);
        return -1;
    }
    *out = p;
    return 0;
}

extern "C" void wema_pool_destroy(wema_pool_t *p) {
    if (!p) return;
    hipSetDevice(p->device);
    if (p->stream) hipStreamSynchronize(p->stream);
    for (auto &kv : p->cache) {
        if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) hipGraphDestroy(kv.second.graph);
        stage_free(kv.second.buf);
    }
    for (wema_dev *s : p->sess)
        if (s) wema_dev_close(s);
    stage_free(p->direct[0]);
    stage_free(p->direct[1]);
    if (p->stream) hipStreamDestroy(p->stream);
    delete p;
}

extern "C" int wema_pool_add(wema_pool_t *p, const wema_dev_params *prm, int *id) {
    if (!p || !prm || !id) { wema_set_error("null argument"); return -1; }
    if (prm->device != p->device) { wema_set_error("session device differs from the pool's"); return -1; }
    size_t bytes = 0;
    if (wema_dev_query_bytes(prm, &bytes) < 0) return -1;
    if (p->in_use + bytes > p->budget) {
        wema_set_error("HBM budget exceeded: %zu in use + %zu requested > %zu", p->in_use, bytes, p->budget);
        return -1;
    }
    wema_dev_t *s = nullptr;
    if (wema_dev_open_ex(&s, prm, true) < 0) return -1;
    int slot = -1;
    for (size_t i = 0; i < p->sess.size(); i++)
        if (!p->sess[i]) { slot = (int)i; break; }
    if (slot < 0) {
        slot = (int)p->sess.size();
        p->sess.push_back(nullptr);
        p->sess_bytes.push_back(0);
    }
    p->sess[slot] = s;
    p->sess_bytes[slot] = bytes;
    p->in_use += bytes;
    *id = slot;
    return 0;
}

static wema_dev *get(wema_pool *p, int id) {
    if (id < 0 || id >= (int)p->sess.size() || !p->sess[id]) {
        wema_set_error("no session %d in the pool", id);
        return nullptr;
    }
    return p->sess[id];
}

extern "C" int wema_pool_remove(wema_pool_t *p, int id) {
    if (!p) { wema_set_error("null pool"); return -1; }
    wema_dev *s = get(p, id);
    if (!s) return -1;
    HIP_OK(hipStreamSynchronize(p->stream));
    // graphs that include this session are stale
    for (auto it = p->cache.begin(); it != p->cache.end();) {
        if (it->first.find("|" + std::to_string(id) + ":") != std::string::npos) {
            hipGraphExecDestroy(it->second.exec);
            hipGraphDestroy(it->second.graph);
            stage_free(it->second.buf);
            it = p->cache.erase(it);
        } else {
            ++it;
        }
    }
    wema_dev_close(s);
    p->sess[id] = nullptr;
    p->in_use -= p->sess_bytes[id];
    p->sess_bytes[id] = 0;
    return 0;
}

extern "C" int wema_pool_session(wema_pool_t *p, int id, wema_dev_t **s) {
    if (!p || !s) { wema_set_error("null argument"); return -1; }
    *s = get(p, id);
    return *s ? 0 : -1;
}

extern "C" size_t wema_pool_bytes_in_use(const wema_pool_t *p) { return p ? p->in_use : 0; }

extern "C" int wema_pool_graphs(wema_pool_t *p, int enable) {
    if (!p) { wema_set_error("null pool"); return -1; }
    p->graphs = enable != 0;
    return 0;
}

extern "C" int wema_pool_timing(wema_pool_t *p, int enable) {
    if (!p) { wema_set_error("null pool"); return -1; }
    if (!enable && p->timer.enabled && wema_timer_collect(&p->timer) < 0) return -1;
    p->timer.enabled = enable != 0;
    if (enable) {
        memset(p->timer.ms, 0, sizeof(p->timer.ms));
        memset(p->timer.bytes, 0, sizeof(p->timer.bytes));
        memset(p->timer.launches, 0, sizeof(p->timer.launches));
    }
    return 0;
}

extern "C" int wema_pool_timing_report(wema_pool_t *p, int kind, char *name, int name_len,
                                       long *launches, double *total_ms, double *bytes) {
    if (!p || kind < 0 || kind >= TK_COUNT) return -1;
    if (wema_timer_collect(&p->timer) < 0) return -1;
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", wema_timer_names[kind]);
    if (launches) *launches = p->timer.launches[kind];
    if (total_ms) *total_ms = p->timer.ms[kind];
    if (bytes) *bytes = p->timer.bytes[kind];
    return 0;
}

// Output kind a session's next frame uses (part of a tick's launch shape); EG_KINDS stands
// for the colour output of the inverse stage.
static int egress_kind(const wema_dev *s) {
    if (!s->p.use_iir && s->frames + 1 < s->p.window) return EG_PASSTHRU;
    if (s->p.color) return EG_KINDS;
    // (the guided filter with its segment height, 0: the launch's split)
    return (s->p.edge_aware || s->p.batch_path) ? EG_KINDS + 1 + s->gf_seg : EG_GAUSS;
}

extern "C" int wema_pool_tick(wema_pool_t *p, int n, const int *ids, const uint8_t *const *d_in,
                              uint8_t *const *d_out, void *hip_stream) {
    if (!p || n < 1 || !ids || !d_in || !d_out) { wema_set_error("bad argument"); return -1; }
    HIP_OK(hipSetDevice(p->device));
    TraceRange trace("wema_pool_tick");
    hipStream_t st = hip_stream ? (hipStream_t)hip_stream : p->stream;
    // runs of equal ids: consecutive frames of one session (at most its max_batch)
    struct Run {
        wema_dev *s;
        int first, len;
    };
    std::vector<Run> runs;
    std::string key;
    for (int i = 0; i < n;) {
        int k = i + 1;
        while (k < n && ids[k] == ids[i]) k++;
        wema_dev *s = get(p, ids[i]);
        if (!s) return -1;
        if (s->failed) { wema_set_error("session %d failed in an earlier tick; remove it", ids[i]); return -1; }
        for (const Run &r : runs)
            if (r.s == s) { wema_set_error("session %d in two runs of one tick", ids[i]); return -1; }
        if (k - i > s->p.max_batch) {
            wema_set_error("session %d: %d frames in one tick, max_batch %d", ids[i], k - i, s->p.max_batch);
            return -1;
        }
        runs.push_back({s, i, k - i});
        key += "|" + std::to_string(ids[i]) + ":" + std::to_string(egress_kind(s)) + "x" + std::to_string(k - i);
        i = k;
    }
    JobBuilder jb;
    for (const Run &r : runs)
        wema_build_jobs(r.s, jb, d_in[r.first], 0, d_out[r.first], nullptr, r.len, nullptr, (const void *const *)&d_in[r.first],
                        &d_out[r.first]);

    // a shape is captured the second time it occurs (one-off shapes run directly), and the
    // cache keeps at most kMaxGraphs shapes
    constexpr size_t kMaxGraphs = 64;
    bool use_graph = p->graphs && !p->timer.enabled && !hip_stream;
    if (use_graph && !p->cache.count(key)) {
        int &cnt = p->seen[key];
        cnt++;
        if (p->seen.size() > 4096) p->seen.clear();
        if (cnt < 2 || p->cache.size() >= kMaxGraphs) use_graph = false;
    }
    if (!use_graph) {
        StageBuf &b = p->direct[p->direct_cur];
        p->direct_cur ^= 1;
        HIP_OK(hipEventSynchronize(b.ev));
        const int rc = wema_run_jobs(jb, b.h, b.d, b.ev, st, &p->timer);
        if (rc < 0) {
            for (const Run &r : runs) r.s->failed = r.s->failed || rc == -2;
            return -1;
        }
    } else {
        auto it = p->cache.find(key);
        if (it == p->cache.end()) {
            GraphEntry g;
            if (stage_alloc(g.buf) < 0) return -1;
            HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int rc = wema_run_jobs(jb, g.buf.h, g.buf.d, nullptr, st, nullptr);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(st, &graph);
            if (rc < 0 || e != hipSuccess) {
                if (graph) hipGraphDestroy(graph);
                stage_free(g.buf);
                if (rc == 0) wema_set_error("graph capture failed: %s", hipGetErrorString(e));
                return -1;
            }
            g.graph = graph;
            if (hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) != hipSuccess) {
                hipGraphDestroy(graph);
                stage_free(g.buf);
                wema_set_error("hipGraphInstantiate failed");
                return -1;
            }
            it = p->cache.emplace(key, g).first;
        } else {
            // tables of this tick into the graph's pinned buffer once its last upload ran
            HIP_OK(hipEventSynchronize(it->second.buf.ev));
            size_t used = 0;
            if (jb.pack(it->second.buf.h, it->second.buf.d, WEMA_STAGE_BYTES, &used) < 0) {
                wema_set_error("job tables exceed %d bytes", WEMA_STAGE_BYTES);
                return -1;
            }
        }
        HIP_OK(hipGraphLaunch(it->second.exec, st));
        HIP_OK(hipEventRecord(it->second.buf.ev, st));
        p->graph_ticks++;
    }
    for (const Run &r : runs) {
        r.s->frames += r.len;
        r.s->last_n = r.len;
        r.s->last_stream = st;
        r.s->thr_ovr_active = 0;
    }
    p->ticks++;
    return 0;
}

extern "C" int wema_pool_sync(wema_pool_t *p) {
    if (!p) { wema_set_error("null pool"); return -1; }
    HIP_OK(hipSetDevice(p->device));
    HIP_OK(hipStreamSynchronize(p->stream));
    return 0;
}

extern "C" void *wema_pool_stream(wema_pool_t *p) { return p ? (void *)p->stream : nullptr; }

extern "C" int wema_pool_stats(const wema_pool_t *p, long *ticks, long *graph_ticks) {
    if (!p) { wema_set_error("null pool"); return -1; }
    if (ticks) *ticks = p->ticks;
    if (graph_ticks) *graph_ticks = p->graph_ticks;
    return 0;
}
