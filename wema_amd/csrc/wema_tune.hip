// Tile-shape autotuner of the guided-filter kernels. They come in tile heights of 8, 16
// and 32 rows (64 columns each); which is fastest depends on the frame size and the pass
// (edge tiles, blocks per XCD, LDS per block). The first session of a frame size on a
// device times the three variants of each pass on frames of that size and keeps the
// fastest of each for the process. Tile shapes never change results (the box sums are
// anchored to the image), so the two passes may use different heights.
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

namespace {

std::mutex g_mu;
std::map<std::tuple<int, int, int>, std::pair<int, int>> g_rows;   // (device, w, h) -> rows of each pass
bool g_enabled = true;
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
    if (!g_enabled || w < 4 || h < 4) return;
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

extern "C" void wema_dev_autotune(int enable) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_enabled = enable != 0;
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
