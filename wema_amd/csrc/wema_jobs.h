// Work descriptors shared by the host scheduler (wema_dev.hip, wema_pool.hip) and the
// kernels (wema_kernels.hip).
//
// Every stage is one launch over a table of jobs. A job is one (stream, channel, frame)
// unit of that stage with its own sizes and buffers; the blocks of a launch are dealt out
// to the jobs in table order (job j owns blocks [tile_begin[j], tile_begin[j+1])). A single
// stream and a frame batch of many streams use the same kernels and differ only in the
// table, so a stream's results do not depend on which other streams share the launch.
#pragma once
#include <stdint.h>

// CDF 9/7 lifting factors (Daubechies & Sweldens factorisation) and the normalisation the
// reference uses (src/dwt.c:31-36).
#define WEMA_ALPHA (-1.586134342059924f)
#define WEMA_BETA (-0.052980118572961f)
#define WEMA_GAMMA (0.882911075530934f)
#define WEMA_DELTA (0.443506852043971f)
#define WEMA_K (1.149604398860241f)
#define WEMA_KINV (1.0f / WEMA_K)

// DWT strips: one wave computes WEMA_TILE_PAIRS_X = 60 pair columns (64 lanes less the
// 2-pair border on each side) and sp pair rows (sp = 32, 16 or 8, per launch) of a level,
// streaming sp + 4 pair rows; a block holds WEMA_STRIPS_PER_BLOCK strips of a job.
#define WEMA_TILE_PAIRS_X 60
#define WEMA_SP_MAX 32
#define WEMA_SP_MIN 8
#ifndef WEMA_STRIPS_PER_BLOCK
#define WEMA_STRIPS_PER_BLOCK 4
#endif
// The levels below full resolution can also run as LDS tiles of WEMA_DT_TP x WEMA_DT_TP
// output pairs, one per 256-thread workgroup (a job's strip height WEMA_SP_TILE selects
// them); levels of fewer than 2 samples in either direction always run as strips.
#ifndef WEMA_DT_TP
#define WEMA_DT_TP 16
#endif
#define WEMA_SP_TILE 1
__host__ __device__ static inline int wema_dt_tiles(int w, int h) {
    const int sw = (w + 1) >> 1, sh = (h + 1) >> 1;
    return ((sw + WEMA_DT_TP - 1) / WEMA_DT_TP) * ((sh + WEMA_DT_TP - 1) / WEMA_DT_TP);
}
// ... or with one thread per output pair, which recomputes the lifting of its own
// neighbourhood in registers (no LDS, no barriers; strip height value WEMA_SP_POINT), for
// levels so small that a strip's or a tile's chain of dependent steps is what their launch
// waits for. 256 pairs per workgroup.
#define WEMA_SP_POINT 2
__host__ __device__ static inline int wema_pt_blocks(int w, int h) {
    const int sw = (w + 1) >> 1, sh = (h + 1) >> 1;
    return (sw * sh + 255) / 256;
}
// Strip height per launch: the tallest that still gives this many channel-strips (a colour
// strip counts 3). Measured on 1080p colour and 4K luma: the forward DWT balances halo
// re-reads against parallelism at about 64 K; the inverse is fastest with the shortest
// strips (8 pair rows) unless a launch is far larger.
#ifndef WEMA_STRIP_TARGET
#define WEMA_STRIP_TARGET 65536
#endif
#ifndef WEMA_STRIP_TARGET_INV
#define WEMA_STRIP_TARGET_INV 262144
#endif
// The same for the levels below full resolution (float LL in, one channel per wave).
#ifndef WEMA_STRIP_TARGET_COARSE
#define WEMA_STRIP_TARGET_COARSE 65536
#endif
#ifndef WEMA_STRIP_TARGET_INV_COARSE
#define WEMA_STRIP_TARGET_INV_COARSE 262144
#endif

// 5x5 Gaussian tiles (outputs per block).
#define WEMA_GF_TX 64
#define WEMA_GF_TY 16
// Guided filter: output columns of a strip (a workgroup's 64 lanes x 4 columns less the
// 8-column halo on each side).
#define WEMA_GF_OUT 240

enum SrcKind {
    SRC_F32_PLANE = 0,   // float image (the previous level's LL)
    SRC_U8_Y = 1,        // RGB24 -> luma (src/wema.c:42-51)
    SRC_U8_CB = 2,       // RGB24 -> Cb (src/wema.c:60-78)
    SRC_U8_CR = 3,
    SRC_F32_Y = 4,       // float RGB -> luma
    SRC_F32_CB = 5,
    SRC_F32_CR = 6,
};

// Sum of |c| of a frame and channel: acc_slots Q32 integers that strips add into (strip s
// into slot s mod acc_slots) and whose total is the sum. One address per frame and channel
// would take every strip's atomic in turn: a one-frame 4K launch has 4,320 strips. A call of
// n frames uses the largest power of two up to 64 / n slots (at least 1), so the temporal
// kernel reads at most max(n, 64) of them per block.
#define WEMA_ACC_SLOTS_MAX 64
__host__ __device__ static inline int wema_acc_slots(int nframes) {
    int s = WEMA_ACC_SLOTS_MAX;
    while (s > 1 && s * nframes > WEMA_ACC_SLOTS_MAX) s >>= 1;
    return s;
}
// Accumulator entries a call of up to nframes frames of nch channels uses.
__host__ __device__ static inline int wema_acc_entries(int nch, int nframes) {
    return nch * (nframes > WEMA_ACC_SLOTS_MAX ? nframes : WEMA_ACC_SLOTS_MAX);
}

// A forward job covers nch channels of one frame that share the source (level 0 of the
// colour mode: Y, Cb, Cr from one RGB frame); channel c uses ll + c * ll_cstride,
// planes + c * planes_cstride and acc + c * acc_slots. Each wave computes its strip of
// every channel from one load of the source.
struct DwtfJob {
    const void *src;                // level 0: RGB frame; else previous level's LL (w x h)
    float *ll;                      // LL of this level (sw x sh), channel 0
    float *planes;                  // LH, HH, HL of this level (3 x sw x sh), channel 0
    const float *prev;              // previous level's planes (3 x w x h) or null (level 0)
    unsigned long long *acc;        // sum of |c| over unique positions, Q32, in acc_slots parts (or null)
    long long ll_cstride, planes_cstride;
    int w, h;                       // input size of this level
    int acc_slots;                  // a power of two
    int src_kind;
    int tiles_x;
    int tile_begin;
    int nch;
    int sp;                         // strip height in output pairs
};

// Inverse job. The colour output job (k_dwt_inv0_ycc) synthesises level 0 of Y, Cb and Cr
// (channel c at ll + c * ll_cstride, planes + c * planes_cstride) and writes RGB24 and/or
// float RGB directly; the other jobs write one float plane to out.
struct DwtiJob {
    const float *ll;                // approximation at this level (sw x sh)
    const float *planes;            // LH, HH, HL (amplified) of this level
    float *out;                     // reconstruction (w x h)
    unsigned long long *acc_clear;  // accumulators to reset once read (acc_n of them, or null)
    uint8_t *out_u8;                // colour output job: RGB24 frame (or null)
    float *out_f32;                 // colour output job: float RGB frame (or null)
    long long ll_cstride, planes_cstride;
    int w, h;
    int tiles_x;
    int tile_begin;
    int nch;
    int sp;                         // strip height in coefficient pair rows
    int acc_n;
};

#define WEMA_MAX_BATCH 256
// Frames whose coefficient loads the IIR kernel issues together (the state recurrence
// then runs through them).
#ifndef WEMA_IIR_G
#define WEMA_IIR_G 12   // measured against 8, 10, 16, 20, 32: 2-3% faster than 16 at 48 and 60 frames
#endif

// Output stages after the inverse DWT (colour mode has none: its inverse writes RGB). The
// guided filter has two kernels: k_gf for RGB24 in and out at widths that are multiples of
// 4 with aligned buffers (wema_gf_fast), k_gf_any for the rest.
enum EgressKind { EG_GF = 0, EG_GF_ANY = 1, EG_GAUSS = 2, EG_PASSTHRU = 3, EG_KINDS = 4 };

// Sliding-window kernels: 0 = windows of <= 32 samples through LDS, 1 = <= 64 through LDS,
// 2 = any size from the ring; exactly 32, 16 or 64 samples in registers: 3, 4, 5.
#define WEMA_WIN_KINDS 6
#ifndef WEMA_WIN_REG_MASK
#define WEMA_WIN_REG_MASK 7   // register kernels in use: 1 = 32, 2 = 16, 4 = 64
#endif
__host__ __device__ static inline int wema_win_kind(int ws) {
    if (ws == 32 && (WEMA_WIN_REG_MASK & 1)) return 3;
    if (ws == 16 && (WEMA_WIN_REG_MASK & 2)) return 4;
    if (ws == 64 && (WEMA_WIN_REG_MASK & 4)) return 5;
    return ws <= 32 ? 0 : (ws <= 64 ? 1 : 2);
}

// Temporal stage (IIR or sliding window) over one (stream, channel) and its frames.
struct TemporalJob {
    float *coef;                    // [nframes][coef_stride] coefficients (IIR: in place)
    float *amp_out;                 // window mode: [nframes][coef_stride] amplified out
    float *filt_out;                // optional [nframes][coef_stride] temporal output
    float4 *state;                  // IIR: per position HP w1, HP w2, LP w1, LP w2
    const float *ring;              // window mode: ring [ring_slots][coef_stride]
    const unsigned long long *acc;  // [nframes] stride acc_stride, acc_slots parts each
    const float *thr_override;      // [nframes] stride 3, NaN = compute (or null)
    float *thr_out;                 // [nframes] stride 3 (or null)
    const float *amp;               // [nframes] effective amplification of each frame
    const float *ltw;               // window mode: log2 of the bilateral time weights [ws]
    long long out_stride;           // frame stride of amp_out / filt_out
    int amp_index;                  // host-side index of amp in the uploaded amp array
    float c[10];                    // HP b0 b1 b2 a1 a2, LP b0 b1 b2 a1 a2
    long long n;                    // unique positions
    long long level0_end;           // positions below this are level 0 (alpha * 0.5)
    long long coef_stride;
    int acc_stride;
    int acc_slots;                  // a power of two, acc_slots * nframes <= max(nframes, 64)
    int nframes;
    int preseed;                    // run the warm-up on frame 0 first
    int warmup;
    float n6f;                      // (float) reference position count
    int window;                     // window mode: ws
    int ring_slots;
    int ring_first;                 // ring slot of the oldest sample of frame 0's window
    int bilateral;
    int tile_begin;
    int tiles;
    int ppt;                        // IIR, one frame: positions per thread (256 ppt per block)
};

// IIR jobs of one frame take this many positions per thread (measured against 1, 4 and 8,
// with the loads issued ahead of the threshold sum: 1-3% faster per frame than 4).
#ifndef WEMA_IIR_PPT
#define WEMA_IIR_PPT 2
#endif

// Guided filter / Gaussian / colour egress over one output frame.
struct EgressJob {
    const float *gray;              // luma reconstruction (or Y for colour)
    const float *cb, *cr;           // colour mode reconstructions
    const uint8_t *src_u8;          // source frame (RGB24) or null
    const float *src_f32;           // source frame (float RGB) or null
    uint8_t *out_u8;                // RGB24 out (or null)
    float *out_f32;                 // float RGB out (or null)
    int w, h;
    int tiles_x;                    // tiles (guided filter: strips) per row of tiles
    int tile_begin;
    int seg;                        // guided filter: rows per segment (a multiple of 4)
};

__host__ __device__ static inline int wema_ceil_div(int a, int b) { return (a + b - 1) / b; }

static inline bool wema_gf_fast(const EgressJob &j) {
    return j.src_u8 && j.out_u8 && !j.src_f32 && !j.out_f32 && (j.w & 3) == 0 && ((uintptr_t)j.gray & 15) == 0 &&
           (((uintptr_t)j.src_u8 | (uintptr_t)j.out_u8) & 3) == 0;
}
