// gfx950 kernels of the WEMA per-frame amplification pipeline.
//
// All stages are bandwidth-bound stencils or recurrences (no MFMA). Arithmetic is FP32 with
// the fused multiply-adds written out explicitly (the library is built with
// -ffp-contract=off), in the order and contraction shape the reference's -Ofast build uses,
// so that DWT, IIR and amplification results are bit-identical to it.
//
// Stages (one launch per stage over a job table, see wema_jobs.h):
//   k_dwt_fwd0/k_dwt_fwd  one 2D CDF 9/7 analysis level (level 0 / coarser levels) as
//               strips: a wave streams a 60-pair-wide column strip top to bottom, row
//               lifting across its 64 lanes (DPP lane shifts), column lifting as a register
//               pipeline; level 0 converts RGB24 to Y (and Cb, Cr) on load; writes LL +
//               LH/HH/HL and adds sum|c| (Q32 integers) to the frame's counter.
//   k_iir       Butterworth HP+LP biquads (DF2T) per coefficient over the job's frames with
//               the state kept in registers, warm-up preseed, threshold mask, amplification;
//               k_iir1 the same for launches of one-frame jobs only.
//   k_win       sliding-window temporal filter (mean removal, optional bilateral weights)
//               + amplification.
//   k_dwt_inv   one 2D synthesis level as strips (column synthesis as a register
//               pipeline, then rows across lanes); k_dwt_inv0 level 0.
//   k_dwt_inv0_ycc  colour mode: level 0 of Y, Cb, Cr + YCbCr -> RGB24 in one pass.
//   k_dwt_fwd_tile/k_dwt_inv_tile, k_dwt_fwd_point/k_dwt_inv_point  the coarser levels as
//               LDS tiles, or one output pair per thread (autotuner candidates).
//   k_gf        guided filter (r = 4, eps = 0.01) of the luma delta with 9x9 clipped box
//               means, streamed down column strips with a and b kept on the chip; adds
//               the smoothed delta to RGB and writes RGB24.
//   k_gauss     5x5 binomial alternative (--no-edge-aware).
//   k_passthru  window mode before the window is full: output = input.
//   k_synth     synthetic frames (deterministic integer generator).
#include <hip/hip_runtime.h>
#include <type_traits>

#include "synth.h"
#include "wema_jobs.h"

// ---------------------------------------------------------------------------------------
// helpers

template <typename Job>
__device__ __forceinline__ int find_job(const Job *jobs, int njobs, int b) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Pointers read from job tables are generic; casting them to the global address space lets
// the compiler emit global (not flat) memory instructions, whose completion does not wait on
// LDS traffic.
#define GLOBAL_AS __attribute__((address_space(1)))
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ GLOBAL_AS T *gp(T *p) {
    return (GLOBAL_AS T *)p;
}

// The element at a 32-bit byte offset from base: with base the same for the whole wave,
// the access takes it from scalar registers and the offset from one vector register (no
// 64-bit address arithmetic per lane).
template <typename T>
__device__ __forceinline__ GLOBAL_AS T &at_b(GLOBAL_AS T *base, unsigned byte_off) {
    return *(GLOBAL_AS T *)((GLOBAL_AS char *)base + byte_off);
}

// Workgroups are dealt to the 8 XCDs of the chip round-robin (block b runs on XCD b % 8),
// and each XCD has its own L2. Tile kernels re-read a halo around their tile, so blocks
// that run on one XCD are given consecutive tiles: neighbours then share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

// Whole-sample symmetric extension of index i into [0, n) (src/dwt.c:63-98 boundary rules).
// Tiles reach at most a few samples past an edge, so the modulo is only needed for tiny n.
__device__ __forceinline__ int reflect_idx(int i, int n) {
    if ((unsigned)i < (unsigned)n) return i;
    if (n <= 1) return 0;
    if (i < 0) i = -i;
    if (i < n) return i;
    const int j = 2 * n - 2 - i;
    if (j >= 0) return j;
    const int period = 2 * n - 2;
    i %= period;
    return i < n ? i : period - i;
}

__device__ __forceinline__ float u8f(unsigned v) { return (float)v * (1.0f / 255.0f); }

// Colour conversions with the contraction shapes of the reference's vectorised loops
// (src/wema.c:42-51, 60-78, 80-102 built with -Ofast).
__device__ __forceinline__ float luma_of(float r, float g, float b) {
    return fmaf(0.114f, b, fmaf(0.299f, r, 0.587f * g));
}
__device__ __forceinline__ float cb_of(float r, float g, float b) {
    return fmaf(-0.169f, r, 0.5f * b) + fmaf(-0.331f, g, 0.5f);
}
__device__ __forceinline__ float cr_of(float r, float g, float b) {
    return fmaf(-0.081f, b, fmaf(-0.419f, g, (r + 1.0f) * 0.5f));
}
__device__ __forceinline__ float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }
// src/ffmpeg_io.c:304-310
__device__ __forceinline__ unsigned char egress_u8(float v) {
    return (unsigned char)fmaf(clamp01(v), 255.0f, 0.5f);
}

// |v| rounded to a multiple of 2^-32, as an integer (order-independent threshold sums).
__device__ __forceinline__ unsigned long long abs_fixed(float v) {
    const unsigned u = __float_as_uint(v);
    const unsigned e = (u >> 23) & 0xFFu;
    if (e == 0) return 0ull;
    const unsigned long long m = (u & 0x7FFFFFu) | 0x800000u;
    const int sh = (int)e - 150 + 32;
    if (sh >= 0) return sh > 38 ? (1ull << 62) : (m << sh);
    if (sh < -25) return 0ull;
    return (m + (1ull << (-sh - 1))) >> (-sh);
}

// The same without branches for |v| < 2^30: the integer part and the fraction are exact,
// the fraction times 2^32 (x, below 2^32) is exact, and x rounded half up is its truncation
// plus one where its fractional part is at least 0.5. Equal to abs_fixed for every float
// below 2^30 (checked exhaustively by tests/q32_check.c); abs_fixed saturates from 2^30 on.
__device__ __forceinline__ unsigned long long abs_q32_small(float v) {
    const float a = fabsf(v);
    const unsigned hi = (unsigned)a;
    const float x = (a - (float)hi) * 4294967296.0f;
    const unsigned lo = (unsigned)x + (__builtin_amdgcn_fractf(x) >= 0.5f ? 1u : 0u);
    return ((unsigned long long)hi << 32) + lo;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// T = 0.2 * (sum over the reference's N6 positions of |c|) / N6, where that sum is twice
// the exact unique-position sum (src/wema.c:718-723).
__device__ __forceinline__ float threshold_of(unsigned long long acc, float n6f) {
    const double s = (double)acc * (1.0 / 4294967296.0);
    const float mag = (float)(2.0 * s);
    const float q = mag / n6f;
    return q * 0.2f;
}

// Value of lane i+1 / lane i-1 of the wave (DPP wave_shl:1 / wave_shr:1; the end lane gets
// 0, a value that only reaches halo lanes). With the zero fill (bound_ctrl) the compiler
// folds the shift into the VALU instruction that uses it, as that instruction's DPP source
// operand, instead of issuing a separate move.
__device__ __forceinline__ float lane_next(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_prev(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}

// One lifting pass over a row held one (even, odd) pair per lane; lanes 0-1 and 62-63 are
// halo (src/dwt.c:62-98).
__device__ __forceinline__ void lift_fwd_lanes(float &e, float &o) {
    o = fmaf(WEMA_ALPHA, e + lane_next(e), o);
    e = fmaf(WEMA_BETA, lane_prev(o) + o, e);
    o = fmaf(WEMA_GAMMA, e + lane_next(e), o);
    e = fmaf(WEMA_DELTA, lane_prev(o) + o, e);
}

// Inverse (src/dwt.c:124-161).
__device__ __forceinline__ void lift_inv_lanes(float &e, float &o) {
    e = fmaf(-WEMA_DELTA, lane_prev(o) + o, e);
    o = fmaf(-WEMA_GAMMA, e + lane_next(e), o);
    e = fmaf(-WEMA_BETA, lane_prev(o) + o, e);
    o = fmaf(-WEMA_ALPHA, e + lane_next(e), o);
}

// Channel value of an RGB24 / float RGB pixel.
template <int KIND>
__device__ __forceinline__ float chan_of(float r, float g, float b) {
    if (KIND == SRC_U8_Y || KIND == SRC_F32_Y) return luma_of(r, g, b);
    if (KIND == SRC_U8_CB || KIND == SRC_F32_CB) return cb_of(r, g, b);
    return cr_of(r, g, b);
}

// ---------------------------------------------------------------------------------------
// Forward DWT level as strips (src/dwt.c:369-473 + 645-703).
//
// A wave computes output pairs [px0, px0 + 60) x [py0, py0 + sp) of one job, for all
// of the job's channels (Y, Cb and Cr of an RGB frame are decoded from one load). Lane l
// owns input pair column px0 - 2 + l, so the row lifting runs across lanes (DPP shifts);
// the 2-pair border on each side feeds the 4 lifting steps, each of which widens the
// dependency by one pair. The wave streams input pair rows py0 - 2 .. py0 + sp + 1 from
// top to bottom, and the column lifting runs as a pipeline in registers: when input pair m
// arrives, steps 1 and 2 of pair m - 1 and steps 3 and 4 of pair m - 2 are computed, which
// makes output pair m - 2 final. Each step is the same fused multiply-add of the same
// operands as the reference's lifting (src/dwt.c:45-107), so the coefficients are
// bit-identical; no LDS and no barrier are needed, and the loads of the next input pairs
// are in flight while a pair is computed.
//
// Samples outside the image are taken from the whole-sample symmetric extension, which
// reproduces the reference's per-step boundary rules bit for bit (the extended values are
// equal sums of the same operands). Where the parent size is odd, the reference's
// extraction leaves the last HL/HH column or LH/HH row of the subband unwritten and its
// scratch buffer keeps the previous level's value at the same flat index
// (src/dwt.c:459-472, 648-653); those entries are copied from the previous level's plane
// (0 at level 0) and enter the threshold sum.

// Source families of a strip job.
enum { FAM_U8_RGB = 0, FAM_F32_RGB = 1, FAM_F32_PLANE = 2 };

// One input row's data for a lane's sample pair (xa, xb).
template <int FAM> struct StripRaw;
template <> struct StripRaw<FAM_U8_RGB> { unsigned d0, d1, d2; };   // 6 bytes from byte L.sh of d0 (0: d0, d1)
template <> struct StripRaw<FAM_F32_RGB> { float ra, ga, ba, rb, gb, bb; };
template <> struct StripRaw<FAM_F32_PLANE> { float a, b; };

// Input pairs a strip keeps in flight.
#ifndef WEMA_FSLOTS_COLOUR
#define WEMA_FSLOTS_COLOUR 2
#endif
#define SLOTS_OF(NCH) ((NCH) == 1 ? 4 : WEMA_FSLOTS_COLOUR)
// Pair columns of a workgroup's staged output rows (level-0 DWT strips).
#define WEMA_STAGE_W (WEMA_STRIPS_PER_BLOCK * WEMA_TILE_PAIRS_X)
#ifndef WEMA_FWD_STAGE
#define WEMA_FWD_STAGE 1
#endif
// ... in launches of at least this many workgroups (8 per CU): with fewer, the barrier per
// row leaves the SIMDs idle (one frame per call: 1080p colour 133 -> 136 us, 4K 207 -> 210 us)
#ifndef WEMA_INV_STAGE
#define WEMA_INV_STAGE 1
#endif
#ifndef WEMA_STAGE_MIN_BLOCKS
#define WEMA_STAGE_MIN_BLOCKS 2048
#endif

// Per-lane constants of a strip.
struct StripLane {
    int xa, xb;        // reflected sample columns of the lane's pair
    unsigned boff;     // fast path: byte offset of the pair in its row (RGB24: aligned down)
    unsigned sh;       // fast path, RGB24: byte offset of the pair in its first dword
};

// Loads row y. FAST: every lane's pair is two adjacent samples and its wide load stays
// inside the source (uniform over the strip); the row's offset is then a 32-bit scalar.
template <int FAM, bool FAST>
__device__ __forceinline__ StripRaw<FAM> strip_load(const DwtfJob &J, const StripLane &L, int y) {
    StripRaw<FAM> r;
    const int w = J.w;
    if (FAM == FAM_U8_RGB) {
        unsigned *d = (unsigned *)&r;
        if (FAST) {
            const GLOBAL_AS unsigned char *row = gp((const unsigned char *)J.src) + (unsigned)y * (unsigned)(3 * w);
            const GLOBAL_AS unsigned *q = (const GLOBAL_AS unsigned *)(row + L.boff);
            d[0] = q[0];
            d[1] = q[1];
            d[2] = q[2];
        } else {
            // reflected columns, rows that need not start on a dword, the end of the frame:
            // the 6 bytes one by one, the same loads in every lane (no branch, so the loads
            // of later rows stay in flight while this one is computed)
            const GLOBAL_AS unsigned char *row = gp((const unsigned char *)J.src) + (size_t)y * w * 3;
            const GLOBAL_AS unsigned char *pa = row + (size_t)L.xa * 3, *pb = row + (size_t)L.xb * 3;
            d[0] = (unsigned)pa[0] | ((unsigned)pa[1] << 8) | ((unsigned)pa[2] << 16) | ((unsigned)pb[0] << 24);
            d[1] = (unsigned)pb[1] | ((unsigned)pb[2] << 8);
            d[2] = 0u;
        }
    } else if (FAM == FAM_F32_PLANE) {
        float *d = (float *)&r;
        if (FAST) {
            const GLOBAL_AS float *row = gp((const float *)J.src) + (unsigned)y * (unsigned)w;
            const GLOBAL_AS float *q = row + L.xa;
            d[0] = q[0];
            d[1] = q[1];
        } else {
            const GLOBAL_AS float *row = gp((const float *)J.src) + (size_t)y * w;
            d[0] = row[L.xa];
            d[1] = row[L.xb];
        }
    } else {
        const GLOBAL_AS float *row = gp((const float *)J.src) + (size_t)y * w * 3;
        const GLOBAL_AS float *pa = row + (size_t)L.xa * 3, *pb = row + (size_t)L.xb * 3;
        float *d = (float *)&r;
        d[0] = pa[0];
        d[1] = pa[1];
        d[2] = pa[2];
        d[3] = pb[0];
        d[4] = pb[1];
        d[5] = pb[2];
    }
    return r;
}

// Float RGB of both samples of a row (RGB families); sh: byte offset of an RGB24 pair in d0.
template <int FAM>
__device__ __forceinline__ void strip_rgb(const StripRaw<FAM> &r, unsigned sh, float (&a)[3], float (&b)[3]) {
    if constexpr (FAM == FAM_U8_RGB) {
        const unsigned v0 = __builtin_amdgcn_alignbyte(r.d1, r.d0, sh);
        const unsigned v1 = __builtin_amdgcn_alignbyte(r.d2, r.d1, sh);
        a[0] = u8f(v0 & 0xFFu);
        a[1] = u8f((v0 >> 8) & 0xFFu);
        a[2] = u8f((v0 >> 16) & 0xFFu);
        b[0] = u8f(v0 >> 24);
        b[1] = u8f(v1 & 0xFFu);
        b[2] = u8f((v1 >> 8) & 0xFFu);
    } else if constexpr (FAM == FAM_F32_RGB) {
        a[0] = r.ra;
        a[1] = r.ga;
        a[2] = r.ba;
        b[0] = r.rb;
        b[1] = r.gb;
        b[2] = r.bb;
    }
}

// Channel c (0 = Y, 1 = Cb, 2 = Cr) of a float RGB sample.
__device__ __forceinline__ float chan_idx(int c, const float (&p)[3]) {
    return c == 0 ? chan_of<SRC_U8_Y>(p[0], p[1], p[2])
                  : c == 1 ? chan_of<SRC_U8_CB>(p[0], p[1], p[2]) : chan_of<SRC_U8_CR>(p[0], p[1], p[2]);
}

// Column-lifting pipeline of one lane column: raw e, o of the last pair, o after step 1
// and e after step 2 of the pair before, o after step 3 of the pair before that.
struct VPipe {
    float e, o, o1, e2, o3;
};

// Input pair m arrives (e, o); returns e after step 4 and o after step 3 of pair m - 2.
__device__ __forceinline__ void vpipe_step(VPipe &p, float e, float o, float &e4, float &o3) {
    const float o1 = fmaf(WEMA_ALPHA, p.e + e, p.o);       // o1[m-1]
    const float e2 = fmaf(WEMA_BETA, p.o1 + o1, p.e);      // e2[m-1]
    o3 = fmaf(WEMA_GAMMA, p.e2 + e2, p.o1);                 // o3[m-2]
    e4 = fmaf(WEMA_DELTA, p.o3 + o3, p.e2);                 // e4[m-2]
    p.e = e;
    p.o = o;
    p.o1 = o1;
    p.e2 = e2;
    p.o3 = o3;
}

// SL: input pairs in flight (the slot count of NCH; 2 for luma strips of 4 pair rows, whose
// fast path needs 2 SL rows)
template <int FAM, int NCH, bool FAST, bool STALE, bool STAGE, int SL = SLOTS_OF(NCH)>
__device__ __forceinline__ void dwt_fwd_strip_t(const DwtfJob &J, int strip, int lane, float *stg) {
    const int SP = J.sp;
    const int w = J.w, h = J.h;
    const int sw = (w + 1) >> 1, sh = (h + 1) >> 1;
    const int sx = strip % J.tiles_x, sy = strip / J.tiles_x;
    const int px0 = sx * WEMA_TILE_PAIRS_X, py0 = sy * SP;
    const int sx0 = 2 * (px0 - 2);

    StripLane L;
    L.xa = reflect_idx(sx0 + 2 * lane, w);
    L.xb = reflect_idx(sx0 + 2 * lane + 1, w);
    if (FAM == FAM_U8_RGB) {   // (FAST strips: rows start on a dword)
        const unsigned b0 = 3u * (unsigned)L.xa;
        L.sh = b0 & 3u;
        L.boff = b0 - L.sh;
    } else {
        L.sh = 0u;
        L.boff = 0u;
    }

    const int hw_valid = w >> 1, hh_valid = h >> 1;   // written extents (src/dwt.c:463-469)
    const size_t n = (size_t)sw * sh;
    const size_t pn = (size_t)w * h;
    const int px = px0 + lane - 2;
    const bool col_ok = lane >= 2 && lane < 62 && px < sw;
    const bool col_stale = px >= hw_valid;
    const GLOBAL_AS float *prev = gp(J.prev);

    VPipe pl[NCH], ph[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        pl[c] = VPipe{0.f, 0.f, 0.f, 0.f, 0.f};
        ph[c] = VPipe{0.f, 0.f, 0.f, 0.f, 0.f};
    }
    unsigned long long part[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) part[c] = 0ull;

    // Staged output rows (stg: the workgroup's LDS, its 4 strips side by side in one strip
    // row): each wave puts its 60 pairs of a row into LDS, and after a barrier the workgroup
    // stores the 240 pairs as contiguous rows, so that most 128-byte lines are written whole
    // by one store instruction instead of in parts by two waves. Two buffers alternate, so a
    // row is written while the one before it is still being read.
    int sbuf = 0;
    const int pxb = (sx - (int)(threadIdx.x >> 6)) * WEMA_TILE_PAIRS_X;   // the workgroup's first pair column
    auto flush = [&](int py) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        const int t = (int)threadIdx.x, gx = pxb + t;
        if (t < WEMA_STAGE_W && gx < sw) {
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const float *b = stg + (sbuf * NCH + c) * (4 * WEMA_STAGE_W) + t;
                GLOBAL_AS float *LLc = gp(J.ll) + c * J.ll_cstride + (size_t)py * sw;
                GLOBAL_AS float *LHc = gp(J.planes) + c * J.planes_cstride + (size_t)py * sw;
                const unsigned xb = 4u * (unsigned)gx;
                at_b(LLc, xb) = b[0];
                at_b(LHc, xb) = b[WEMA_STAGE_W];
                at_b(LHc + n, xb) = b[2 * WEMA_STAGE_W];
                at_b(LHc + 2 * n, xb) = b[3 * WEMA_STAGE_W];
            }
        }
        sbuf ^= 1;
    };

    // output pair row py (uniform) of channel c from its four final values
    auto emit = [&](int c, int py, float ll, float lh, float hl, float hh) {
        // level 0 has no previous level (stale entries are 0); at the coarser levels only
        // strips that reach past the written extents read the previous level's planes, and
        // only lanes that write an output (px in [0, sw)) read them
        if ((FAM != FAM_F32_PLANE || STALE) && col_ok) {
            if (py >= hh_valid) {
                lh = FAM == FAM_F32_PLANE ? prev[(size_t)py * sw + px] : 0.0f;
                hh = FAM == FAM_F32_PLANE ? prev[pn + (size_t)py * sw + px] : 0.0f;
            }
            if (col_stale) {
                hl = FAM == FAM_F32_PLANE ? prev[2 * pn + (size_t)py * sw + px] : 0.0f;
                hh = FAM == FAM_F32_PLANE ? prev[pn + (size_t)py * sw + px] : 0.0f;
            }
        }
        if (STAGE) {
            if (lane >= 2 && lane < 62) {
                float *b = stg + (sbuf * NCH + c) * (4 * WEMA_STAGE_W) + (threadIdx.x >> 6) * WEMA_TILE_PAIRS_X + lane - 2;
                b[0] = ll;
                b[WEMA_STAGE_W] = lh;
                b[2 * WEMA_STAGE_W] = hh;
                b[3 * WEMA_STAGE_W] = hl;
            }
        } else {
            GLOBAL_AS float *LLc = gp(J.ll) + c * J.ll_cstride + (size_t)py * sw;
            GLOBAL_AS float *LHc = gp(J.planes) + c * J.planes_cstride + (size_t)py * sw;
            if (col_ok) {
                const unsigned xb = 4u * (unsigned)px;
                at_b(LLc, xb) = ll;
                at_b(LHc, xb) = lh;
                at_b(LHc + n, xb) = hh;
                at_b(LHc + 2 * n, xb) = hl;
            }
        }
        bool small = true;
        if (FAM != FAM_U8_RGB)   // (RGB24 samples give coefficients below 64)
            small = __all(!col_ok || fabsf(lh) + fabsf(hl) + fabsf(hh) < 1073741824.0f);
        if (col_ok) {
            if (small) part[c] += abs_q32_small(lh) + abs_q32_small(hl) + abs_q32_small(hh);
            else part[c] += abs_fixed(lh) + abs_fixed(hl) + abs_fixed(hh);
        }
    };

    // one input pair (rows ra, rb) through the row lifting and the column pipelines
    auto step = [&](int m, bool out, const StripRaw<FAM> &ra, const StripRaw<FAM> &rb) {
        float a0[3], b0[3], a1[3], b1[3];
        if constexpr (FAM != FAM_F32_PLANE) {
            strip_rgb<FAM>(ra, FAST ? L.sh : 0u, a0, b0);
            strip_rgb<FAM>(rb, FAST ? L.sh : 0u, a1, b1);
        }
        const int py = py0 + m - 4;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            float e0, o0, e1, o1;
            if constexpr (FAM == FAM_F32_PLANE) {
                const float *fa = (const float *)&ra, *fb = (const float *)&rb;
                e0 = fa[0];
                o0 = fa[1];
                e1 = fb[0];
                o1 = fb[1];
            } else {
                e0 = chan_idx(c, a0);
                o0 = chan_idx(c, b0);
                e1 = chan_idx(c, a1);
                o1 = chan_idx(c, b1);
            }
            float l0, g0, l1, g1;   // row-lifted L (low) and H (high) samples of both rows
            if (FAST || w >= 2) {
                lift_fwd_lanes(e0, o0);
                lift_fwd_lanes(e1, o1);
                l0 = e0 * WEMA_KINV;
                g0 = o0 * WEMA_K;
                l1 = e1 * WEMA_KINV;
                g1 = o1 * WEMA_K;
            } else {
                // n < 2: the row transform is the identity (src/dwt.c:46)
                l0 = g0 = e0;
                l1 = g1 = e1;
            }
            float le4, lo3, he4, ho3;
            vpipe_step(pl[c], l0, l1, le4, lo3);
            vpipe_step(ph[c], g0, g1, he4, ho3);
            if (out) emit(c, py, le4 * WEMA_KINV, lo3 * WEMA_K, he4 * WEMA_KINV, ho3 * WEMA_K);
        }
        if (out && STAGE) flush(py);
    };

    if (h < 2) {
        // one output pair row: the column transform is the identity (src/dwt.c:46), and
        // both rows of the pair are row 0
        const StripRaw<FAM> r = strip_load<FAM, FAST>(J, L, 0);
        float a0[3], b0[3];
        if constexpr (FAM != FAM_F32_PLANE) strip_rgb<FAM>(r, FAST ? L.sh : 0u, a0, b0);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            float e0, o0;
            if constexpr (FAM == FAM_F32_PLANE) {
                e0 = ((const float *)&r)[0];
                o0 = ((const float *)&r)[1];
            } else {
                e0 = chan_idx(c, a0);
                o0 = chan_idx(c, b0);
            }
            float l0 = e0, g0 = e0;
            if (w >= 2) {
                lift_fwd_lanes(e0, o0);
                l0 = e0 * WEMA_KINV;
                g0 = o0 * WEMA_K;
            }
            emit(c, 0, l0, l0, g0, g0);
        }
        if (STAGE) flush(0);
    } else {
        // input pairs py0 - 2 + m for m < M; output pair rows py0 .. py0 + nout - 1
        constexpr int SLOTS = SL;
        const int nout = min(SP, sh - py0);
        // FAST strips have nout >= 2 * SLOTS; shorter strips run on to 2 * SLOTS steps past
        // the warm-up, the extra steps reading reflected rows and writing nothing
        const int M = (FAST ? nout : max(nout, 2 * SLOTS)) + 4;
        auto load = [&](int m, int r) {
            int y = 2 * (py0 - 2 + m) + r;
            // (FAST strips: h >= 8, so one reflection reaches every row a strip reads)
            y = FAST ? (y < 0 ? -y : y >= h ? 2 * h - 2 - y : y) : reflect_idx(y, h);
            return strip_load<FAM, FAST>(J, L, y);
        };
        // SLOTS input pairs in flight: a slot is reloaded right after its pair is used, so
        // its registers are rewritten in place while the other slots' pairs are computed.
        // Pairs 0-3 only fill the column pipelines. Every strip has M >= 4 + 2 * SLOTS, so
        // the loads of the warm-up and of the main loop need no conditions (which keeps them
        // in flight across iterations: a load issued under a condition makes the wait for it
        // a wait for every load); the tail takes the rest.
        auto reload = [&](StripRaw<FAM> &a, StripRaw<FAM> &b, int m) {
            a = load(m, 0);
            b = load(m, 1);
        };
        auto emits = [&](int m) { return FAST || m - 4 < nout; };
        StripRaw<FAM> sa[SLOTS], sb[SLOTS];
        // (after the 4 warm-up pairs, pair m + k of the loops below is in slot k)
        static_assert(4 % SLOTS == 0, "the slot count must divide the 4 warm-up pairs");
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) reload(sa[k], sb[k], k);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            step(m, false, sa[m % SLOTS], sb[m % SLOTS]);
            reload(sa[m % SLOTS], sb[m % SLOTS], m + SLOTS);
        }
        int m = 4;
        for (; m + 2 * SLOTS <= M; m += SLOTS) {
#pragma unroll
            for (int k = 0; k < SLOTS; ++k) {
                step(m + k, emits(m + k), sa[k], sb[k]);
                reload(sa[k], sb[k], m + k + SLOTS);
            }
        }
        for (; m < M; m += SLOTS) {
#pragma unroll
            for (int k = 0; k < SLOTS; ++k) {
                if (m + k < M) {
                    step(m + k, emits(m + k), sa[k], sb[k]);
                    if (m + k + SLOTS < M) reload(sa[k], sb[k], m + k + SLOTS);
                }
            }
        }
    }
    if (J.acc) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const unsigned long long s = wave_sum_u64(part[c]);
            // (also when s is 0: with a branch on s here the compiler waits for every
            // load in flight at the top of each iteration of the strip loop above)
            if (lane == 0) atomicAdd(J.acc + c * J.acc_slots + (strip & (J.acc_slots - 1)), s);
        }
    }
}

template <int FAM, int NCH, bool STAGE>
__device__ __forceinline__ void dwt_fwd_strip(const DwtfJob &J, int strip, int lane, float *stg) {
    const int w = J.w, h = J.h, sh = (h + 1) >> 1;
    const int sx0 = 2 * ((strip % J.tiles_x) * WEMA_TILE_PAIRS_X - 2);
    const int py0 = (strip / J.tiles_x) * J.sp;
    if constexpr (FAM == FAM_F32_RGB) {
        // (float RGB frames: the general loads only)
        dwt_fwd_strip_t<FAM, NCH, false, false, STAGE>(J, strip, lane, stg);
    } else if constexpr (FAM == FAM_F32_PLANE) {
        // every lane's pair adjacent and inside the row, sources below 2^29 samples (32-bit
        // row offsets); strips that reach the unwritten last column or row of an odd-sized
        // parent take those entries from the previous level
        const int px0 = (strip % J.tiles_x) * WEMA_TILE_PAIRS_X;
        const bool stale = px0 + WEMA_TILE_PAIRS_X > (w >> 1) || py0 + J.sp > (h >> 1);
        const bool fast = sx0 >= 0 && sx0 + 128 <= w && h >= 8 && (long long)w * h < (1ll << 29) &&
                          min(J.sp, sh - py0) >= (NCH == 1 && J.sp == 4 ? 4 : 2 * SLOTS_OF(NCH));
        if (stale) dwt_fwd_strip_t<FAM, NCH, false, true, STAGE>(J, strip, lane, stg);
        else if (fast && NCH == 1 && J.sp == 4) dwt_fwd_strip_t<FAM, NCH, true, false, STAGE, 2>(J, strip, lane, stg);
        else if (fast) dwt_fwd_strip_t<FAM, NCH, true, false, STAGE>(J, strip, lane, stg);
        else dwt_fwd_strip_t<FAM, NCH, false, false, STAGE>(J, strip, lane, stg);
    } else {
        // RGB24 as above, and rows that start on a dword (frame aligned, width a multiple
        // of 4); the 3-dword loads end at most 6 bytes past the pair, so the strip must not
        // reach the frame's last bytes
        const int last_row = 2 * (min(J.sp, sh - py0) + py0 + 1) + 1;   // last input row (unreflected)
        const bool fast = sx0 >= 0 && sx0 + 128 <= w && h >= 8 && (long long)w * h < (1ll << 29) &&
                          min(J.sp, sh - py0) >= (NCH == 1 && J.sp == 4 ? 4 : 2 * SLOTS_OF(NCH)) &&
                          (((size_t)J.src | (size_t)w) & 3) == 0 && (sx0 + 132 <= w || last_row < h - 1);
        if (fast && NCH == 1 && J.sp == 4) dwt_fwd_strip_t<FAM, NCH, true, false, STAGE, 2>(J, strip, lane, stg);
        else if (fast) dwt_fwd_strip_t<FAM, NCH, true, false, STAGE>(J, strip, lane, stg);
        else dwt_fwd_strip_t<FAM, NCH, false, false, STAGE>(J, strip, lane, stg);
    }
}

// Block = WEMA_STRIPS_PER_BLOCK waves, each an independent strip of the block's job.
template <bool LEVEL0>
__device__ __forceinline__ void dwt_fwd_strips(const DwtfJob *__restrict__ jobs, int njobs) {
    const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int jid = find_job(jobs, njobs, bid);
    const DwtfJob J = jobs[jid];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int strip = (bid - J.tile_begin) * WEMA_STRIPS_PER_BLOCK + wave;
    if (strip >= J.tiles_x * wema_ceil_div((J.h + 1) >> 1, J.sp)) return;
    // (level-0 jobs take Y, or Y, Cb and Cr, from RGB24 or float RGB frames)
    // (level 0: staged rows where the job's strip rows are whole workgroups)
    __shared__ float s_stg[LEVEL0 ? 2 * 3 * 4 * WEMA_STAGE_W : 1];
    const bool stage = LEVEL0 && WEMA_FWD_STAGE && (J.tiles_x & (WEMA_STRIPS_PER_BLOCK - 1)) == 0 &&
                       gridDim.x >= WEMA_STAGE_MIN_BLOCKS;
    if (!LEVEL0) dwt_fwd_strip<FAM_F32_PLANE, 1, false>(J, strip, lane, s_stg);
    else if (J.src_kind == SRC_U8_Y && J.nch == 3) {
        if (stage) dwt_fwd_strip<FAM_U8_RGB, 3, true>(J, strip, lane, s_stg);
        else dwt_fwd_strip<FAM_U8_RGB, 3, false>(J, strip, lane, s_stg);
    } else if (J.src_kind == SRC_U8_Y) {
        if (stage) dwt_fwd_strip<FAM_U8_RGB, 1, true>(J, strip, lane, s_stg);
        else dwt_fwd_strip<FAM_U8_RGB, 1, false>(J, strip, lane, s_stg);
    } else if (J.nch == 3) {
        if (stage) dwt_fwd_strip<FAM_F32_RGB, 3, true>(J, strip, lane, s_stg);
        else dwt_fwd_strip<FAM_F32_RGB, 3, false>(J, strip, lane, s_stg);
    } else {
        if (stage) dwt_fwd_strip<FAM_F32_RGB, 1, true>(J, strip, lane, s_stg);
        else dwt_fwd_strip<FAM_F32_RGB, 1, false>(J, strip, lane, s_stg);
    }
}

__global__ void __launch_bounds__(64 * WEMA_STRIPS_PER_BLOCK) k_dwt_fwd0(const DwtfJob *__restrict__ jobs, int njobs) {
    dwt_fwd_strips<true>(jobs, njobs);
}
__global__ void __launch_bounds__(64 * WEMA_STRIPS_PER_BLOCK) k_dwt_fwd(const DwtfJob *__restrict__ jobs, int njobs) {
    dwt_fwd_strips<false>(jobs, njobs);
}

// Colour mode output (src/wema.c:80-102).
__device__ __forceinline__ void ycc_px(float yv, float cb, float cr, float (&v)[3]) {
    const float cbv = cb - 0.5f, crv = cr - 0.5f;
    v[0] = clamp01(fmaf(1.402f, crv, yv));
    v[1] = clamp01(yv - fmaf(0.344f, cbv, 0.714f * crv));
    v[2] = clamp01(fmaf(1.772f, cbv, yv));
}

// ---------------------------------------------------------------------------------------
// Inverse DWT level as strips (src/dwt.c:475-582 + 705-767).
//
// A wave reconstructs image columns 2 px0 .. 2 px0 + 119 and rows 2 py0 .. 2 (py0 + sp) - 1
// of one job (all channels of a colour job). Lane l owns coefficient column
// I = px0 - 2 + l of the packed [LL | HL ; LH | HH] layout: its low column (LL over LH) and
// its high column (HL over HH). The wave streams coefficient pair rows py0 - 2 ..
// py0 + sp + 1 from top to bottom through two column pipelines per channel: when pair m
// arrives, step 1 of pair m, steps 2 and 3 of pair m - 1 and step 4 of pair m - 2 are
// computed (each 1D inverse first undoes the K scaling, then runs the reference's four
// inverse lifting steps with the same operands, src/dwt.c:124-161, so the samples are
// bit-identical), which makes image rows 2 (m - 2) and 2 (m - 2) + 1 final. Each such row
// then goes through the row synthesis across lanes (DPP shifts) and is stored. Coefficients
// outside the subbands come from the whole-sample symmetric extension of the interleaved
// signal (as in the forward DWT); every load is in bounds.

// Column-synthesis pipeline of one lane column: o (K-scaled) of the last pair, e after
// step 1 of the last pair, o after step 2 and e after step 3 of the pair before.
struct IPipe {
    float o, e1, o2, e3;
};

// Pair m arrives (e, o scaled); returns the even-row sample (e after step 3) and the
// odd-row sample (o after step 4) of pair m - 2.
__device__ __forceinline__ void ipipe_step(IPipe &p, float e, float o, float &ev, float &od) {
    const float e1 = fmaf(-WEMA_DELTA, p.o + o, e);        // e1[m]
    const float o2 = fmaf(-WEMA_GAMMA, p.e1 + e1, p.o);    // o2[m-1]
    const float e3 = fmaf(-WEMA_BETA, p.o2 + o2, p.e1);    // e3[m-1]
    od = fmaf(-WEMA_ALPHA, p.e3 + e3, p.o2);                // o4[m-2]
    ev = p.e3;                                              // e3[m-2]
    p.o = o;
    p.e1 = e1;
    p.o2 = o2;
    p.e3 = e3;
}

// One coefficient pair row of every channel for a lane: LL, LH (low column) and HL, HH
// (high column).
template <int NCH>
struct IStripRaw {
    float ll[NCH], lh[NCH], hl[NCH], hh[NCH];
};

#ifndef WEMA_ISLOTS_COLOUR
#define WEMA_ISLOTS_COLOUR 2
#endif
#define ISLOTS_OF(NCH) ((NCH) == 1 ? 4 : WEMA_ISLOTS_COLOUR)

// OUT: 0 = one float plane (J.out), 1 = colour egress (RGB24 and/or float RGB). SL: pairs in
// flight (the slot count of NCH; 2 for luma strips of 4 pair rows).
template <int NCH, int OUT, bool FAST, bool STAGE, int SL = ISLOTS_OF(NCH)>
__device__ __forceinline__ void dwt_inv_strip_t(const DwtiJob &J, int strip, int lane, float *stg) {
    const int SP = J.sp;
    const int w = J.w, h = J.h;
    const int sw = (w + 1) >> 1, sh = (h + 1) >> 1;
    const size_t n = (size_t)sw * sh;
    const int px0 = (strip % J.tiles_x) * WEMA_TILE_PAIRS_X, py0 = (strip / J.tiles_x) * SP;
    const int I = px0 - 2 + lane;
    // coefficient columns of the lane's low and high columns (FAST: I itself)
    const unsigned colL = FAST ? (unsigned)I : (unsigned)(reflect_idx(2 * I, w) >> 1);
    const unsigned colH = FAST ? (unsigned)I : (unsigned)(reflect_idx(2 * I + 1, w) >> 1);
    const int x = 2 * I;
    const bool lane_ok = lane >= 2 && lane < 62 && x < w;
    const bool two = x + 1 < w;
    const bool vert = h >= 2, horz = w >= 2, even_w = (w & 1) == 0;

    IPipe pl[NCH], ph[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        pl[c] = IPipe{0.f, 0.f, 0.f, 0.f};
        ph[c] = IPipe{0.f, 0.f, 0.f, 0.f};
    }

    // coefficient pair row q (unextended) of every channel
    auto load = [&](int q) {
        IStripRaw<NCH> r;
        int ye2 = 2 * q, yo2 = 2 * q + 1;
        // (FAST strips: h >= 8, so one reflection reaches every row a strip reads)
        ye2 = FAST ? (ye2 < 0 ? -ye2 : ye2 >= h ? 2 * h - 2 - ye2 : ye2) : reflect_idx(ye2, h);
        yo2 = FAST ? (yo2 < 0 ? -yo2 : yo2 >= h ? 2 * h - 2 - yo2 : yo2) : reflect_idx(yo2, h);
        const size_t re = (size_t)(ye2 >> 1) * sw, ro = (size_t)(yo2 >> 1) * sw;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const GLOBAL_AS float *LL = gp(J.ll) + c * J.ll_cstride;
            const GLOBAL_AS float *LH = gp(J.planes) + c * J.planes_cstride;
            r.ll[c] = at_b(LL + re, 4u * colL);
            r.lh[c] = at_b(LH + ro, 4u * colL);
            r.hl[c] = at_b(LH + 2 * n + re, 4u * colH);
            r.hh[c] = at_b(LH + n + ro, 4u * colH);
        }
        return r;
    };

    // Staged output rows (level 0 of large launches, the workgroup's 4 strips side by side
    // in one strip row): the waves put the pixel pairs of a step's two image rows into LDS
    // (float planes: 2 floats per pair; RGB24: 6 bytes, as an aligned dword and short), and
    // after a barrier the workgroup stores its 480 pixels of each row as contiguous rows.
    // Two buffers alternate.
    int sbuf = 0;
    const int pxb = (px0 / WEMA_TILE_PAIRS_X - (int)(threadIdx.x >> 6)) * WEMA_TILE_PAIRS_X;
    auto flush = [&](int y0, int nrows) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        const int t = (int)threadIdx.x;
        if constexpr (OUT == 0) {
            const int gx = pxb + t;
            if (t < WEMA_STAGE_W && 2 * gx < w) {
                for (int r = 0; r < nrows; ++r) {
                    if (y0 + r >= h) break;
                    const v2f v = ((const v2f *)stg)[(sbuf * 2 + r) * WEMA_STAGE_W + t];
                    at_b((GLOBAL_AS v2f *)(gp(J.out) + (size_t)(y0 + r) * w), 8u * (unsigned)gx) = v;
                }
            }
        } else {
            const int nb = 6 * min(WEMA_STAGE_W, sw - pxb);   // bytes of the workgroup's part of a row
            const unsigned *sd = (const unsigned *)stg;
            for (int r = 0; r < nrows; ++r) {
                if (y0 + r >= h) break;
                GLOBAL_AS unsigned char *rb = gp(J.out_u8) + ((size_t)(y0 + r) * w + 2 * pxb) * 3;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int d = t + 256 * k;
                    const unsigned v = sd[(sbuf * 2 + r) * (6 * WEMA_STAGE_W / 4) + min(d, 6 * WEMA_STAGE_W / 4 - 1)];
                    if (4 * d + 4 <= nb) at_b((GLOBAL_AS unsigned *)rb, 4u * (unsigned)d) = v;
                    else if (4 * d < nb) at_b((GLOBAL_AS unsigned short *)rb, 4u * (unsigned)d) = (unsigned short)v;
                }
            }
        }
        sbuf ^= 1;
    };

    // image row y (uniform) of every channel: row synthesis and store
    auto put_row = [&](int y, const float (&lo)[NCH], const float (&hi)[NCH]) {
        float ev[NCH], ov[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            ev[c] = lo[c];
            ov[c] = hi[c];
            if (FAST || horz) {
                ev[c] *= WEMA_K;
                ov[c] *= WEMA_KINV;
                lift_inv_lanes(ev[c], ov[c]);
            }
        }
        if constexpr (STAGE) {
            if (lane < 2 || lane >= 62) return;
            const int k = (int)(threadIdx.x >> 6) * WEMA_TILE_PAIRS_X + lane - 2;   // pair in the workgroup's row
            const int r = y & 1;
            if constexpr (OUT == 0) {
                ((v2f *)stg)[(sbuf * 2 + r) * WEMA_STAGE_W + k] = v2f{ev[0], ov[0]};
            } else {
                float v0[3], v1[3];
                ycc_px(ev[0], ev[1], ev[2], v0);
                ycc_px(ov[0], ov[1], ov[2], v1);
                const unsigned lo = (unsigned)egress_u8(v0[0]) | ((unsigned)egress_u8(v0[1]) << 8) |
                                    ((unsigned)egress_u8(v0[2]) << 16) | ((unsigned)egress_u8(v1[0]) << 24);
                const unsigned hi = (unsigned)egress_u8(v1[1]) | ((unsigned)egress_u8(v1[2]) << 8);
                unsigned char *sb = (unsigned char *)stg + (sbuf * 2 + r) * (6 * WEMA_STAGE_W) + 6 * k;
                if ((k & 1) == 0) {   // (6 k on a dword)
                    *(unsigned *)sb = lo;
                    *(unsigned short *)(sb + 4) = (unsigned short)hi;
                } else {
                    *(unsigned short *)sb = (unsigned short)lo;
                    *(unsigned *)(sb + 2) = (lo >> 16) | (hi << 16);
                }
            }
            return;
        }
        if (!lane_ok || y >= h) return;
        if constexpr (OUT == 0) {
            GLOBAL_AS float *row = gp(J.out) + (size_t)y * w;
            const unsigned xb = 4u * (unsigned)x;
            if (even_w && two) {
                // (even width: the pair starts on 8 bytes)
                at_b((GLOBAL_AS v2f *)row, xb) = v2f{ev[0], ov[0]};
            } else {
                at_b(row, xb) = ev[0];
                if (two) at_b(row, xb + 4u) = ov[0];
            }
        } else {
            float v0[3], v1[3];
            ycc_px(ev[0], ev[1], ev[2], v0);
            ycc_px(ov[0], ov[1], ov[2], v1);
            const size_t idx = ((size_t)y * w + x) * 3;
            if (J.out_u8) {
                GLOBAL_AS unsigned char *o = gp(J.out_u8) + idx;
                const unsigned b0 = egress_u8(v0[0]), b1 = egress_u8(v0[1]), b2 = egress_u8(v0[2]);
                // row base (the same for the wave) and the pair's byte offset in the row
                GLOBAL_AS unsigned char *rb = gp(J.out_u8) + (size_t)y * w * 3;
                const unsigned xo = 3u * (unsigned)x, a = (unsigned)(size_t)rb + xo;
                if (two && (a & 1) == 0) {
                    // the 6 bytes as one dword and one short, both aligned: bytes 0-3 and
                    // 4-5 when the pair starts on a dword, else bytes 0-1 and 2-5
                    const unsigned lo = b0 | (b1 << 8) | (b2 << 16) | ((unsigned)egress_u8(v1[0]) << 24);
                    const unsigned hi = (unsigned)egress_u8(v1[1]) | ((unsigned)egress_u8(v1[2]) << 8);
                    const bool d0 = (a & 2) == 0;
                    const unsigned dw = d0 ? lo : (lo >> 16) | (hi << 16);
                    const unsigned short sv = (unsigned short)(d0 ? hi : lo);
                    at_b((GLOBAL_AS unsigned *)rb, xo + (d0 ? 0u : 2u)) = dw;
                    at_b((GLOBAL_AS unsigned short *)rb, xo + (d0 ? 4u : 0u)) = sv;
                } else {
                    o[0] = b0;
                    o[1] = b1;
                    o[2] = b2;
                    if (two) {
                        o[3] = egress_u8(v1[0]);
                        o[4] = egress_u8(v1[1]);
                        o[5] = egress_u8(v1[2]);
                    }
                }
            }
            if (J.out_f32) {
                GLOBAL_AS float *o = gp(J.out_f32) + idx;
                o[0] = v0[0];
                o[1] = v0[1];
                o[2] = v0[2];
                if (two) {
                    o[3] = v1[0];
                    o[4] = v1[1];
                    o[5] = v1[2];
                }
            }
        }
    };

    // one coefficient pair row through the column pipelines; out: store pair m - 2
    auto step = [&](int m, bool out, const IStripRaw<NCH> &r) {
        float le[NCH], lo[NCH], he[NCH], ho[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            ipipe_step(pl[c], r.ll[c] * WEMA_K, r.lh[c] * WEMA_KINV, le[c], lo[c]);
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            ipipe_step(ph[c], r.hl[c] * WEMA_K, r.hh[c] * WEMA_KINV, he[c], ho[c]);
        if (out) {
            const int y = 2 * (py0 + m - 4);
            put_row(y, le, he);
            put_row(y + 1, lo, ho);
            if (STAGE) flush(y, 2);
        }
    };

    if (!vert) {
        // one coefficient row: the column synthesis is the identity (src/dwt.c:46), and
        // image row 0 is the even sample
        const IStripRaw<NCH> r = load(0);
        put_row(0, r.ll, r.hl);
        if (STAGE) flush(0, 1);
        return;
    }
    // coefficient pairs py0 - 2 + m for m < M; image rows of pairs py0 .. py0 + nout - 1.
    // FAST strips have nout >= 2 * SLOTS; shorter strips run on to 2 * SLOTS steps past the
    // warm-up (reflected rows, nothing stored), so that no load of the warm-up or of the
    // main loop is issued under a condition and the loads stay in flight across steps.
    constexpr int SLOTS = SL;
    const int nout = min(SP, sh - py0);
    const int M = (FAST ? nout : max(nout, 2 * SLOTS)) + 4;
    auto emits = [&](int m) { return FAST || m - 4 < nout; };
    // (after the 4 warm-up pairs, pair m + k of the loops below is in slot k)
    static_assert(4 % SLOTS == 0, "the slot count must divide the 4 warm-up pairs");
    IStripRaw<NCH> s[SLOTS];
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) s[k] = load(py0 - 2 + k);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        step(m, false, s[m % SLOTS]);
        s[m % SLOTS] = load(py0 - 2 + m + SLOTS);
    }
    int m = 4;
    for (; m + 2 * SLOTS <= M; m += SLOTS) {
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            step(m + k, emits(m + k), s[k]);
            s[k] = load(py0 - 2 + m + k + SLOTS);
        }
    }
    for (; m < M; m += SLOTS) {
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            if (m + k < M) {
                step(m + k, emits(m + k), s[k]);
                if (m + k + SLOTS < M) s[k] = load(py0 - 2 + m + k + SLOTS);
            }
        }
    }
}

template <int NCH, int OUT, bool LEVEL0>
__device__ __forceinline__ void dwt_inv_strips(const DwtiJob *__restrict__ jobs, int njobs) {
    const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int jid = find_job(jobs, njobs, bid);
    const DwtiJob &J = jobs[jid];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int strip = (bid - J.tile_begin) * WEMA_STRIPS_PER_BLOCK + wave;
    if (strip == 0 && J.acc_clear)
        for (int i = lane; i < J.acc_n; i += 64) J.acc_clear[i] = 0ull;
    const int w = J.w, h = J.h, sw = (w + 1) >> 1, sh = (h + 1) >> 1;
    if (strip >= J.tiles_x * wema_ceil_div(sh, J.sp)) return;
    const int px0 = (strip % J.tiles_x) * WEMA_TILE_PAIRS_X, py0 = (strip / J.tiles_x) * J.sp;
    // every lane's columns unreflected (I in [0, sw) with 2 I + 1 < w), one reflection
    // reaches every row, full-length strips, 32-bit plane offsets
    const bool fast = px0 >= 2 && 2 * (px0 + 61) + 1 < w && h >= 8 &&
                      min(J.sp, sh - py0) >= (NCH == 1 && J.sp == 4 ? 4 : 2 * ISLOTS_OF(NCH)) &&
                      (long long)sw * sh < (1ll << 28);
    // (level 0: staged rows where the job's strip rows are whole workgroups; RGB24 rows on
    // dwords, float rows of even width)
    __shared__ float s_stg[LEVEL0 ? 2 * 2 * 2 * WEMA_STAGE_W : 1];
    const bool stage = LEVEL0 && WEMA_INV_STAGE && (J.tiles_x & (WEMA_STRIPS_PER_BLOCK - 1)) == 0 &&
                       gridDim.x >= WEMA_STAGE_MIN_BLOCKS &&
                       (OUT == 0 ? (w & 1) == 0
                                 : J.out_u8 && !J.out_f32 && ((w | (int)(size_t)J.out_u8) & 3) == 0);
    if (stage) {
        if (fast && NCH == 1 && J.sp == 4) dwt_inv_strip_t<NCH, OUT, true, true, 2>(J, strip, lane, s_stg);
        else if (fast) dwt_inv_strip_t<NCH, OUT, true, true>(J, strip, lane, s_stg);
        else dwt_inv_strip_t<NCH, OUT, false, true>(J, strip, lane, s_stg);
    } else {
        if (fast && NCH == 1 && J.sp == 4) dwt_inv_strip_t<NCH, OUT, true, false, 2>(J, strip, lane, s_stg);
        else if (fast) dwt_inv_strip_t<NCH, OUT, true, false>(J, strip, lane, s_stg);
        else dwt_inv_strip_t<NCH, OUT, false, false>(J, strip, lane, s_stg);
    }
}

__global__ void __launch_bounds__(64 * WEMA_STRIPS_PER_BLOCK) k_dwt_inv0(const DwtiJob *__restrict__ jobs, int njobs) {
    dwt_inv_strips<1, 0, true>(jobs, njobs);
}
__global__ void __launch_bounds__(64 * WEMA_STRIPS_PER_BLOCK) k_dwt_inv(const DwtiJob *__restrict__ jobs, int njobs) {
    dwt_inv_strips<1, 0, false>(jobs, njobs);
}
__global__ void __launch_bounds__(64 * WEMA_STRIPS_PER_BLOCK) k_dwt_inv0_ycc(const DwtiJob *__restrict__ jobs, int njobs) {
    dwt_inv_strips<3, 1, true>(jobs, njobs);
}

// ---------------------------------------------------------------------------------------
// Coarser DWT levels as LDS tiles (one float plane per job; src/dwt.c:369-582).
//
// A workgroup computes WEMA_DT_TP x WEMA_DT_TP output pairs of one job from the
// (WEMA_DT_TP + 4)^2 pairs around them (2 pairs of halo on each side, the reach of the four
// lifting steps), staged in LDS. The samples come from the same whole-sample symmetric
// extension the strips read, every lifting step is one pass over the tile with a barrier
// after it, and every value takes the same fused multiply-add of the same operands as in
// the strips (forward: rows, then columns; inverse: columns, then rows), so tiles and
// strips give identical results. A strip walks down its columns one pair row per step; a
// tile has all its loads in flight at once and a depth of eight short passes, which suits
// levels too small to fill the chip with strips. Valid extents after each step (pairs of
// the tile): forward rows/columns step 1 [0, R-1), 2 [1, R-1), 3 [1, R-2), 4 [2, R-2);
// inverse step 1 [1, R), 2 [1, R-1), 3 [2, R-1), 4 [2, R-2): the central WEMA_DT_TP.

constexpr int DT_TP = WEMA_DT_TP;
constexpr int DT_R = DT_TP + 4;          // pairs per side with the halo
constexpr int DT_S = 2 * DT_R;           // samples per side
constexpr int DT_LOADS = (DT_S * DT_S + 255) / 256;   // sample loads per thread

// |v| as Q32 for any float (abs_q32_small where it applies).
__device__ __forceinline__ unsigned long long abs_q32(float v) {
    return fabsf(v) < 1073741824.0f ? abs_q32_small(v) : abs_fixed(v);
}

__global__ void __launch_bounds__(256) k_dwt_fwd_tile(const DwtfJob *__restrict__ jobs, int njobs) {
    __shared__ float sE[DT_S][DT_R];   // even samples of each row, then L
    __shared__ float sO[DT_S][DT_R];   // odd samples of each row
    __shared__ float sH[DT_S][DT_R];   // H (the odd samples after the row lifting, scaled)
    const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const DwtfJob J = jobs[find_job(jobs, njobs, bid)];
    const int t = bid - J.tile_begin, tid = (int)threadIdx.x;
    const int w = J.w, h = J.h, sw = (w + 1) >> 1, sh = (h + 1) >> 1;
    const int ntx = wema_ceil_div(sw, DT_TP);
    const int px0 = (t % ntx) * DT_TP, py0 = (t / ntx) * DT_TP;
    const GLOBAL_AS float *src = gp((const float *)J.src);
    {
        // sample (r, x) of the tile = image sample (2 (py0 - 2) + r, 2 (px0 - 2) + x),
        // reflected into the image; all loads issued before the first LDS store
        float v[DT_LOADS];
#pragma unroll
        for (int k = 0; k < DT_LOADS; ++k) {
            const int q = min(tid + 256 * k, DT_S * DT_S - 1);
            const int r = q / DT_S, x = q - r * DT_S;
            const int yi = reflect_idx(2 * (py0 - 2) + r, h), xi = reflect_idx(2 * (px0 - 2) + x, w);
            v[k] = src[(size_t)yi * w + xi];
        }
#pragma unroll
        for (int k = 0; k < DT_LOADS; ++k) {
            const int q = tid + 256 * k;
            if (q < DT_S * DT_S) {
                const int r = q / DT_S, x = q - r * DT_S;
                (x & 1 ? sO : sE)[r][x >> 1] = v[k];
            }
        }
    }
    __syncthreads();
    // row lifting, as lift_fwd_lanes: item (row r, pair i)
#define DT_ROWS(BODY)                                                   \
    for (int q = tid; q < DT_S * DT_R; q += 256) {                      \
        const int r = q / DT_R, i = q - r * DT_R;                        \
        BODY                                                            \
    }                                                                   \
    __syncthreads();
    DT_ROWS(if (i < DT_R - 1) sO[r][i] = fmaf(WEMA_ALPHA, sE[r][i] + sE[r][i + 1], sO[r][i]);)
    DT_ROWS(if (i > 0) sE[r][i] = fmaf(WEMA_BETA, sO[r][i - 1] + sO[r][i], sE[r][i]);)
    DT_ROWS(if (i < DT_R - 1) sO[r][i] = fmaf(WEMA_GAMMA, sE[r][i] + sE[r][i + 1], sO[r][i]);)
    DT_ROWS(if (i > 0) sE[r][i] = fmaf(WEMA_DELTA, sO[r][i - 1] + sO[r][i], sE[r][i]) * WEMA_KINV;
            sH[r][i] = sO[r][i] * WEMA_K;)
#undef DT_ROWS
    // column lifting of the L (sE) and H (sH) columns, as vpipe_step: item (array, pair row
    // j, column i); rows 2 j and 2 j + 1 are the pair's even and odd samples
#define DT_COLS(BODY)                                                   \
    for (int q = tid; q < 2 * DT_R * DT_R; q += 256) {                  \
        const int a = q >= DT_R * DT_R, qq = q - (a ? DT_R * DT_R : 0); \
        const int j = qq / DT_R, i = qq - j * DT_R;                      \
        float(*X)[DT_R] = a ? sH : sE;                                  \
        BODY                                                            \
    }                                                                   \
    __syncthreads();
    DT_COLS(if (j < DT_R - 1) X[2 * j + 1][i] = fmaf(WEMA_ALPHA, X[2 * j][i] + X[2 * j + 2][i], X[2 * j + 1][i]);)
    DT_COLS(if (j > 0) X[2 * j][i] = fmaf(WEMA_BETA, X[2 * j - 1][i] + X[2 * j + 1][i], X[2 * j][i]);)
    DT_COLS(if (j < DT_R - 1) X[2 * j + 1][i] = fmaf(WEMA_GAMMA, X[2 * j][i] + X[2 * j + 2][i], X[2 * j + 1][i]);)
#undef DT_COLS
    // last column step and the outputs of the central pairs (as the strips' emit)
    const int hw_valid = w >> 1, hh_valid = h >> 1;   // written extents (src/dwt.c:463-469)
    const size_t n = (size_t)sw * sh, pn = (size_t)w * h;
    const GLOBAL_AS float *prev = gp(J.prev);
    unsigned long long part = 0ull;
    for (int q = tid; q < DT_TP * DT_TP; q += 256) {
        const int j = 2 + q / DT_TP, i = 2 + q % DT_TP;
        const int px = px0 + i - 2, py = py0 + j - 2;
        if (px >= sw || py >= sh) continue;
        const float le4 = fmaf(WEMA_DELTA, sE[2 * j - 1][i] + sE[2 * j + 1][i], sE[2 * j][i]);
        const float he4 = fmaf(WEMA_DELTA, sH[2 * j - 1][i] + sH[2 * j + 1][i], sH[2 * j][i]);
        const float ll = le4 * WEMA_KINV, hl = he4 * WEMA_KINV;
        float lh = sE[2 * j + 1][i] * WEMA_K, hh = sH[2 * j + 1][i] * WEMA_K;
        const size_t o = (size_t)py * sw + px;
        float hlv = hl;
        if (prev) {
            if (py >= hh_valid) {
                lh = prev[o];
                hh = prev[pn + o];
            }
            if (px >= hw_valid) {
                hlv = prev[2 * pn + o];
                hh = prev[pn + o];
            }
        }
        gp(J.ll)[o] = ll;
        gp(J.planes)[o] = lh;
        gp(J.planes)[n + o] = hh;
        gp(J.planes)[2 * n + o] = hlv;
        part += abs_q32(lh) + abs_q32(hlv) + abs_q32(hh);
    }
    if (J.acc) {
        const unsigned long long s = wave_sum_u64(part);
        const int wave = tid >> 6;
        if ((tid & 63) == 0) atomicAdd(J.acc + ((4 * t + wave) & (J.acc_slots - 1)), s);
    }
}

__global__ void __launch_bounds__(256) k_dwt_inv_tile(const DwtiJob *__restrict__ jobs, int njobs) {
    // column phase: even (e) and odd (o) coefficient rows of the low (L) and high (H)
    // columns, by pair row; row phase: even (A) and odd (B) samples of each image row
    __shared__ float sLe[DT_R][DT_R], sLo[DT_R][DT_R], sHe[DT_R][DT_R], sHo[DT_R][DT_R];
    __shared__ float sA[DT_S][DT_R], sB[DT_S][DT_R];
    const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const DwtiJob J = jobs[find_job(jobs, njobs, bid)];
    const int t = bid - J.tile_begin, tid = (int)threadIdx.x;
    if (t == 0 && J.acc_clear)
        for (int i = tid; i < J.acc_n; i += 256) J.acc_clear[i] = 0ull;
    const int w = J.w, h = J.h, sw = (w + 1) >> 1, sh = (h + 1) >> 1;
    const size_t n = (size_t)sw * sh;
    const int ntx = wema_ceil_div(sw, DT_TP);
    const int px0 = (t % ntx) * DT_TP, py0 = (t / ntx) * DT_TP;
    {
        // coefficient pair (j, i) of the tile = pair (py0 - 2 + j, px0 - 2 + i) of the
        // symmetric extension of the interleaved signal, as the strips load it
        constexpr int NP = DT_R * DT_R, NL = (NP + 255) / 256;
        float v[NL][4];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int q = min(tid + 256 * k, NP - 1);
            const int j = q / DT_R, i = q - j * DT_R;
            const int I = px0 - 2 + i, Q = py0 - 2 + j;
            const size_t colL = (size_t)(reflect_idx(2 * I, w) >> 1), colH = (size_t)(reflect_idx(2 * I + 1, w) >> 1);
            const size_t re = (size_t)(reflect_idx(2 * Q, h) >> 1) * sw, ro = (size_t)(reflect_idx(2 * Q + 1, h) >> 1) * sw;
            const GLOBAL_AS float *LL = gp(J.ll), *LH = gp(J.planes);
            v[k][0] = LL[re + colL];
            v[k][1] = LH[ro + colL];
            v[k][2] = LH[2 * n + re + colH];
            v[k][3] = LH[n + ro + colH];
        }
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int q = tid + 256 * k;
            if (q < NP) {
                const int j = q / DT_R, i = q - j * DT_R;
                sLe[j][i] = v[k][0] * WEMA_K;
                sLo[j][i] = v[k][1] * WEMA_KINV;
                sHe[j][i] = v[k][2] * WEMA_K;
                sHo[j][i] = v[k][3] * WEMA_KINV;
            }
        }
    }
    __syncthreads();
    // column synthesis, as ipipe_step: item (column kind, pair row j, column i)
#define DT_ICOLS(BODY)                                                  \
    for (int q = tid; q < 2 * DT_R * DT_R; q += 256) {                  \
        const int a = q >= DT_R * DT_R, qq = q - (a ? DT_R * DT_R : 0); \
        const int j = qq / DT_R, i = qq - j * DT_R;                      \
        float(*Xe)[DT_R] = a ? sHe : sLe;                               \
        float(*Xo)[DT_R] = a ? sHo : sLo;                               \
        BODY                                                            \
    }                                                                   \
    __syncthreads();
    DT_ICOLS(if (j > 0) Xe[j][i] = fmaf(-WEMA_DELTA, Xo[j - 1][i] + Xo[j][i], Xe[j][i]);)
    DT_ICOLS(if (j < DT_R - 1) Xo[j][i] = fmaf(-WEMA_GAMMA, Xe[j][i] + Xe[j + 1][i], Xo[j][i]);)
    DT_ICOLS(if (j > 0) Xe[j][i] = fmaf(-WEMA_BETA, Xo[j - 1][i] + Xo[j][i], Xe[j][i]);)
    // last column step; image rows 2 j (e) and 2 j + 1 (o) to the row arrays, scaled as
    // put_row scales them (low column: A, times K; high column: B, times 1/K)
    DT_ICOLS(if (j >= 2 && j < DT_R - 2) {
        const float o4 = fmaf(-WEMA_ALPHA, Xe[j][i] + Xe[j + 1][i], Xo[j][i]);
        if (a) {
            sB[2 * j][i] = Xe[j][i] * WEMA_KINV;
            sB[2 * j + 1][i] = o4 * WEMA_KINV;
        } else {
            sA[2 * j][i] = Xe[j][i] * WEMA_K;
            sA[2 * j + 1][i] = o4 * WEMA_K;
        }
    })
#undef DT_ICOLS
    // row synthesis of image rows 4 .. S - 5 (the central pair rows), as lift_inv_lanes
#define DT_IROWS(BODY)                                                  \
    for (int q = tid; q < (DT_S - 8) * DT_R; q += 256) {                \
        const int r = 4 + q / DT_R, i = q % DT_R;                        \
        BODY                                                            \
    }                                                                   \
    __syncthreads();
    DT_IROWS(if (i > 0) sA[r][i] = fmaf(-WEMA_DELTA, sB[r][i - 1] + sB[r][i], sA[r][i]);)
    DT_IROWS(if (i < DT_R - 1) sB[r][i] = fmaf(-WEMA_GAMMA, sA[r][i] + sA[r][i + 1], sB[r][i]);)
    DT_IROWS(if (i > 0) sA[r][i] = fmaf(-WEMA_BETA, sB[r][i - 1] + sB[r][i], sA[r][i]);)
#undef DT_IROWS
    // last row step and the stores of the central pairs
    GLOBAL_AS float *out = gp(J.out);
    for (int q = tid; q < 2 * DT_TP * DT_TP; q += 256) {
        const int r = 4 + q / DT_TP, i = 2 + q % DT_TP;
        const int x = 2 * (px0 + i - 2), y = 2 * (py0 - 2) + r;
        if (x >= w || y >= h) continue;
        const float ev = sA[r][i];
        const float od = fmaf(-WEMA_ALPHA, sA[r][i] + sA[r][i + 1], sB[r][i]);
        out[(size_t)y * w + x] = ev;
        if (x + 1 < w) out[(size_t)y * w + x + 1] = od;
    }
}

// ---------------------------------------------------------------------------------------
// Coarser DWT levels, one output pair per thread (src/dwt.c:369-582). Each thread recomputes
// the lifting steps its pair depends on from the symmetric extension the strips and tiles
// read, every value by the same fused multiply-add of the same operands, so the results
// are identical to theirs; the neighbours' intermediate values are computed again instead
// of shared, which costs a few times the arithmetic and leaves a chain of a few dozen
// instructions per thread, with all 81 loads in flight at once.

// Forward lifting of one pair p from the samples v[k] = x[2p - 4 + k], k = 0..8 (the reach
// of the four steps): lo = the low-pass value of pair p, hi = the high-pass one. With E[m]
// = v[2m + 4], O[m] = v[2m + 5] the samples of pair p + m, as lift_fwd_lanes and the tiles'
// row and column passes.
__device__ __forceinline__ void lift9(const float (&v)[9], float &lo, float &hi) {
    float o1[4];   // O after step 1, m = -2..1
#pragma unroll
    for (int m = -2; m <= 1; ++m) o1[m + 2] = fmaf(WEMA_ALPHA, v[2 * m + 4] + v[2 * m + 6], v[2 * m + 5]);
    float e2[3];   // E after step 2, m = -1..1
#pragma unroll
    for (int m = -1; m <= 1; ++m) e2[m + 1] = fmaf(WEMA_BETA, o1[m + 1] + o1[m + 2], v[2 * m + 4]);
    float o3[2];   // O after step 3, m = -1..0
#pragma unroll
    for (int m = -1; m <= 0; ++m) o3[m + 1] = fmaf(WEMA_GAMMA, e2[m + 1] + e2[m + 2], o1[m + 2]);
    lo = fmaf(WEMA_DELTA, o3[0] + o3[1], e2[1]) * WEMA_KINV;
    hi = o3[1] * WEMA_K;
}

__global__ void __launch_bounds__(256) k_dwt_fwd_point(const DwtfJob *__restrict__ jobs, int njobs) {
    const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const DwtfJob J = jobs[find_job(jobs, njobs, bid)];
    const int t = bid - J.tile_begin, tid = (int)threadIdx.x;
    const int w = J.w, h = J.h, sw = (w + 1) >> 1, sh = (h + 1) >> 1;
    const int q = t * 256 + tid;
    const bool valid = q < sw * sh;
    const int qv = valid ? q : 0;   // (lanes past the end compute pair 0 and store nothing)
    const int py = qv / sw, px = qv - py * sw;
    const GLOBAL_AS float *src = gp((const float *)J.src);
    int xi[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) xi[k] = reflect_idx(2 * px - 4 + k, w);
    // rows 2 py - 4 .. 2 py + 4, lifted along the row: the L and H columns of pair px
    float L[9], H[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const GLOBAL_AS float *row = src + (size_t)reflect_idx(2 * py - 4 + r, h) * w;
        float v[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) v[k] = row[xi[k]];
        lift9(v, L[r], H[r]);
    }
    // the columns, lifted: (LL, LH) from the L column, (HL, HH) from the H column
    float ll, lh, hl, hh;
    lift9(L, ll, lh);
    lift9(H, hl, hh);
    unsigned long long part = 0ull;
    if (valid) {
        // stale entries of the reference's planes outside the written extents (as the tiles)
        const int hw_valid = w >> 1, hh_valid = h >> 1;
        const size_t n = (size_t)sw * sh, pn = (size_t)w * h, o = (size_t)py * sw + px;
        const GLOBAL_AS float *prev = gp(J.prev);
        if (prev) {
            if (py >= hh_valid) {
                lh = prev[o];
                hh = prev[pn + o];
            }
            if (px >= hw_valid) {
                hl = prev[2 * pn + o];
                hh = prev[pn + o];
            }
        }
        gp(J.ll)[o] = ll;
        gp(J.planes)[o] = lh;
        gp(J.planes)[n + o] = hh;
        gp(J.planes)[2 * n + o] = hl;
        part = abs_q32(lh) + abs_q32(hl) + abs_q32(hh);
    }
    if (J.acc) {
        const unsigned long long s = wave_sum_u64(part);
        const int wave = tid >> 6;
        if ((tid & 63) == 0) atomicAdd(J.acc + ((4 * t + wave) & (J.acc_slots - 1)), s);
    }
}

// Inverse lifting of one coefficient pair row (or column) m = 0 from the even (e) and odd
// (o) coefficients of pairs m = -2..2 (index m + 2; e[0] is not read): the even and the odd
// output sample, as ipipe_step and the tiles' column pass (the row pass has the same steps).
__device__ __forceinline__ void ilift5(const float (&e)[5], const float (&o)[5], float &ev, float &od) {
    float e1[5];   // E after step 1, index 1..4
#pragma unroll
    for (int j = 1; j <= 4; ++j) e1[j] = fmaf(-WEMA_DELTA, o[j - 1] + o[j], e[j]);
    float o2[4];   // O after step 2, index 1..3
#pragma unroll
    for (int j = 1; j <= 3; ++j) o2[j] = fmaf(-WEMA_GAMMA, e1[j] + e1[j + 1], o[j]);
    float e3[4];   // E after step 3, index 2..3
#pragma unroll
    for (int j = 2; j <= 3; ++j) e3[j] = fmaf(-WEMA_BETA, o2[j - 1] + o2[j], e1[j]);
    ev = e3[2];
    od = fmaf(-WEMA_ALPHA, e3[2] + e3[3], o2[2]);
}

__global__ void __launch_bounds__(256) k_dwt_inv_point(const DwtiJob *__restrict__ jobs, int njobs) {
    const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const DwtiJob J = jobs[find_job(jobs, njobs, bid)];
    const int t = bid - J.tile_begin, tid = (int)threadIdx.x;
    if (t == 0 && J.acc_clear)
        for (int i = tid; i < J.acc_n; i += 256) J.acc_clear[i] = 0ull;
    const int w = J.w, h = J.h, sw = (w + 1) >> 1, sh = (h + 1) >> 1;
    const size_t n = (size_t)sw * sh;
    const int q = t * 256 + tid;
    const bool valid = q < sw * sh;
    const int qv = valid ? q : 0;
    const int py = qv / sw, px = qv - py * sw;
    // coefficient pair (py - 2 + j, px - 2 + c) of the symmetric extension of the
    // interleaved signal, as the strips and tiles load it
    int colL[5], colH[5];
    size_t re[5], ro[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const int I = px - 2 + c, Q = py - 2 + c;
        colL[c] = reflect_idx(2 * I, w) >> 1;
        colH[c] = reflect_idx(2 * I + 1, w) >> 1;
        re[c] = (size_t)(reflect_idx(2 * Q, h) >> 1) * sw;
        ro[c] = (size_t)(reflect_idx(2 * Q + 1, h) >> 1) * sw;
    }
    const GLOBAL_AS float *LL = gp(J.ll), *LH = gp(J.planes), *HH = LH + n, *HL = LH + 2 * n;
    // column synthesis: image rows 2 py (index 0) and 2 py + 1 (index 1) of the low (A, pair
    // columns c = 1..4) and high (B, c = 0..4) columns, scaled as the row pass takes them
    float A[2][5], B[2][5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        float e[5], o[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            e[j] = j ? HL[re[j] + colH[c]] * WEMA_K : 0.0f;
            o[j] = HH[ro[j] + colH[c]] * WEMA_KINV;
        }
        float ev, od;
        ilift5(e, o, ev, od);
        B[0][c] = ev * WEMA_KINV;
        B[1][c] = od * WEMA_KINV;
        if (c == 0) {
            A[0][0] = A[1][0] = 0.0f;   // (not read)
            continue;
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            e[j] = j ? LL[re[j] + colL[c]] * WEMA_K : 0.0f;
            o[j] = LH[ro[j] + colL[c]] * WEMA_KINV;
        }
        ilift5(e, o, ev, od);
        A[0][c] = ev * WEMA_K;
        A[1][c] = od * WEMA_K;
    }
    if (!valid) return;
    // row synthesis of pair px of each of the two rows: even samples A, odd samples B
    GLOBAL_AS float *out = gp(J.out);
    const int x = 2 * px;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int y = 2 * py + r;
        float ev, od;
        ilift5(A[r], B[r], ev, od);
        if (y < h) {
            out[(size_t)y * w + x] = ev;
            if (x + 1 < w) out[(size_t)y * w + x + 1] = od;
        }
    }
}

// ---------------------------------------------------------------------------------------
// IIR band-pass + amplification (src/iir_filter.c:181-264, 266-370, 376-471;
// src/wema.c:714-744, 783-800, 830-848; src/wema_batch.c:268-297, 420-440).
__device__ __forceinline__ float biquad2(const float *c, float4 &st, float x) {
    // DF2T, FMA sequence of the reference's AVX-512 kernel (src/iir_filter.c:211-219)
    float y = fmaf(c[0], x, st.x);
    float nw1 = fmaf(-c[3], y, fmaf(c[1], x, st.y));
    float nw2 = fmaf(-c[4], y, c[2] * x);
    st.x = nw1;
    st.y = nw2;
    x = y;
    y = fmaf(c[5], x, st.z);
    nw1 = fmaf(-c[8], y, fmaf(c[6], x, st.w));
    nw2 = fmaf(-c[9], y, c[7] * x);
    st.z = nw1;
    st.w = nw2;
    return y;
}

// Thresholds and per-level amplification of the job's frames, computed once per block:
// thread t reads slot t mod S of frame t / S (S = acc_slots, a power of two up to 64, so the
// S slots of a frame sit in S consecutive lanes of one wave), and the S lanes add them up.
__device__ __forceinline__ void frame_params(const TemporalJob &J, float *sT, float *sA) {
    const int S = J.acc_slots, total = J.nframes * S;
    for (int base = 0; base < total; base += (int)blockDim.x) {   // (the same trip count in every lane)
        const int i = base + (int)threadIdx.x;
        const int f = i / S, k = i & (S - 1);
        unsigned long long v = i < total ? J.acc[(size_t)f * J.acc_stride + k] : 0ull;
        for (int off = 1; off < S; off <<= 1) v += __shfl_xor(v, off);
        if (i < total && k == 0) {
            float T = threshold_of(v, J.n6f);
            if (J.thr_override) {
                const float o = J.thr_override[3 * f];
                if (o == o) T = o;
            }
            sT[f] = T;
            sA[f] = J.amp[f];
            if (J.thr_out && blockIdx.x == (unsigned)J.tile_begin) J.thr_out[3 * f] = T;
        }
    }
    __syncthreads();
}

// One frame of one position: biquads, then the threshold mask and amplification
// (src/wema.c:727-743: |c| <= T keeps c, otherwise fma(alpha_l, y, c)).
__device__ __forceinline__ float iir_amp(const float *c, float4 &st, float x, float T, float a,
                                         float &y) {
    y = biquad2(c, st, x);
    return (fabsf(x) <= T) ? x : fmaf(a, y, x);
}

template <bool SEP, bool FILT>
__device__ __forceinline__ void iir_frames(const TemporalJob &J, const float *c, float4 &st,
                                           const float *sT, const float *sA, float lscale,
                                           long long p) {
    // frames in groups of WEMA_IIR_G with the coefficient loads issued first (independent
    // of the state recurrence), so several loads per thread are in flight; the last, partial
    // group likewise. Every coefficient is read once and written once per launch: the
    // accesses are marked nontemporal (at 1080p colour 685 -> 598 us per launch; either
    // kind alone gained nothing), and so are the state's.
    const int nframes = J.nframes;
    const size_t stride = (size_t)J.coef_stride, ostride = (size_t)J.out_stride;
    GLOBAL_AS float *in = gp(J.coef) + p;
    GLOBAL_AS float *out = (SEP ? gp(J.amp_out) : gp(J.coef)) + p;
    GLOBAL_AS float *fo = FILT ? gp(J.filt_out) + p : nullptr;
    const size_t ostep = SEP ? ostride : stride;
    constexpr int G = WEMA_IIR_G;
    int f = 0;
    for (; f + G <= nframes; f += G) {
        float x[G];
#pragma unroll
        for (int k = 0; k < G; ++k) x[k] = __builtin_nontemporal_load(&in[k * stride]);
#pragma unroll
        for (int k = 0; k < G; ++k) {
            float y;
            const float o = iir_amp(c, st, x[k], sT[f + k], sA[f + k] * lscale, y);
            __builtin_nontemporal_store(o, &out[k * ostep]);
            if (FILT) fo[k * ostride] = y;
        }
        in += G * stride;
        out += G * ostep;
        if (FILT) fo += G * ostride;
    }
    if (f < nframes) {
        const int m = nframes - f;
        float x[G];
#pragma unroll
        for (int k = 0; k < G; ++k)
            if (k < m) x[k] = __builtin_nontemporal_load(&in[k * stride]);
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if (k >= m) break;
            float y;
            const float o = iir_amp(c, st, x[k], sT[f + k], sA[f + k] * lscale, y);
            __builtin_nontemporal_store(o, &out[k * ostep]);
            if (FILT) fo[k * ostride] = y;
        }
    }
}

// One frame, PPT positions per thread (p0 + 256 k), the path of one frame per call and of
// the serving loop's ticks: 40 bytes move per position, so a block of one position per
// thread would spend most of its life on its set-up (finding its job, the frame's threshold)
// and on one memory round trip; here the loads of all PPT positions are issued first, and
// the frame's threshold is summed within each wave (lane l takes slot l of the frame's at
// most 64 accumulator slots) while they are in flight: no barrier, one round trip.
template <int PPT, bool SEP, bool FILT>
__device__ __forceinline__ void iir_one_frame(const TemporalJob &J, const float *c, long long p0) {
    GLOBAL_AS float *in = gp(J.coef);
    GLOBAL_AS float *out = SEP ? gp(J.amp_out) : gp(J.coef);
    GLOBAL_AS float *sp = gp((float *)J.state);
    // slot lane mod acc_slots of the frame, loaded first: the wait for it then leaves the
    // loads below in flight (loads complete in issue order as the wait counter sees them)
    const int lane = (int)threadIdx.x & 63;
    unsigned long long v = gp(J.acc)[lane & (J.acc_slots - 1)];
    float x[PPT];
    float4 st[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const long long p = min(p0 + 256 * k, J.n - 1);   // (a valid address)
        x[k] = __builtin_nontemporal_load(&in[p]);
        // (loaded on the preseed frame too, and replaced: a load under a condition would
        // make the wait below wait for every load)
        st[k] = make_float4(__builtin_nontemporal_load(&sp[4 * p]), __builtin_nontemporal_load(&sp[4 * p + 1]),
                            __builtin_nontemporal_load(&sp[4 * p + 2]), __builtin_nontemporal_load(&sp[4 * p + 3]));
    }
    // (each group of acc_slots lanes adds the frame's slots; uniform trip count)
    for (int off = 1; off < J.acc_slots; off <<= 1) v += __shfl_xor(v, off);
    float T = threshold_of(v, J.n6f);
    if (J.thr_override) {
        const float o = gp(J.thr_override)[0];
        if (o == o) T = o;
    }
    const float A = gp(J.amp)[0];
    if (J.thr_out && blockIdx.x == (unsigned)J.tile_begin && threadIdx.x == 0) gp(J.thr_out)[0] = T;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const long long p = p0 + 256 * k;
        if (J.preseed) {
            // warm-up on frame 0 (src/iir_filter.c:376-471), register-resident
            st[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int it = 0; it < J.warmup; ++it) (void)biquad2(c, st[k], x[k]);
        }
        float y;
        const float o = iir_amp(c, st[k], x[k], T, A * (p < J.level0_end ? 0.5f : 1.0f), y);
        if (p < J.n) {
            __builtin_nontemporal_store(o, &out[p]);
            if (FILT) gp(J.filt_out)[p] = y;
            __builtin_nontemporal_store(st[k].x, &sp[4 * p]);
            __builtin_nontemporal_store(st[k].y, &sp[4 * p + 1]);
            __builtin_nontemporal_store(st[k].z, &sp[4 * p + 2]);
            __builtin_nontemporal_store(st[k].w, &sp[4 * p + 3]);
        }
    }
}

__device__ __forceinline__ void iir_one_frame_job(const TemporalJob &J, const float *c) {
    // (every lane of a wave takes part in the threshold sum; positions past the end load a
    // valid address and store nothing)
    const long long p0 = (long long)((int)blockIdx.x - J.tile_begin) * (256 * WEMA_IIR_PPT) + threadIdx.x;
    if (J.amp_out) {
        if (J.filt_out) iir_one_frame<WEMA_IIR_PPT, true, true>(J, c, p0);
        else iir_one_frame<WEMA_IIR_PPT, true, false>(J, c, p0);
    } else {
        if (J.filt_out) iir_one_frame<WEMA_IIR_PPT, false, true>(J, c, p0);
        else iir_one_frame<WEMA_IIR_PPT, false, false>(J, c, p0);
    }
}

// Launches whose jobs are all of one frame: the one-frame path alone, so the waits the
// compiler places in it are not merged with those of the many-frame path.
__global__ void __launch_bounds__(256) k_iir1(const TemporalJob *__restrict__ jobs, int njobs) {
    const TemporalJob J = jobs[find_job(jobs, njobs, (int)blockIdx.x)];
    float c[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) c[i] = J.c[i];
    iir_one_frame_job(J, c);
}

__global__ void __launch_bounds__(256) k_iir(const TemporalJob *__restrict__ jobs, int njobs) {
    __shared__ float sT[WEMA_MAX_BATCH], sA[WEMA_MAX_BATCH];
    const int jid = find_job(jobs, njobs, (int)blockIdx.x);
    const TemporalJob J = jobs[jid];
    float c[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) c[i] = J.c[i];
    if (J.ppt == WEMA_IIR_PPT && J.nframes == 1) {
        iir_one_frame_job(J, c);
        return;
    }
    frame_params(J, sT, sA);
    const long long p = (long long)((int)blockIdx.x - J.tile_begin) * 256 + threadIdx.x;
    if (p >= J.n) return;
    GLOBAL_AS float *sp = gp((float *)J.state) + 4 * p;
    float4 st;
    if (J.preseed) {
        // warm-up on frame 0 (src/iir_filter.c:376-471), register-resident
        st = make_float4(0.f, 0.f, 0.f, 0.f);
        const float x0 = gp(J.coef)[p];
        for (int it = 0; it < J.warmup; ++it) (void)biquad2(c, st, x0);
    } else {
        st = make_float4(__builtin_nontemporal_load(&sp[0]), __builtin_nontemporal_load(&sp[1]),
                         __builtin_nontemporal_load(&sp[2]), __builtin_nontemporal_load(&sp[3]));
    }
    const float lscale = p < J.level0_end ? 0.5f : 1.0f;
    if (J.amp_out) {
        if (J.filt_out) iir_frames<true, true>(J, c, st, sT, sA, lscale, p);
        else iir_frames<true, false>(J, c, st, sT, sA, lscale, p);
    } else {
        if (J.filt_out) iir_frames<false, true>(J, c, st, sT, sA, lscale, p);
        else iir_frames<false, false>(J, c, st, sT, sA, lscale, p);
    }
    __builtin_nontemporal_store(st.x, &sp[0]);
    __builtin_nontemporal_store(st.y, &sp[1]);
    __builtin_nontemporal_store(st.z, &sp[2]);
    __builtin_nontemporal_store(st.w, &sp[3]);
}

// ---------------------------------------------------------------------------------------
// Sliding-window temporal filter + amplification (src/temporal_filter.c:139-319,
// src/phase_amp.c:12-56). Window of frame f: ring slots ring_first + f + t (mod ring_slots),
// t = 0..ws-1, oldest first; tw holds the Gaussian time weights of the bilateral mode.
//
// One thread per coefficient position runs all frames of the job. Per frame the window is
// read twice, oldest first: once for the mean (and, bilateral, the mean absolute deviation
// from the centre sample), once for the bilateral weights and the weighted sum (the
// reference's two weight loops read the same values with the same weights, so one pass
// with one exponential per sample gives the same sums). Each weight is one hardware exp2
// of the summed base-2 exponents of its value and time weights (relative error of a few
// 1e-7 where it matters, against about 6e-8 for libm's expf and the product).

// Temporal output of one frame from its window, given in order by win.each(fn(t, v)).
// ltw: base-2 logarithms of the time weights; the weight of a sample is
// exp2(fma(d^2, -log2(e) / (2 sigma^2), log2(tw))), the product of the value and time
// weights in one exponential.
template <class Win, class LTW>
__device__ __forceinline__ float win_delta(const Win &win, int ws, float inv_ws, float cur, bool bilateral,
                                           LTW ltw, float &mean_out) {
    float mean = 0.0f, mad = 0.0f;
    win.each([&](int, float v) {
        mean += v;
        if (bilateral) mad += fabsf(v - cur);   // d > 0 ? d : -d, the same sum from +0
    });
    mean *= inv_ws;
    mean_out = mean;
    float delta = cur - mean;
    if (bilateral) {
        const float sigma = fmaxf(mad * inv_ws * 0.5f, 0.001f);
        const float inv2 = 1.0f / (2.0f * sigma * sigma);
        const float c = inv2 * -1.4426950408889634f;   // -log2(e) / (2 sigma^2)
        float wsum = 0.0f, acc = 0.0f;
        win.each([&](int t, float v) {
            const float d = v - cur;
            const float wt = __builtin_amdgcn_exp2f(fmaf(d * d, c, ltw[t]));
            wsum += wt;
            acc = fmaf(wt, v - mean, acc);
        });
        if (wsum > 1e-8f) delta = acc * (1.0f / wsum);
    }
    return delta;
}

// Window read from the ring in global memory.
struct WinGlobal {
    const float *ring;
    size_t plane;
    long long p;
    int first, slots, ws;
    template <class F>
    __device__ __forceinline__ void each(F &&fn) const {
        int slot = first;
        for (int t = 0; t < ws; ++t) {
            fn(t, gp(ring)[(size_t)slot * plane + p]);
            if (++slot == slots) slot = 0;
        }
    }
};

// Window held in LDS, one column per thread: entries head..ws-1 then 0..head-1.
template <int WSMAX>
struct WinLds {
    float (*W)[256];
    int head, ws;
    template <class F>
    __device__ __forceinline__ void each(F &&fn) const {
        int t = 0;
        for (int j = head; j < ws; ++j) fn(t++, W[j][threadIdx.x]);
        for (int j = 0; j < head; ++j) fn(t++, W[j][threadIdx.x]);
    }
};

__device__ __forceinline__ void win_store(const TemporalJob &J, int f, long long p, float orig, float delta,
                                          const float *sT, const float *sA, float lscale) {
    const float T = sT[f];
    const float a = sA[f] * lscale;
    const float o = (fabsf(orig) <= T) ? orig : fmaf(a, delta, orig);
    __builtin_nontemporal_store(o, gp(J.amp_out) + ((size_t)f * J.out_stride + p));   // written once
    if (J.filt_out) J.filt_out[(size_t)f * J.out_stride + p] = delta;
}

// Any window size: the window is read from the ring for every frame.
__global__ void __launch_bounds__(256) k_win(const TemporalJob *__restrict__ jobs, int njobs) {
    __shared__ float sT[WEMA_MAX_BATCH], sA[WEMA_MAX_BATCH];
    const int jid = find_job(jobs, njobs, (int)blockIdx.x);
    const TemporalJob J = jobs[jid];
    frame_params(J, sT, sA);
    const long long p = (long long)((int)blockIdx.x - J.tile_begin) * 256 + threadIdx.x;
    const int nframes = J.nframes, ws = J.window, center = ws / 2;
    if (p >= J.n) return;
    const float inv_ws = 1.0f / (float)ws;
    const float lscale = p < J.level0_end ? 0.5f : 1.0f;
    const size_t plane = (size_t)J.coef_stride;
    for (int f = 0; f < nframes; ++f) {
        const int first = (J.ring_first + f) % J.ring_slots;
        const WinGlobal win{J.ring, plane, p, first, J.ring_slots, ws};
        const float cur = gp(J.ring)[(size_t)((first + center) % J.ring_slots) * plane + p];
        const float orig = gp(J.ring)[(size_t)((first + ws - 1) % J.ring_slots) * plane + p];
        float mean;
        const float delta = win_delta(win, ws, inv_ws, cur, J.bilateral != 0, J.ltw, mean);
        win_store(J, f, p, orig, delta, sT, sA, lscale);
    }
}

// Windows of up to WSMAX samples: each position's window slides through an LDS column, so
// every ring sample is read from HBM once per launch (ws - 1 + nframes reads per position
// instead of ws per frame).
template <int WSMAX>
__device__ __forceinline__ void win_lds_body(const TemporalJob *__restrict__ jobs, int njobs, float (*W)[256],
                                             float *sT, float *sA, float *sTw) {
    const int jid = find_job(jobs, njobs, (int)blockIdx.x);
    const TemporalJob J = jobs[jid];
    frame_params(J, sT, sA);
    const int nframes = J.nframes, ws = J.window, center = ws / 2;
    if (J.bilateral)
        for (int t = threadIdx.x; t < ws; t += blockDim.x) sTw[t] = J.ltw[t];
    __syncthreads();
    const long long p = (long long)((int)blockIdx.x - J.tile_begin) * 256 + threadIdx.x;
    if (p >= J.n) return;
    const float inv_ws = 1.0f / (float)ws;
    const float lscale = p < J.level0_end ? 0.5f : 1.0f;
    const size_t plane = (size_t)J.coef_stride;
    const GLOBAL_AS float *ring = gp(J.ring);
    // the first ws - 1 samples of frame 0's window
    int slot = J.ring_first;
    for (int t = 0; t < ws - 1; ++t) {
        W[t][threadIdx.x] = ring[(size_t)slot * plane + p];
        if (++slot == J.ring_slots) slot = 0;
    }
    int head = 0;   // LDS entry of the oldest sample of the current window
    for (int f = 0; f < nframes; ++f) {
        // the newest sample of frame f's window (ring slot ring_first + f + ws - 1)
        const float orig = ring[(size_t)slot * plane + p];
        if (++slot == J.ring_slots) slot = 0;
        int tail = head + ws - 1;
        if (tail >= ws) tail -= ws;
        W[tail][threadIdx.x] = orig;
        int c = head + center;
        if (c >= ws) c -= ws;
        const float cur = W[c][threadIdx.x];
        const WinLds<WSMAX> win{W, head, ws};
        float mean;
        const float delta = win_delta(win, ws, inv_ws, cur, J.bilateral != 0, sTw, mean);
        win_store(J, f, p, orig, delta, sT, sA, lscale);
        if (++head == ws) head = 0;
    }
}

// Windows of exactly WS samples held in registers, frames in groups of WIN_U: frame j of a
// group sees x[j .. j + WS - 1], so every access is static; after the group the window moves
// down WIN_U registers (WS - 1 moves per WIN_U frames) and takes the WIN_U newest samples,
// which were loaded while the group computed. The bilateral pass keeps each sample's
// deviation from the centre sample from the first pass. The sums are those of win_delta,
// in the same order.
constexpr int WIN_U = 4;

template <int WS, bool BIL>
__device__ __forceinline__ float win_frame_reg(const float *x, float inv_ws, const GLOBAL_AS float *ltw) {
    const float cur = x[WS / 2];
    float mean = 0.0f, mad = 0.0f;
    float d[BIL ? WS : 1];
#pragma unroll
    for (int t = 0; t < WS; ++t) {
        mean += x[t];
        if (BIL) {
            d[t] = x[t] - cur;
            mad += fabsf(d[t]);   // d > 0 ? d : -d, the same sum from +0
        }
    }
    mean *= inv_ws;
    float delta = cur - mean;
    if (BIL) {
        const float sigma = fmaxf(mad * inv_ws * 0.5f, 0.001f);
        const float inv2 = 1.0f / (2.0f * sigma * sigma);
        const float c = inv2 * -1.4426950408889634f;   // -log2(e) / (2 sigma^2)
        float wsum = 0.0f, acc = 0.0f;
#pragma unroll
        for (int t = 0; t < WS; ++t) {
            const float wt = __builtin_amdgcn_exp2f(fmaf(d[t] * d[t], c, ltw[t]));
            wsum += wt;
            acc = fmaf(wt, x[t] - mean, acc);
        }
        if (wsum > 1e-8f) delta = acc * (1.0f / wsum);
    }
    return delta;
}

template <int WS, bool BIL>
__device__ __forceinline__ void win_reg_run(const TemporalJob &J, long long p, const float *sT, const float *sA) {
    const int nframes = J.nframes;
    const float inv_ws = 1.0f / (float)WS;
    const float lscale = p < J.level0_end ? 0.5f : 1.0f;
    const size_t plane = (size_t)J.coef_stride;
    const GLOBAL_AS float *ring = gp(J.ring);
    const GLOBAL_AS float *ltw = gp(J.ltw);
    float x[WS + WIN_U - 1];
    // the first WS - 1 samples of frame 0's window, then the newest samples of frames
    // 0 .. WIN_U - 1 (ring slot ring_first + f + WS - 1 for frame f)
    int slot = J.ring_first;
#pragma unroll
    for (int t = 0; t < WS - 1; ++t) {
        x[t] = ring[(size_t)slot * plane + p];
        if (++slot == J.ring_slots) slot = 0;
    }
#pragma unroll
    for (int j = 0; j < WIN_U; ++j) {
        if (j < nframes) x[WS - 1 + j] = __builtin_nontemporal_load(&ring[(size_t)slot * plane + p]);
        if (++slot == J.ring_slots) slot = 0;
    }
    for (int f0 = 0; f0 < nframes; f0 += WIN_U) {
        float nxt[WIN_U];
#pragma unroll
        for (int j = 0; j < WIN_U; ++j) {
            if (f0 + WIN_U + j < nframes) nxt[j] = __builtin_nontemporal_load(&ring[(size_t)slot * plane + p]);
            if (++slot == J.ring_slots) slot = 0;
        }
#pragma unroll
        for (int j = 0; j < WIN_U; ++j) {
            const int f = f0 + j;
            if (f < nframes) {
                const float delta = win_frame_reg<WS, BIL>(x + j, inv_ws, ltw);
                win_store(J, f, p, x[j + WS - 1], delta, sT, sA, lscale);
            }
        }
#pragma unroll
        for (int t = 0; t < WS - 1; ++t) x[t] = x[t + WIN_U];
#pragma unroll
        for (int j = 0; j < WIN_U; ++j) x[WS - 1 + j] = nxt[j];
    }
}

template <int WS>
__device__ __forceinline__ void win_reg_body(const TemporalJob *__restrict__ jobs, int njobs, float *sT,
                                             float *sA) {
    const int jid = find_job(jobs, njobs, (int)blockIdx.x);
    const TemporalJob J = jobs[jid];
    frame_params(J, sT, sA);
    const long long p = (long long)((int)blockIdx.x - J.tile_begin) * 256 + threadIdx.x;
    if (p >= J.n) return;
    if (J.bilateral) win_reg_run<WS, true>(J, p, sT, sA);
    else win_reg_run<WS, false>(J, p, sT, sA);
}

#define WEMA_WIN_REG_KERNEL(WS)                                                                 \
    __global__ void __launch_bounds__(256) k_win_r##WS(const TemporalJob *__restrict__ jobs, int njobs) { \
        __shared__ float sT[WEMA_MAX_BATCH], sA[WEMA_MAX_BATCH];                                \
        win_reg_body<WS>(jobs, njobs, sT, sA);                                                  \
    }
WEMA_WIN_REG_KERNEL(16)
WEMA_WIN_REG_KERNEL(32)
WEMA_WIN_REG_KERNEL(64)

#define WEMA_WIN_LDS_KERNEL(WSMAX)                                                              \
    __global__ void __launch_bounds__(256) k_win_l##WSMAX(const TemporalJob *__restrict__ jobs, int njobs) { \
        __shared__ float W[WSMAX][256];                                                         \
        __shared__ float sT[WEMA_MAX_BATCH], sA[WEMA_MAX_BATCH], sTw[WSMAX];                    \
        win_lds_body<WSMAX>(jobs, njobs, W, sT, sA, sTw);                                       \
    }
WEMA_WIN_LDS_KERNEL(32)
WEMA_WIN_LDS_KERNEL(64)

// ---------------------------------------------------------------------------------------
// Guided filter (src/wema.c:114-230, 289-335; src/wema_batch.c:26-132, 311-346).
// Box means over the 9x9 window clipped to the frame. The reference divides each row's
// window sums by their count and then the column sums of those by theirs (a mean of means
// over FP32 prefix sums, src/wema.c:114-171). Here the window sum is the vertical 9-tap sum
// of horizontal 9-tap sums, and the mean is that sum times the correctly rounded reciprocal
// of the clipped window's area: the same mean in real arithmetic, with one multiply where
// the reference divides twice. The order of the additions (the oracle restates it exactly):
//   horizontal: blocks of 4 outputs starting at multiples of 4 of the image; the first sum
//     of a block adds its 9 samples left to right, each next one is
//     (previous - leaving sample) + entering sample;
//   vertical: rows in groups of 4 starting at multiples of 4; a group h0..h3 has the prefix
//     sums P1 = h0, P2 = P1 + h1, P3 = P2 + h2, F = P3 + h3 and the suffix sums U1 = h3,
//     U2 = h2 + h3, U3 = h1 + U2, U4 = F; the sum of row 4b + j (j = 0..3) is
//     (U_{b-1}(4 - j) + F_b) + P_{b+1}(j + 1), with P(4) = F.
// Samples outside the frame are 0 in both directions.

__device__ __forceinline__ float src_luma(const EgressJob &J, size_t idx) {
    if (J.src_u8) {
        const GLOBAL_AS unsigned char *p = gp(J.src_u8) + idx * 3;
        return luma_of(u8f(p[0]), u8f(p[1]), u8f(p[2]));
    }
    const GLOBAL_AS float *p = gp(J.src_f32) + idx * 3;
    return luma_of(p[0], p[1], p[2]);
}

// Clipped 9-tap window count at position x of n (src/wema.c:144-147).
__device__ __forceinline__ int clip_count(int x, int n) {
    const int lo = x - 4 > 0 ? x - 4 : 0;
    const int hi = x + 5 < n ? x + 5 : n;
    return hi > lo ? hi - lo : 1;
}

// a / b for b in [2^-7, 2) and 2^-66 <= |a| < 2^64 (or a = +0): the steps of the IEEE
// division sequence (reciprocal estimate, one refinement, two corrections with exact
// residuals) without its range scaling and special-case fix-up, which that range never
// needs; bit-identical to a / b there when the estimate is the correctly rounded 1/b, and
// from the neighbouring estimates too except at b = 2^k (2 - 2^-23), where 1/b lies just
// above a rounding tie (tests/div_mid_check.c on the CPU; the hardware estimate at those b
// is checked on the GPU); within an ulp of a / b for smaller nonzero |a|, whose residuals
// flush to zero, and +0 for a = -0 (a zero that only ever meets other terms).
__device__ __forceinline__ float div_mid(float a, float b) {
    float y = __builtin_amdgcn_rcpf(b);
    y = fmaf(fmaf(-b, y, 1.0f), y, y);
    float q = a * y;
    q = fmaf(fmaf(-b, q, a), y, q);
    return fmaf(fmaf(-b, q, a), y, q);
}

// One kernel streams the whole filter. A workgroup of 4 waves owns a strip of GF_OUT = 240
// output columns of one frame and a segment of its rows, and walks down the segment one
// group of 4 rows per step; a and b never leave the chip. Step m:
//   A  (row layout: wave w takes input row 4m + w, lane l the 4 columns x0 - 8 + 4l ..)
//      guide I = luma(src), p = rec - I, I^2 and I p; their horizontal sums, the
//      neighbouring lanes' samples taken by DPP lane shifts; valid in lanes 1-62.
//   B  (column layout: thread t takes column x0 - 8 + t of all 4 rows) the vertical sums,
//      which complete the window sums of row group m - 1: means, a and b (0 outside the
//      frame).
//   C  (row layout) horizontal sums of a and b, valid in lanes 2-61.
//   D  (column layout) their vertical sums: mean a and mean b of row group m - 2.
//   E  (row layout) q = mean(a) I + mean(b) and RGB out = clamp(src + q) of row group
//      m - 2, from the source pixels this lane loaded two steps earlier.
// The input rows of step m + 1 are loaded while step m computes. A segment starts 2 groups
// above its first output row (the filter reaches 8 rows up) and ends 2 groups below its last.
constexpr int GF_OUT = WEMA_GF_OUT;   // output columns of a strip
constexpr int GF_W = GF_OUT + 16;     // with the 8-column halo on each side: 64 lanes x 4

// Horizontal 9-tap sums of the 4 columns of this lane from its own samples o and those of
// the lanes on either side, in the block order above: v[0..3] = the previous lane's o,
// v[4..7] = o, v[8..11] = the next lane's o; the previous lane's ((o0 + o1) + o2) + o3 is
// the block's first partial sum.
__device__ __forceinline__ void hsum_lanes(const float (&o)[4], float (&s)[4]) {
    const float a = ((o[0] + o[1]) + o[2]) + o[3];
    float s0 = lane_prev(a) + o[0];
    s0 = ((s0 + o[1]) + o[2]) + o[3];
    s[0] = s0 + lane_next(o[0]);
    s[1] = (s[0] - lane_prev(o[0])) + lane_next(o[1]);
    s[2] = (s[1] - lane_prev(o[1])) + lane_next(o[2]);
    s[3] = (s[2] - lane_prev(o[2])) + lane_next(o[3]);
}

// Vertical window sums of one column of one quantity, fed one group of 4 rows at a time.
struct VSum {
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;   // U_{b-1}(4 - j) + F_b for the next group's outputs
    float u1 = 0.0f, u2 = 0.0f, u3 = 0.0f, f = 0.0f;    // suffix sums and sum of the last group
    // Takes group b + 1 (rows h0..h3); gives the window sums of the 4 rows of group b.
    __device__ __forceinline__ void step(float h0, float h1, float h2, float h3, float (&s)[4]) {
        const float p2 = h0 + h1, p3 = p2 + h2, F = p3 + h3;
        s[0] = t0 + h0;
        s[1] = t1 + p2;
        s[2] = t2 + p3;
        s[3] = t3 + F;
        t0 = f + F;
        t1 = u3 + F;
        t2 = u2 + F;
        t3 = u1 + F;
        u2 = h2 + h3;   // (the last uses of the previous group's sums are above)
        u3 = h1 + u2;
        u1 = h3;
        f = F;
    }
};

// The source pixels of 4 consecutive columns: 3 dwords of RGB24 or 12 floats.
template <bool U8>
struct Src4 {
    unsigned d[U8 ? 3 : 12];
    __device__ __forceinline__ float ch(int i) const {   // channel i (0..11) in [0, 1]
        return U8 ? u8f((d[i / 4] >> (8 * (i % 4))) & 0xFFu) : __uint_as_float(d[i]);
    }
};

// Input row y of the 4 columns from x: the source pixels and the reconstruction. VEC (width
// a multiple of 4, aligned buffers): the 4 columns lie wholly inside or outside the frame
// and load as 3 dwords (or 3 x 16 bytes) plus 16 bytes; coordinates are clamped into the
// frame so every load is valid (samples outside are zeroed by the caller).
// VEC: y is the same in every lane of the wave, and xo is the lane's clamped first column
// (loop-invariant): the row base is a scalar, the lane's part a 32-bit offset.
template <bool U8, bool VEC>
__device__ __forceinline__ void gf_load(const EgressJob &J, int x, int y, unsigned xo, Src4<U8> &s, v4f &g) {
    const int w = J.w, h = J.h;
    const int yc = min(max(y, 0), h - 1);
    if constexpr (VEC && U8) {
        const size_t row = (size_t)yc * (unsigned)w;
        const GLOBAL_AS unsigned char *rgb = gp(J.src_u8) + row * 3;
        const GLOBAL_AS unsigned *q = &at_b((const GLOBAL_AS unsigned *)rgb, 3u * xo);
        s.d[0] = q[0]; s.d[1] = q[1]; s.d[2] = q[2];
        g = at_b((const GLOBAL_AS v4f *)(gp(J.gray) + row), 4u * xo);
    } else if constexpr (VEC) {
        const int xc = min(max(x, 0), w - 4);
        const unsigned idx = (unsigned)yc * (unsigned)w + (unsigned)xc;
        const GLOBAL_AS v4f *q = (const GLOBAL_AS v4f *)(gp(J.src_f32) + (size_t)idx * 3);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const v4f f = q[k];
            s.d[4 * k] = __float_as_uint(f.x); s.d[4 * k + 1] = __float_as_uint(f.y);
            s.d[4 * k + 2] = __float_as_uint(f.z); s.d[4 * k + 3] = __float_as_uint(f.w);
        }
        g = *(const GLOBAL_AS v4f *)(gp(J.gray) + idx);
    } else {
        float gv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int xc = min(max(x + k, 0), w - 1);
            const size_t idx = (size_t)yc * w + xc;
            gv[k] = gp(J.gray)[idx];
            if (U8) {
                const GLOBAL_AS unsigned char *p = gp(J.src_u8) + idx * 3;
                const unsigned b0 = p[0], b1 = p[1], b2 = p[2];
                // channel 3k + c of the group sits in byte (3k + c) % 4 of dword (3k + c) / 4
                const unsigned v[3] = {b0, b1, b2};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int i = 3 * k + c;
                    if (i % 4 == 0) s.d[i / 4] = v[c];
                    else s.d[i / 4] |= v[c] << (8 * (i % 4));
                }
            } else {
                const GLOBAL_AS float *p = gp(J.src_f32) + idx * 3;
                s.d[3 * k] = __float_as_uint(p[0]);
                s.d[3 * k + 1] = __float_as_uint(p[1]);
                s.d[3 * k + 2] = __float_as_uint(p[2]);
            }
        }
        g = v4f{gv[0], gv[1], gv[2], gv[3]};
    }
}

// 1 / (area of the clipped window) at column count cx and row y of h: the correctly rounded
// reciprocal from the table (kx9: the entry of cx rows of 9, for rows whose window lies in
// the frame; the row test is the same for the whole workgroup).
__device__ __forceinline__ float inv_area(int cx, float kx9, int y, int h, const float *sRcp) {
    const int cy = clip_count(y, h);
    return cy == 9 ? kx9 : sRcp[cx * cy];
}

constexpr int GF_RCP = 82;   // correctly rounded 1/c for c = 0..81 (entry 0 unused)

// OUTF: float RGB output possible (otherwise RGB24 only).
template <bool U8, bool VEC, bool OUTF>
__device__ __forceinline__ void gf_strip(const EgressJob &J, int x0, int ybeg, int yend, float (*hs)[4][GF_W],
                                         float (*ab)[2][GF_W], const float *sRcp) {
    const int w = J.w, h = J.h;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), ln = threadIdx.x & 63;
    const int xr = x0 - 8 + 4 * ln;            // row layout: first of this lane's 4 columns
    const unsigned xo = (unsigned)min(max(xr, 0), w - 4);   // (VEC loads: clamped into the frame)
    const int xcol = x0 - 8 + (int)threadIdx.x;   // column layout: this thread's column
    const bool col_in = xcol >= 0 && xcol < w;
    const int cx = clip_count(xcol, w);
    const float kx9 = sRcp[cx * 9];
    const bool out_lane = ln >= 2 && ln < 62;
    VSum v1[4], v2[2];
    const int m0 = (ybeg >> 2) - 2, m1 = ((yend - 1) >> 2) + 2;
    Src4<U8> nxt, cur, old1, old2;
    v4f gn;
    gf_load<U8, VEC>(J, xr, 4 * m0 + wv, xo, nxt, gn);
    for (int m = m0; m <= m1; ++m) {
        cur = nxt;
        const v4f g = gn;
        if (m < m1) gf_load<U8, VEC>(J, xr, 4 * (m + 1) + wv, xo, nxt, gn);
        // A: the four quantities of input row 4m + wv and their horizontal sums
        {
            const int y = 4 * m + wv;
            bool in = (unsigned)y < (unsigned)h;
            if (VEC) in = in && xr >= 0 && xr < w;
            const float gv[4] = {g.x, g.y, g.z, g.w};
            float q[4][4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float l = luma_of(cur.ch(3 * k), cur.ch(3 * k + 1), cur.ch(3 * k + 2));
                const bool ink = VEC ? in : (in && (unsigned)(xr + k) < (unsigned)w);
                const float I = ink ? l : 0.0f, p = ink ? gv[k] - l : 0.0f;
                q[0][k] = I;
                q[1][k] = p;
                q[2][k] = I * I;
                q[3][k] = I * p;
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float s[4];
                hsum_lanes(q[c], s);
                *(v4f *)&hs[wv][c][4 * ln] = v4f{s[0], s[1], s[2], s[3]};
            }
        }
        __syncthreads();
        // B: vertical sums, means, a and b of row group m - 1
        {
            float hv[4][4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int c = 0; c < 4; ++c) hv[c][j] = hs[j][c][threadIdx.x];
            float S[4][4];
#pragma unroll
            for (int c = 0; c < 4; ++c) v1[c].step(hv[c][0], hv[c][1], hv[c][2], hv[c][3], S[c]);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int y = 4 * (m - 1) + j;
                const float k = inv_area(cx, kx9, y, h, sRcp);
                const float mI = S[0][j] * k, mp = S[1][j] * k, cI = S[2][j] * k, cIp = S[3][j] * k;
                const float var = fmaf(-mI, mI, cI);
                const float cov = fmaf(-mI, mp, cIp);
                // var + eps lies in [0.0099, 1.02): I is in [0, 1]
                const float av = div_mid(cov, var + 0.01f);
                const float bv = fmaf(-av, mI, mp);
                const bool in = col_in && (unsigned)y < (unsigned)h;
                ab[j][0][threadIdx.x] = in ? av : 0.0f;
                ab[j][1][threadIdx.x] = in ? bv : 0.0f;
            }
        }
        __syncthreads();
        // C: horizontal sums of a and b of row 4(m - 1) + wv, in place (each lane rewrites
        // only the 4 columns it read)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const v4f a = *(const v4f *)&ab[wv][c][4 * ln];
            const float o[4] = {a.x, a.y, a.z, a.w};
            float s[4];
            hsum_lanes(o, s);
            *(v4f *)&ab[wv][c][4 * ln] = v4f{s[0], s[1], s[2], s[3]};
        }
        __syncthreads();
        // D: vertical sums of a and b: their means at row group m - 2, in place
        {
            float hv[2][4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                hv[0][j] = ab[j][0][threadIdx.x];
                hv[1][j] = ab[j][1][threadIdx.x];
            }
            float S[2][4];
            v2[0].step(hv[0][0], hv[0][1], hv[0][2], hv[0][3], S[0]);
            v2[1].step(hv[1][0], hv[1][1], hv[1][2], hv[1][3], S[1]);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float k = inv_area(cx, kx9, 4 * (m - 2) + j, h, sRcp);
                ab[j][0][threadIdx.x] = S[0][j] * k;
                ab[j][1][threadIdx.x] = S[1][j] * k;
            }
        }
        __syncthreads();
        // E: output row 4(m - 2) + wv
        {
            const int y = 4 * (m - 2) + wv;
            if (y >= ybeg && y < yend && out_lane && xr < w) {
                const v4f ma = *(const v4f *)&ab[wv][0][4 * ln];
                const v4f mb = *(const v4f *)&ab[wv][1][4 * ln];
                const float mav[4] = {ma.x, ma.y, ma.z, ma.w}, mbv[4] = {mb.x, mb.y, mb.z, mb.w};
                // clamp to [0, 1]: as clamp01 for the float output (which keeps a -0); the
                // bytes take the median of (v, 0, 1), one instruction, with the same
                // egress (fmaf(-0, 255, 0.5) = fmaf(+0, 255, 0.5))
                float o[12];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float c0 = old2.ch(3 * k), c1 = old2.ch(3 * k + 1), c2 = old2.ch(3 * k + 2);
                    const float q = fmaf(mav[k], luma_of(c0, c1, c2), mbv[k]);
                    o[3 * k] = c0 + q;
                    o[3 * k + 1] = c1 + q;
                    o[3 * k + 2] = c2 + q;
                }
                const size_t base = ((size_t)y * w + xr) * 3;
                if (!OUTF || J.out_u8) {
                    if (VEC) {
                        unsigned d[3] = {0u, 0u, 0u};
#pragma unroll
                        for (int i = 0; i < 12; ++i)
                            d[i / 4] |= (unsigned)(unsigned char)fmaf(__builtin_amdgcn_fmed3f(o[i], 0.0f, 1.0f), 255.0f, 0.5f)
                                        << (8 * (i % 4));
                        // (the row is the wave's; out_lane columns are inside the frame)
                        GLOBAL_AS unsigned *os = &at_b((GLOBAL_AS unsigned *)(gp(J.out_u8) + (size_t)y * (unsigned)w * 3),
                                                       3u * (unsigned)xr);
                        os[0] = d[0];
                        os[1] = d[1];
                        os[2] = d[2];
                    } else {
#pragma unroll
                        for (int i = 0; i < 12; ++i)
                            if (xr + i / 3 < w)
                                gp(J.out_u8)[base + i] = (unsigned char)fmaf(__builtin_amdgcn_fmed3f(o[i], 0.0f, 1.0f), 255.0f, 0.5f);
                    }
                }
                if (OUTF && J.out_f32) {
                    if (VEC) {
                        GLOBAL_AS v4f *os = (GLOBAL_AS v4f *)(gp(J.out_f32) + base);
                        os[0] = v4f{clamp01(o[0]), clamp01(o[1]), clamp01(o[2]), clamp01(o[3])};
                        os[1] = v4f{clamp01(o[4]), clamp01(o[5]), clamp01(o[6]), clamp01(o[7])};
                        os[2] = v4f{clamp01(o[8]), clamp01(o[9]), clamp01(o[10]), clamp01(o[11])};
                    } else {
#pragma unroll
                        for (int i = 0; i < 12; ++i)
                            if (xr + i / 3 < w) gp(J.out_f32)[base + i] = clamp01(o[i]);
                    }
                }
            }
        }
        old2 = old1;
        old1 = cur;
    }
}

// The workgroup's job, strip and segment.
struct GfTile {
    const EgressJob *J;
    int x0, ybeg, yend;
};
__device__ __forceinline__ GfTile gf_tile(const EgressJob *__restrict__ jobs, int njobs, float *sRcp) {
    if (threadIdx.x < GF_RCP) sRcp[threadIdx.x] = threadIdx.x ? 1.0f / (float)threadIdx.x : 0.0f;
    const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const EgressJob &J = jobs[find_job(jobs, njobs, bid)];
    const int t = bid - J.tile_begin;
    const int ybeg = (t / J.tiles_x) * J.seg;
    __syncthreads();
    return GfTile{&J, (t % J.tiles_x) * GF_OUT, ybeg, min(ybeg + J.seg, J.h)};
}

#define WEMA_GF_SHARED                                              \
    __shared__ __attribute__((aligned(16))) float hs[4][4][GF_W];   \
    __shared__ __attribute__((aligned(16))) float ab[4][2][GF_W];   \
    __shared__ float sRcp[GF_RCP];

// RGB24 in and out, width a multiple of 4, frames and planes aligned (the host checks every
// job of the launch): the pipeline's path, with the fewest registers.
__global__ void __launch_bounds__(256) k_gf(const EgressJob *__restrict__ jobs, int njobs) {
    WEMA_GF_SHARED
    const GfTile T = gf_tile(jobs, njobs, sRcp);
    gf_strip<true, true, false>(*T.J, T.x0, T.ybeg, T.yend, hs, ab, sRcp);
}

// Any job: float RGB in or out, any width.
__global__ void __launch_bounds__(256) k_gf_any(const EgressJob *__restrict__ jobs, int njobs) {
    WEMA_GF_SHARED
    const GfTile T = gf_tile(jobs, njobs, sRcp);
    const EgressJob &J = *T.J;
    const uintptr_t al = (uintptr_t)J.gray | (J.src_u8 ? 0 : (uintptr_t)J.src_f32) | (J.out_f32 ? (uintptr_t)J.out_f32 : 0);
    const bool vec = (J.w & 3) == 0 && (al & 15) == 0 && (((uintptr_t)J.src_u8 | (uintptr_t)J.out_u8) & 3) == 0;
    if (J.src_u8) {
        if (vec) gf_strip<true, true, true>(J, T.x0, T.ybeg, T.yend, hs, ab, sRcp);
        else gf_strip<true, false, true>(J, T.x0, T.ybeg, T.yend, hs, ab, sRcp);
    } else {
        if (vec) gf_strip<false, true, true>(J, T.x0, T.ybeg, T.yend, hs, ab, sRcp);
        else gf_strip<false, false, true>(J, T.x0, T.ybeg, T.yend, hs, ab, sRcp);
    }
}

// RGB out = clamp(source + q) for the 5x5 path (src/wema.c:329-334).
__device__ __forceinline__ void write_rgb(const EgressJob &J, size_t idx, float q) {
    float s[3];
    if (J.src_u8) {
        const GLOBAL_AS unsigned char *p = gp(J.src_u8) + idx * 3;
        s[0] = u8f(p[0]); s[1] = u8f(p[1]); s[2] = u8f(p[2]);
    } else {
        const GLOBAL_AS float *p = gp(J.src_f32) + idx * 3;
        s[0] = p[0]; s[1] = p[1]; s[2] = p[2];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float v = clamp01(s[k] + q);
        if (J.out_u8) gp(J.out_u8)[idx * 3 + k] = egress_u8(v);
        if (J.out_f32) gp(J.out_f32)[idx * 3 + k] = v;
    }
}

// 5x5 binomial blur of the luma delta with weights renormalised at the borders
// (src/wema.c:236-283), then RGB transfer.
__global__ void __launch_bounds__(256) k_gauss(const EgressJob *__restrict__ jobs, int njobs) {
    constexpr int EX = WEMA_GF_TX + 4, EY = WEMA_GF_TY + 4;
    __shared__ float sP[EY][EX + 1], sT[EY][WEMA_GF_TX];
    const float kw[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};
    const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int jid = find_job(jobs, njobs, bid);
    const EgressJob &J = jobs[jid];
    const int t = bid - J.tile_begin;
    const int x0 = (t % J.tiles_x) * WEMA_GF_TX, y0 = (t / J.tiles_x) * WEMA_GF_TY;
    const int w = J.w, h = J.h;
    constexpr int NLD = (EY * EX + 255) / 256;
    float lp[NLD];
#pragma unroll
    for (int j = 0; j < NLD; ++j) {
        const int i = threadIdx.x + 256 * j;
        const int ey = i / EX, ex = i % EX;
        const int x = x0 - 2 + ex, y = y0 - 2 + ey;
        float p = 0.0f;
        if (i < EY * EX && x >= 0 && x < w && y >= 0 && y < h) {
            const size_t idx = (size_t)y * w + x;
            p = gp(J.gray)[idx] - src_luma(J, idx);
        }
        lp[j] = p;
    }
#pragma unroll
    for (int j = 0; j < NLD; ++j) {
        const int i = threadIdx.x + 256 * j;
        if (i < EY * EX) sP[i / EX][i % EX] = lp[j];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < EY * WEMA_GF_TX; i += 256) {
        const int ey = i / WEMA_GF_TX, x = i % WEMA_GF_TX;
        const int gx = x0 + x;
        float s = 0.0f, wt = 0.0f;
#pragma unroll
        for (int d = 0; d < 5; ++d) {
            const int xx = gx + d - 2;
            if (xx >= 0 && xx < w) {
                s = fmaf(kw[d], sP[ey][x + d], s);
                wt += kw[d];
            }
        }
        sT[ey][x] = wt > 0.0f ? s / wt : 0.0f;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < WEMA_GF_TY * WEMA_GF_TX; i += 256) {
        const int y = i / WEMA_GF_TX, x = i % WEMA_GF_TX;
        const int gx = x0 + x, gy = y0 + y;
        if (gx >= w || gy >= h) continue;
        float s = 0.0f, wt = 0.0f;
#pragma unroll
        for (int d = 0; d < 5; ++d) {
            const int yy = gy + d - 2;
            if (yy >= 0 && yy < h) {
                s = fmaf(kw[d], sT[y + d][x], s);
                wt += kw[d];
            }
        }
        write_rgb(J, (size_t)gy * w + gx, s / wt);
    }
}

// Window mode before the window is full: the input frame is returned (src/wema.c:924-929).
__global__ void __launch_bounds__(256) k_passthru(const EgressJob *__restrict__ jobs, int njobs) {
    const int jid = find_job(jobs, njobs, (int)blockIdx.x);
    const EgressJob &J = jobs[jid];
    const size_t i = (size_t)((int)blockIdx.x - J.tile_begin) * 256 + threadIdx.x;
    if (i >= (size_t)J.w * J.h * 3) return;
    const float v = J.src_u8 ? u8f(J.src_u8[i]) : J.src_f32[i];
    if (J.out_u8) J.out_u8[i] = egress_u8(v);
    if (J.out_f32) J.out_f32[i] = v;
}

// Test hook: the guided filter's covariance division and the hardware reciprocal estimate
// it starts from, element by element.
__global__ void __launch_bounds__(256) k_debug_div(const float *__restrict__ a, const float *__restrict__ b,
                                                   float *__restrict__ q, float *__restrict__ rcp, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    q[i] = div_mid(a[i], b[i]);
    rcp[i] = __builtin_amdgcn_rcpf(b[i]);
}

// Clears n 64-bit words (threshold accumulators) with vector stores: a kernel node, which
// graph capture records as it records the pipeline's own launches.
__global__ void __launch_bounds__(256) k_zero_u64(unsigned long long *__restrict__ p, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---------------------------------------------------------------------------------------
// Synthetic frames: one thread per pixel.
__global__ void __launch_bounds__(256) k_synth(const SynthFrame *__restrict__ frames,
                                               const unsigned short *__restrict__ tex,
                                               unsigned char *__restrict__ out, int n) {
    const SynthFrame &F = frames[blockIdx.y];
    const size_t P = (size_t)F.width * F.height;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= P || (int)blockIdx.y >= n) return;
    const int x = (int)(i % F.width), y = (int)(i / F.width);
    const int tq = tex ? (int)tex[i] : 0;
    unsigned char *o = out + (size_t)blockIdx.y * P * 3 + i * 3;
    o[0] = synth_pixel(F, tq, x, y, 0);
    o[1] = F.color ? synth_pixel(F, tq, x, y, 1) : o[0];
    o[2] = F.color ? synth_pixel(F, tq, x, y, 2) : o[0];
}

// ---------------------------------------------------------------------------------------
// launchers (called by the host scheduler)

extern "C" {
hipError_t wema_launch_dwt_fwd(const DwtfJob *jobs, int njobs, int nblocks, int level0, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (level0) hipLaunchKernelGGL(k_dwt_fwd0, dim3(nblocks), dim3(64 * WEMA_STRIPS_PER_BLOCK), 0, s, jobs, njobs);
    else hipLaunchKernelGGL(k_dwt_fwd, dim3(nblocks), dim3(64 * WEMA_STRIPS_PER_BLOCK), 0, s, jobs, njobs);
    return hipGetLastError();
}
hipError_t wema_launch_dwt_inv(const DwtiJob *jobs, int njobs, int nblocks, int level0, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (level0) hipLaunchKernelGGL(k_dwt_inv0, dim3(nblocks), dim3(64 * WEMA_STRIPS_PER_BLOCK), 0, s, jobs, njobs);
    else hipLaunchKernelGGL(k_dwt_inv, dim3(nblocks), dim3(64 * WEMA_STRIPS_PER_BLOCK), 0, s, jobs, njobs);
    return hipGetLastError();
}
hipError_t wema_launch_dwt_fwd_tile(const DwtfJob *jobs, int njobs, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dwt_fwd_tile, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}
hipError_t wema_launch_dwt_inv_tile(const DwtiJob *jobs, int njobs, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dwt_inv_tile, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}
hipError_t wema_launch_dwt_fwd_point(const DwtfJob *jobs, int njobs, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dwt_fwd_point, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}
hipError_t wema_launch_dwt_inv_point(const DwtiJob *jobs, int njobs, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dwt_inv_point, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}
hipError_t wema_launch_dwt_inv_ycc(const DwtiJob *jobs, int njobs, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dwt_inv0_ycc, dim3(nblocks), dim3(64 * WEMA_STRIPS_PER_BLOCK), 0, s, jobs, njobs);
    return hipGetLastError();
}
// one_frame: every job of the launch is of one frame with WEMA_IIR_PPT positions per thread
hipError_t wema_launch_iir(const TemporalJob *jobs, int njobs, int nblocks, int one_frame, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (one_frame) hipLaunchKernelGGL(k_iir1, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    else hipLaunchKernelGGL(k_iir, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}
// variant: 0 = windows up to 32 samples (LDS), 1 = up to 64 (LDS), 2 = any (ring reads)
hipError_t wema_launch_win(int variant, const TemporalJob *jobs, int njobs, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (variant == 0) hipLaunchKernelGGL(k_win_l32, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    else if (variant == 1) hipLaunchKernelGGL(k_win_l64, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    else if (variant == 3) hipLaunchKernelGGL(k_win_r32, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    else if (variant == 4) hipLaunchKernelGGL(k_win_r16, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    else if (variant == 5) hipLaunchKernelGGL(k_win_r64, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    else hipLaunchKernelGGL(k_win, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}
hipError_t wema_launch_egress(int kind, const EgressJob *jobs, int njobs, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    switch (kind) {
    case EG_GF: hipLaunchKernelGGL(k_gf, dim3(nblocks), dim3(256), 0, s, jobs, njobs); break;
    case EG_GF_ANY: hipLaunchKernelGGL(k_gf_any, dim3(nblocks), dim3(256), 0, s, jobs, njobs); break;
    case EG_GAUSS: hipLaunchKernelGGL(k_gauss, dim3(nblocks), dim3(256), 0, s, jobs, njobs); break;
    default: hipLaunchKernelGGL(k_passthru, dim3(nblocks), dim3(256), 0, s, jobs, njobs); break;
    }
    return hipGetLastError();
}
hipError_t wema_launch_debug_div(const float *a, const float *b, float *q, float *rcp, long long n,
                                 hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_debug_div, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, q, rcp, n);
    return hipGetLastError();
}
hipError_t wema_launch_zero_u64(unsigned long long *p, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
hipError_t wema_launch_synth(const SynthFrame *frames, const unsigned short *tex,
                             unsigned char *out, int width, int height, int n, hipStream_t s) {
    const size_t P = (size_t)width * height;
    dim3 grid((unsigned)((P + 255) / 256), (unsigned)n);
    hipLaunchKernelGGL(k_synth, grid, dim3(256), 0, s, frames, tex, out, n);
    return hipGetLastError();
}
}
