/*
 * CPU oracle for the WEMA per-frame amplification pipeline (see wema_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker; never called by the product.
 *
 * Built with -ffp-contract=off so that every fused multiply-add below is one written as
 * fmaf(): those are the contractions the reference's -Ofast build performs (checked against
 * the compiled reference by tests/test_oracle_ref.py).
 *
 * Pinning (what checks each part besides the compiled reference):
 *   lifting DWT     published CDF 9/7 analysis taps, brute-force symmetric-extension
 *                   convolution in FP64, perfect reconstruction      (tests/test_oracle_math.py)
 *   IIR design/run  scipy.signal.butter / sosfilt                    (tests/test_oracle_math.py)
 *   box / guided    brute-force O(r^2) FP64 windows                  (tests/test_oracle_math.py)
 *   colour, u8      exhaustive u8 tables against the compiled reference
 *   window/bilateral direct FP64 per-position formula
 */
#include "wema_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>

/* Floating-point mode. The reference runs with denormals flushed (FTZ) and treated as zero
 * (DAZ): its -Ofast build links crtfastmath, which sets both at start-up; the kernels flush
 * denormals too. Every FP32 entry point below (and the DWT, whose FP32 and FP64 forms share
 * one body; FP64 values here never come near its denormal range) sets both for its own
 * duration and restores
 * the caller's mode on return, so results do not depend on the mode of the calling thread
 * (for instance on whether the compiled reference was loaded into the process first). */
#define MXCSR_FTZ_DAZ 0x8040u
static unsigned fp_enter(void) {
    const unsigned m = _mm_getcsr();
    _mm_setcsr(m | MXCSR_FTZ_DAZ);
    return m;
}
static void fp_restore(unsigned *m) { _mm_setcsr(*m); }
#define FP_REFERENCE_MODE() unsigned fpm_ __attribute__((cleanup(fp_restore))) = fp_enter()

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_SQRT2
#define M_SQRT2 1.41421356237309504880
#endif

/* CDF 9/7 lifting factors and the reference's normalisation (src/dwt.c:31-36). */
#define ALPHA_F (-1.586134342059924f)
#define BETA_F  (-0.052980118572961f)
#define GAMMA_F (0.882911075530934f)
#define DELTA_F (0.443506852043971f)
#define K_F     (1.149604398860241f)
#define KINV_F  (1.0f / K_F)
#define ALPHA_D (-1.586134342059924)
#define BETA_D  (-0.052980118572961)
#define GAMMA_D (0.882911075530934)
#define DELTA_D (0.443506852043971)
#define K_D     (1.149604398860241)

/*============================================================================
 * Geometry (src/dwt.c:598-632, 769-778)
 *===========================================================================*/

int orc_level_dims(int w, int h, int levels, int *lw, int *lh) {
    lw[0] = w;
    lh[0] = h;
    for (int l = 0; l < levels; l++) {
        lw[l + 1] = (lw[l] + 1) / 2;
        lh[l + 1] = (lh[l] + 1) / 2;
    }
    return 0;
}

size_t orc_num_unique(int w, int h, int levels) {
    int lw[32], lh[32];
    orc_level_dims(w, h, levels, lw, lh);
    size_t s = 0;
    for (int l = 1; l <= levels; l++) s += (size_t)lw[l] * lh[l];
    return 3 * s;
}

/*============================================================================
 * 1D lifting, FP32 in reference order (src/dwt.c:45-107, 109-170)
 *===========================================================================*/

#define LIFT_PREDICT(T, C, e, o, half, odd_n, FMA)                                   \
    do {                                                                             \
        for (int i = 0; i < (odd_n) - 1; i++) o[i] = FMA(C, e[i] + e[i + 1], o[i]);  \
        o[(odd_n) - 1] = FMA(C, e[(odd_n) - 1] + e[(half) > (odd_n) ? (odd_n) : (odd_n) - 1], \
                             o[(odd_n) - 1]);                                        \
    } while (0)

#define LIFT_UPDATE(T, C, e, o, half, odd_n, FMA)                                    \
    do {                                                                             \
        e[0] = FMA(C, o[0] + o[0], e[0]);                                            \
        for (int i = 1; i < (half); i++) {                                           \
            int r_ = (i < (odd_n)) ? i : (odd_n) - 1;                                \
            e[i] = FMA(C, o[i - 1] + o[r_], e[i]);                                   \
        }                                                                            \
    } while (0)

#define FMA_F(a, b, c) fmaf((a), (b), (c))
#define FMA_D(a, b, c) ((a) * (b) + (c))

void orc_fwd1d_f32(float *x, int n) {
    FP_REFERENCE_MODE();
    if (n < 2) return;
    const int half = (n + 1) / 2, odd_n = n / 2;
    float *e = malloc(sizeof(float) * (size_t)n), *o = e + half;
    for (int i = 0; i < odd_n; i++) { e[i] = x[2 * i]; o[i] = x[2 * i + 1]; }
    if (half > odd_n) e[odd_n] = x[2 * odd_n];
    LIFT_PREDICT(float, ALPHA_F, e, o, half, odd_n, FMA_F);
    LIFT_UPDATE(float, BETA_F, e, o, half, odd_n, FMA_F);
    LIFT_PREDICT(float, GAMMA_F, e, o, half, odd_n, FMA_F);
    LIFT_UPDATE(float, DELTA_F, e, o, half, odd_n, FMA_F);
    for (int i = 0; i < half; i++) x[i] = e[i] * KINV_F;
    for (int i = 0; i < odd_n; i++) x[half + i] = o[i] * K_F;
    free(e);
}

void orc_inv1d_f32(float *x, int n) {
    FP_REFERENCE_MODE();
    if (n < 2) return;
    const int half = (n + 1) / 2, odd_n = n / 2;
    float *e = malloc(sizeof(float) * (size_t)n), *o = e + half;
    for (int i = 0; i < half; i++) e[i] = x[i] * K_F;
    for (int i = 0; i < odd_n; i++) o[i] = x[half + i] * KINV_F;
    LIFT_UPDATE(float, -DELTA_F, e, o, half, odd_n, FMA_F);
    LIFT_PREDICT(float, -GAMMA_F, e, o, half, odd_n, FMA_F);
    LIFT_UPDATE(float, -BETA_F, e, o, half, odd_n, FMA_F);
    LIFT_PREDICT(float, -ALPHA_F, e, o, half, odd_n, FMA_F);
    for (int i = 0; i < odd_n; i++) { x[2 * i] = e[i]; x[2 * i + 1] = o[i]; }
    if (half > odd_n) x[2 * odd_n] = e[odd_n];
    free(e);
}

void orc_fwd1d_f64(double *x, int n) {
    if (n < 2) return;
    const int half = (n + 1) / 2, odd_n = n / 2;
    double *e = malloc(sizeof(double) * (size_t)n), *o = e + half;
    for (int i = 0; i < odd_n; i++) { e[i] = x[2 * i]; o[i] = x[2 * i + 1]; }
    if (half > odd_n) e[odd_n] = x[2 * odd_n];
    LIFT_PREDICT(double, ALPHA_D, e, o, half, odd_n, FMA_D);
    LIFT_UPDATE(double, BETA_D, e, o, half, odd_n, FMA_D);
    LIFT_PREDICT(double, GAMMA_D, e, o, half, odd_n, FMA_D);
    LIFT_UPDATE(double, DELTA_D, e, o, half, odd_n, FMA_D);
    for (int i = 0; i < half; i++) x[i] = e[i] / K_D;
    for (int i = 0; i < odd_n; i++) x[half + i] = o[i] * K_D;
    free(e);
}

void orc_inv1d_f64(double *x, int n) {
    if (n < 2) return;
    const int half = (n + 1) / 2, odd_n = n / 2;
    double *e = malloc(sizeof(double) * (size_t)n), *o = e + half;
    for (int i = 0; i < half; i++) e[i] = x[i] * K_D;
    for (int i = 0; i < odd_n; i++) o[i] = x[half + i] / K_D;
    LIFT_UPDATE(double, -DELTA_D, e, o, half, odd_n, FMA_D);
    LIFT_PREDICT(double, -GAMMA_D, e, o, half, odd_n, FMA_D);
    LIFT_UPDATE(double, -BETA_D, e, o, half, odd_n, FMA_D);
    LIFT_PREDICT(double, -ALPHA_D, e, o, half, odd_n, FMA_D);
    for (int i = 0; i < odd_n; i++) { x[2 * i] = e[i]; x[2 * i + 1] = o[i]; }
    if (half > odd_n) x[2 * odd_n] = e[odd_n];
    free(e);
}

/*============================================================================
 * 2D pyramid (src/dwt.c:369-473, 475-582, 645-767)
 *
 * The forward transform keeps the reference's extraction scratch buffers (LH, HL, HH, one
 * per orientation, reused by every level, src/dwt.c:648-653): where a level's parent
 * dimension is odd, the last row (LH, HH) or column (HL, HH) of that level's subband is not
 * written by the extraction guards (src/dwt.c:463-469) and keeps what an earlier level left
 * at the same flat index. Those values enter the threshold sum (src/wema.c:719-721). At
 * level 0 the reference reads fresh malloc memory there; the oracle (and the GPU) use 0.
 *===========================================================================*/

#define DEF_DWT2D(SUF, T, FWD1, INV1)                                                      \
static void fwd2d_##SUF(T *img, int w, int h, T *col) {                                    \
    for (int y = 0; y < h; y++) FWD1(img + (size_t)y * w, w);                              \
    for (int x = 0; x < w; x++) {                                                          \
        for (int y = 0; y < h; y++) col[y] = img[(size_t)y * w + x];                       \
        FWD1(col, h);                                                                      \
        for (int y = 0; y < h; y++) img[(size_t)y * w + x] = col[y];                       \
    }                                                                                      \
}                                                                                          \
static void inv2d_##SUF(T *img, int w, int h, T *col) {                                    \
    for (int x = 0; x < w; x++) {                                                          \
        for (int y = 0; y < h; y++) col[y] = img[(size_t)y * w + x];                       \
        INV1(col, h);                                                                      \
        for (int y = 0; y < h; y++) img[(size_t)y * w + x] = col[y];                       \
    }                                                                                      \
    for (int y = 0; y < h; y++) INV1(img + (size_t)y * w, w);                              \
}                                                                                          \
void orc_dwt_fwd_##SUF(const T *image, int W, int H, int levels, T *planes, T *ll_out) {   \
    FP_REFERENCE_MODE();                                                                   \
    size_t P = (size_t)W * H;                                                              \
    T *work = malloc(P * sizeof(T)), *col = malloc((size_t)(W > H ? W : H) * sizeof(T));   \
    T *sb[3];                                                                              \
    for (int s = 0; s < 3; s++) sb[s] = calloc(P, sizeof(T));                              \
    memcpy(work, image, P * sizeof(T));                                                    \
    int w = W, h = H;                                                                      \
    size_t off = 0;                                                                        \
    for (int lev = 0; lev < levels; lev++) {                                               \
        int sw = (w + 1) / 2, sh = (h + 1) / 2;                                            \
        fwd2d_##SUF(work, w, h, col);                                                      \
        /* extraction with the reference's guards; sb = {LH, HH, HL} */                    \
        for (int y = 0; y < sh; y++)                                                       \
            for (int x = 0; x < sw; x++) {                                                 \
                if (x < w - sw) sb[2][(size_t)y * sw + x] = work[(size_t)y * w + sw + x];  \
                if (y < h - sh) sb[0][(size_t)y * sw + x] = work[(size_t)(sh + y) * w + x]; \
                if (x < w - sw && y < h - sh)                                              \
                    sb[1][(size_t)y * sw + x] = work[(size_t)(sh + y) * w + sw + x];       \
            }                                                                              \
        size_t n = (size_t)sw * sh;                                                        \
        for (int s = 0; s < 3; s++) memcpy(planes + off + s * n, sb[s], n * sizeof(T));    \
        off += 3 * n;                                                                      \
        for (int y = 0; y < sh; y++)                                                       \
            memmove(work + (size_t)y * sw, work + (size_t)y * w, (size_t)sw * sizeof(T));  \
        w = sw;                                                                            \
        h = sh;                                                                            \
    }                                                                                      \
    memcpy(ll_out, work, (size_t)w * h * sizeof(T));                                       \
    for (int s = 0; s < 3; s++) free(sb[s]);                                               \
    free(work);                                                                            \
    free(col);                                                                             \
}                                                                                          \
void orc_dwt_inv_##SUF(const T *planes, const T *ll, int W, int H, int levels, T *img) {   \
    FP_REFERENCE_MODE();                                                                   \
    int lw[32], lh[32];                                                                    \
    orc_level_dims(W, H, levels, lw, lh);                                                  \
    size_t P = (size_t)W * H, off[33];                                                     \
    off[0] = 0;                                                                            \
    for (int l = 0; l < levels; l++) off[l + 1] = off[l] + 3 * (size_t)lw[l + 1] * lh[l + 1]; \
    T *work = malloc(P * sizeof(T)), *out = malloc(P * sizeof(T));                         \
    T *col = malloc((size_t)(W > H ? W : H) * sizeof(T));                                  \
    memcpy(work, ll, (size_t)lw[levels] * lh[levels] * sizeof(T));                         \
    for (int lev = levels - 1; lev >= 0; lev--) {                                          \
        int w = lw[lev], h = lh[lev], sw = lw[lev + 1], sh = lh[lev + 1];                  \
        size_t n = (size_t)sw * sh;                                                        \
        const T *LH = planes + off[lev], *HH = LH + n, *HL = HH + n;                       \
        for (int y = 0; y < h; y++)                                                        \
            for (int x = 0; x < w; x++) {                                                  \
                T v;                                                                       \
                if (y < sh && x < sw) v = work[(size_t)y * sw + x];                        \
                else if (y < sh) v = HL[(size_t)y * sw + (x - sw)];                        \
                else if (x < sw) v = LH[(size_t)(y - sh) * sw + x];                        \
                else v = HH[(size_t)(y - sh) * sw + (x - sw)];                             \
                out[(size_t)y * w + x] = v;                                                \
            }                                                                              \
        inv2d_##SUF(out, w, h, col);                                                       \
        memcpy(work, out, (size_t)w * h * sizeof(T));                                      \
    }                                                                                      \
    memcpy(img, work, P * sizeof(T));                                                      \
    free(work);                                                                            \
    free(out);                                                                             \
    free(col);                                                                             \
}

DEF_DWT2D(f32, float, orc_fwd1d_f32, orc_inv1d_f32)
DEF_DWT2D(f64, double, orc_fwd1d_f64, orc_inv1d_f64)

/* Reference flat order (src/dwt.c:680-686, src/wema.c:698-708): per level
 * slots [LH, HH, HL, HL, HH, LH]. */
void orc_planes_to_flat6(const float *planes, int w, int h, int levels, float *flat6) {
    int lw[32], lh[32];
    orc_level_dims(w, h, levels, lw, lh);
    static const int slot_src[6] = {0, 1, 2, 2, 1, 0};
    size_t off = 0, pos = 0;
    for (int l = 0; l < levels; l++) {
        size_t n = (size_t)lw[l + 1] * lh[l + 1];
        for (int s = 0; s < 6; s++) {
            memcpy(flat6 + pos, planes + off + slot_src[s] * n, n * sizeof(float));
            pos += n;
        }
        off += 3 * n;
    }
}

/*============================================================================
 * Ingest / colour (src/ffmpeg_io.c:292-294; src/wema.c:42-51, 60-102)
 *===========================================================================*/

static inline float u8f(uint8_t v) { return (float)v * (1.0f / 255.0f); }

/* Contraction shapes of the reference's vectorised loops (src/wema.c:42-51, 60-78, built
 * with -Ofast: gcc reassociates cb's "+0.5" into the g term and cr's into (r + 1) * 0.5). */
static inline float luma_f32(float r, float g, float b) {
    return fmaf(0.114f, b, fmaf(0.299f, r, 0.587f * g));
}
static inline float cb_f32(float r, float g, float b) {
    return fmaf(-0.169f, r, 0.5f * b) + fmaf(-0.331f, g, 0.5f);
}
static inline float cr_f32(float r, float g, float b) {
    return fmaf(-0.081f, b, fmaf(-0.419f, g, (r + 1.0f) * 0.5f));
}

void orc_ingest_f32(const uint8_t *rgb, size_t npx, int channel, float *out) {
    FP_REFERENCE_MODE();
    for (size_t i = 0; i < npx; i++) {
        float r = u8f(rgb[3 * i]), g = u8f(rgb[3 * i + 1]), b = u8f(rgb[3 * i + 2]);
        out[i] = channel == 0 ? luma_f32(r, g, b) : channel == 1 ? cb_f32(r, g, b) : cr_f32(r, g, b);
    }
}

static inline float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }
static inline uint8_t egress_u8(float v) { return (uint8_t)fmaf(clamp01(v), 255.0f, 0.5f); }

/*============================================================================
 * IIR (src/iir_filter.c:36-154, 174-471)
 *===========================================================================*/

/* Butterworth sections as the reference's -Ofast build evaluates them
 * (src/iir_filter.c:36-63 inlined into iir_design_bandpass, :70-101): gcc contracts the
 * high-pass section's wc*wc-1, wc*wc+sqrt2*wc and wc*wc+1 into fmas, and the low-pass
 * section's sqrt2*wc+wc2 and wc2+1-sqrt2*wc. The warm-up count (:150-151) is evaluated as
 * fs * fl(7/(2*pi_f)) / f_low. */
static void design_hp(float *c, float fc, float fs) {
    const float sqrt2 = 1.41421356237f;
    float wc = tanf(((float)M_PI * fc) / fs);
    float s2w = sqrt2 * wc;
    float norm = 1.0f / (fmaf(wc, wc, s2w) + 1.0f);
    c[0] = norm;
    c[1] = -2.0f * norm;
    c[2] = norm;
    float wm1 = fmaf(wc, wc, -1.0f);
    c[3] = (wm1 + wm1) * norm;
    c[4] = (fmaf(wc, wc, 1.0f) - s2w) * norm;
}

static void design_lp(float *c, float fc, float fs) {
    const float sqrt2 = 1.41421356237f;
    float wc = tanf(((float)M_PI * fc) / fs);
    float wc2 = wc * wc;
    float norm = 1.0f / (fmaf(sqrt2, wc, wc2) + 1.0f);
    c[0] = wc2 * norm;
    c[1] = c[0] + c[0];
    c[2] = c[0];
    float wm1 = wc2 - 1.0f;
    c[3] = (wm1 + wm1) * norm;
    c[4] = fmaf(-wc, sqrt2, wc2 + 1.0f) * norm;
}

int orc_iir_design(float f_low, float f_high, float fs, float *c10, int *warmup) {
    FP_REFERENCE_MODE();
    if (f_low >= f_high || f_low <= 0 || f_high >= fs * 0.5f) return -1;
    design_hp(c10, f_low, fs);
    design_lp(c10 + 5, f_high, fs);
    if (warmup) {
        const float k = 7.0f / (2.0f * (float)M_PI);
        int wu = (int)((fs * k) / f_low);
        *warmup = wu < 30 ? 30 : wu;
    }
    return 0;
}

int orc_iir_design_f64(double f_low, double f_high, double fs, double *c) {
    if (f_low >= f_high || f_low <= 0 || f_high >= fs / 2) return -1;
    for (int s = 0; s < 2; s++) {
        double wc = tan(M_PI * (s == 0 ? f_low : f_high) / fs), wc2 = wc * wc;
        double norm = 1.0 / (1.0 + M_SQRT2 * wc + wc2);
        double *q = c + 5 * s;
        if (s == 0) { q[0] = norm; q[1] = -2.0 * norm; q[2] = norm; }
        else { q[0] = wc2 * norm; q[1] = 2.0 * wc2 * norm; q[2] = wc2 * norm; }
        q[3] = 2.0 * (wc2 - 1.0) * norm;
        q[4] = (1.0 - M_SQRT2 * wc + wc2) * norm;
    }
    return 0;
}

/* DF2T with the reference's FMA sequence (src/iir_filter.c:211-219). st: [4] per position
 * (HP w1, HP w2, LP w1, LP w2). */
static inline float biquad2_f32(const float *c, float *st, float x) {
    for (int s = 0; s < 2; s++) {
        const float *q = c + 5 * s;
        float w1 = st[2 * s], w2 = st[2 * s + 1];
        float y = fmaf(q[0], x, w1);
        float nw1 = fmaf(-q[3], y, fmaf(q[1], x, w2));
        float nw2 = fmaf(-q[4], y, q[2] * x);
        st[2 * s] = nw1;
        st[2 * s + 1] = nw2;
        x = y;
    }
    return x;
}

static inline double biquad2_f64(const double *c, double *st, double x) {
    for (int s = 0; s < 2; s++) {
        const double *q = c + 5 * s;
        double y = q[0] * x + st[2 * s];
        double nw1 = q[1] * x - q[3] * y + st[2 * s + 1];
        double nw2 = q[2] * x - q[4] * y;
        st[2 * s] = nw1;
        st[2 * s + 1] = nw2;
        x = y;
    }
    return x;
}

void orc_iir_preseed_f32(const float *c10, const float *x, size_t n, int warmup, float *st) {
    FP_REFERENCE_MODE();
    for (size_t i = 0; i < n; i++) {
        float *s = st + 4 * i;
        s[0] = s[1] = s[2] = s[3] = 0.0f;
        for (int it = 0; it < warmup; it++) (void)biquad2_f32(c10, s, x[i]);
    }
}

void orc_iir_step_f32(const float *c10, const float *x, size_t n, float *st, float *y) {
    FP_REFERENCE_MODE();
    for (size_t i = 0; i < n; i++) y[i] = biquad2_f32(c10, st + 4 * i, x[i]);
}

void orc_iir_preseed_f64(const double *c10, const double *x, size_t n, int warmup, double *st) {
    for (size_t i = 0; i < n; i++) {
        double *s = st + 4 * i;
        s[0] = s[1] = s[2] = s[3] = 0.0;
        for (int it = 0; it < warmup; it++) (void)biquad2_f64(c10, s, x[i]);
    }
}

void orc_iir_step_f64(const double *c10, const double *x, size_t n, double *st, double *y) {
    for (size_t i = 0; i < n; i++) y[i] = biquad2_f64(c10, st + 4 * i, x[i]);
}

/*============================================================================
 * Threshold and amplification (src/wema.c:714-744, src/phase_amp.c:12-56)
 *
 * The threshold is T = 0.2 * (sum over the N6 reference positions of |c|) / N6. The N6
 * positions are the unique ones twice over (mirrored slots), so the sum is twice the sum
 * over unique positions. It is computed exactly: each |c| is rounded to a multiple of
 * 2^-32 and summed as a 64-bit integer (order-independent), then converted once.
 *===========================================================================*/

uint64_t orc_abs_fixed(float v) {
    FP_REFERENCE_MODE();
    union { float f; uint32_t u; } b = {v};
    uint32_t e = (b.u >> 23) & 0xFF;
    if (e == 0) return 0;                     /* zero / denormal */
    uint64_t m = (b.u & 0x7FFFFF) | 0x800000; /* |v| = m * 2^(e - 150) */
    int sh = (int)e - 150 + 32;
    if (sh >= 0) return sh > 38 ? (uint64_t)1 << 62 : m << sh;
    if (sh < -25) return 0;
    return (m + ((uint64_t)1 << (-sh - 1))) >> (-sh);
}

float orc_threshold_from_fixed(uint64_t sum_unique, size_t n_unique) {
    FP_REFERENCE_MODE();
    double s = (double)sum_unique * (1.0 / 4294967296.0);
    float mag_sum = (float)(2.0 * s);
    float n6 = (float)(2 * n_unique);
    float q = mag_sum / n6;
    return q * 0.2f;
}

/* Fade-in after the preseed (src/wema.c:830-839, src/wema_batch.c:426-435). */
float orc_fade_amp(float amp, int global_frame) {
    FP_REFERENCE_MODE();
    const int fade_in_frames = 15;
    if (global_frame < 0) return 0.0f;
    if (global_frame < fade_in_frames) {
        /* src/wema.c:830-839 and src/wema_batch.c:426-435 write amp * t * t with
         * t = (g + 1) / 15; built -Ofast (reciprocal math, reassociation) both compile to
         * t = (g + 1) * fl(1/15), then amp * (t * t): a multiply by the rounded reciprocal
         * and the square first, which differs from the division and (amp * t) * t by an ulp
         * at some g (g = 2: 4.8000007 against 4.8) */
        const float t = (float)(global_frame + 1) * (1.0f / (float)fade_in_frames);
        const float tt = t * t;
        return amp * tt;
    }
    return amp;
}

/*============================================================================
 * Box / guided / Gaussian (src/wema.c:114-283; src/wema_batch.c:26-132)
 *===========================================================================*/

/* The reference's box mean (src/wema.c:114-171): per line, FP32 prefix sums, the window's
 * sum as a difference of two of them, divided by the clipped count; rows, then columns of
 * the row means (a mean of means). */
static void box_prefix_line_f32(const float *v, size_t st, int n, int r, float *pre, float *out,
                                size_t ost) {
    pre[0] = 0.0f;
    for (int i = 0; i < n; i++) pre[i + 1] = pre[i] + v[(size_t)i * st];
    for (int x = 0; x < n; x++) {
        const int lo = x - r > 0 ? x - r : 0, hi = x + r + 1 < n ? x + r + 1 : n;
        out[(size_t)x * ost] = (pre[hi] - pre[lo]) / (float)(hi - lo);
    }
}

/* Horizontal window sums of one row of n samples, radius r, samples outside [0, n) taken
 * as 0, in the order the GPU kernel uses: the row is cut into blocks of 4 outputs starting
 * at multiples of 4; the first sum of a block adds its window's 2r + 1 samples left to
 * right, starting from the first, each next one is (previous - leaving sample) + entering
 * sample. */
static void box_sum_line_f32(const float *v, int n, int r, float *out) {
#define VAL(i) ((i) >= 0 && (i) < n ? v[(i)] : 0.0f)
    float s = 0.0f;
    for (int x = 0; x < n; x++) {
        if (x % 4 == 0) {
            s = VAL(x - r);
            for (int k = -r + 1; k <= r; k++) s += VAL(x + k);
        } else {
            s = (s - VAL(x - r - 1)) + VAL(x + r);
        }
        out[x] = s;
    }
#undef VAL
}

/* Prefix sums P[1..4] and suffix sums U[1..4] of the group of rows 4g..4g+3 of a column
 * (stride st, n rows, rows outside [0, n) are 0): P1 = h0, P2 = P1 + h1, P3 = P2 + h2,
 * P4 = P3 + h3 (the group's sum F); U1 = h3, U2 = h2 + h3, U3 = h1 + U2, U4 = F. */
static void group_sums_f32(const float *v, size_t st, int n, int g, float P[5], float U[5]) {
    float hv[4];
    for (int i = 0; i < 4; i++) {
        const int y = 4 * g + i;
        hv[i] = y >= 0 && y < n ? v[(size_t)y * st] : 0.0f;
    }
    P[1] = hv[0];
    P[2] = P[1] + hv[1];
    P[3] = P[2] + hv[2];
    P[4] = P[3] + hv[3];
    U[1] = hv[3];
    U[2] = hv[2] + hv[3];
    U[3] = hv[1] + U[2];
    U[4] = P[4];
}

/* Vertical window sums (radius 4) of one column of n rows in the order the GPU kernel
 * streams them, one group of 4 rows at a time: the window of row 4b + j (j = 0..3) covers
 * the last 4 - j rows of group b - 1, all of group b and the first j + 1 rows of group
 * b + 1, and sums to (U_{b-1}(4 - j) + F_b) + P_{b+1}(j + 1). */
static void box_vsum_groups_f32(const float *v, size_t st, int n, float *out, size_t ost) {
    for (int b = 0; 4 * b < n; b++) {
        float Pa[5], Ua[5], Pb[5], Ub[5], Pc[5], Uc[5];
        group_sums_f32(v, st, n, b - 1, Pa, Ua);
        group_sums_f32(v, st, n, b, Pb, Ub);
        group_sums_f32(v, st, n, b + 1, Pc, Uc);
        for (int j = 0; j < 4 && 4 * b + j < n; j++)
            out[(size_t)(4 * b + j) * ost] = (Ua[4 - j] + Pb[4]) + Pc[j + 1];
    }
}

/* Clipped window count at x of n (src/wema.c:144-147). */
static int clip_count(int x, int n, int r) {
    const int lo = x - r > 0 ? x - r : 0, hi = x + r + 1 < n ? x + r + 1 : n;
    return hi > lo ? hi - lo : 1;
}

/* Box mean over the (2r + 1)^2 window clipped to the frame. prefix: the reference's order
 * (above), any r. Otherwise the kernel's, r = 4 (the guided filter's radius,
 * src/wema.c:173): the window sum as the vertical window sum (box_vsum_groups_f32) of
 * horizontal window sums (box_sum_line_f32), times the correctly rounded reciprocal of the
 * clipped window's area; in real arithmetic the same mean as the reference's mean of
 * means. Returns -1 for another radius in the kernel's order. */
int orc_box_f32(const float *in, float *out, int w, int h, int r, int prefix) {
    FP_REFERENCE_MODE();
    if (!prefix && r != 4) return -1;
    size_t P = (size_t)w * h;
    float *tmp = malloc(P * sizeof(float));
    float *pre = malloc((size_t)((w > h ? w : h) + 1) * sizeof(float));
    if (prefix) {
        for (int y = 0; y < h; y++)
            box_prefix_line_f32(in + (size_t)y * w, 1, w, r, pre, tmp + (size_t)y * w, 1);
        for (int x = 0; x < w; x++) box_prefix_line_f32(tmp + x, (size_t)w, h, r, pre, out + x, (size_t)w);
    } else {
        for (int y = 0; y < h; y++) box_sum_line_f32(in + (size_t)y * w, w, r, tmp + (size_t)y * w);
        for (int x = 0; x < w; x++) box_vsum_groups_f32(tmp + x, (size_t)w, h, out + x, (size_t)w);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                const float k = 1.0f / (float)(clip_count(x, w, r) * clip_count(y, h, r));
                out[(size_t)y * w + x] *= k;
            }
    }
    free(tmp);
    free(pre);
    return 0;
}

void orc_box_f64(const double *in, double *out, int w, int h, int r) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double s = 0;
            int n = 0;
            for (int dy = -r; dy <= r; dy++)
                for (int dx = -r; dx <= r; dx++) {
                    int yy = y + dy, xx = x + dx;
                    if (yy < 0 || yy >= h || xx < 0 || xx >= w) continue;
                    s += in[(size_t)yy * w + xx];
                    n++;
                }
            out[(size_t)y * w + x] = s / n;
        }
}

void orc_guided_f32(const float *I, const float *p, float *q, int w, int h, int prefix) {
    FP_REFERENCE_MODE();
    const int r = 4;
    const float eps = 0.01f;
    size_t n = (size_t)w * h;
    float *mI = malloc(n * 4), *mp = malloc(n * 4), *cI = malloc(n * 4), *cIp = malloc(n * 4);
    float *t = malloc(n * 4);
    orc_box_f32(I, mI, w, h, r, prefix);
    orc_box_f32(p, mp, w, h, r, prefix);
    for (size_t i = 0; i < n; i++) t[i] = I[i] * I[i];
    orc_box_f32(t, cI, w, h, r, prefix);
    for (size_t i = 0; i < n; i++) t[i] = I[i] * p[i];
    orc_box_f32(t, cIp, w, h, r, prefix);
    for (size_t i = 0; i < n; i++) {
        float var = fmaf(-mI[i], mI[i], cI[i]);
        float cov = fmaf(-mI[i], mp[i], cIp[i]);
        float a = cov / (var + eps);
        float b = fmaf(-a, mI[i], mp[i]);
        cI[i] = a;
        cIp[i] = b;
    }
    orc_box_f32(cI, mI, w, h, r, prefix);
    orc_box_f32(cIp, mp, w, h, r, prefix);
    for (size_t i = 0; i < n; i++) q[i] = fmaf(mI[i], I[i], mp[i]);
    free(mI); free(mp); free(cI); free(cIp); free(t);
}

void orc_guided_f64(const double *I, const double *p, double *q, int w, int h) {
    const int r = 4;
    const double eps = 0.01;
    size_t n = (size_t)w * h;
    double *mI = malloc(n * 8), *mp = malloc(n * 8), *cI = malloc(n * 8), *cIp = malloc(n * 8);
    double *t = malloc(n * 8);
    orc_box_f64(I, mI, w, h, r);
    orc_box_f64(p, mp, w, h, r);
    for (size_t i = 0; i < n; i++) t[i] = I[i] * I[i];
    orc_box_f64(t, cI, w, h, r);
    for (size_t i = 0; i < n; i++) t[i] = I[i] * p[i];
    orc_box_f64(t, cIp, w, h, r);
    for (size_t i = 0; i < n; i++) {
        double a = (cIp[i] - mI[i] * mp[i]) / (cI[i] - mI[i] * mI[i] + eps);
        cI[i] = a;
        cIp[i] = mp[i] - a * mI[i];
    }
    orc_box_f64(cI, mI, w, h, r);
    orc_box_f64(cIp, mp, w, h, r);
    for (size_t i = 0; i < n; i++) q[i] = mI[i] * I[i] + mp[i];
    free(mI); free(mp); free(cI); free(cIp); free(t);
}

void orc_gauss5_f32(const float *in, float *out, int w, int h) {
    FP_REFERENCE_MODE();
    static const float k[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};
    float *tmp = malloc((size_t)w * h * sizeof(float));
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float s = 0.0f, wt = 0.0f;
            for (int d = -2; d <= 2; d++) {
                int xx = x + d;
                if (xx >= 0 && xx < w) { s = fmaf(k[d + 2], in[(size_t)y * w + xx], s); wt += k[d + 2]; }
            }
            tmp[(size_t)y * w + x] = s / wt;
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float s = 0.0f, wt = 0.0f;
            for (int d = -2; d <= 2; d++) {
                int yy = y + d;
                if (yy >= 0 && yy < h) { s = fmaf(k[d + 2], tmp[(size_t)yy * w + x], s); wt += k[d + 2]; }
            }
            out[(size_t)y * w + x] = s / wt;
        }
    free(tmp);
}

/*============================================================================
 * Whole pipeline, FP32 (src/wema.c:746-952)
 *===========================================================================*/

typedef struct {
    OrcParams p;
    int L, nch, warmup, frames;   /* frames: frames processed since open */
    int lw[32], lh[32];
    size_t nu, P, off[33];
    float c10[10];
    float *orig[3], *filt[3], *ampd[3], *ll[3], *state[3], *ring[3];
    float *chan[3], *recon[3];
} OrcPipe;

void *orc_pipe_open(const OrcParams *prm) {
    FP_REFERENCE_MODE();
    OrcPipe *o = calloc(1, sizeof(OrcPipe));
    o->p = *prm;
    int L = prm->levels;
    if (L <= 0) {
        int m = prm->width < prm->height ? prm->width : prm->height;
        L = 0;
        while (m >= 16 && L < 4) { m /= 2; L++; }
        if (L == 0) L = 1;
    }
    o->L = L;
    o->nch = prm->color ? 3 : 1;
    orc_level_dims(prm->width, prm->height, L, o->lw, o->lh);
    o->off[0] = 0;
    for (int l = 0; l < L; l++) o->off[l + 1] = o->off[l] + 3 * (size_t)o->lw[l + 1] * o->lh[l + 1];
    o->nu = o->off[L];
    o->P = (size_t)prm->width * prm->height;
    if (prm->use_iir && orc_iir_design(prm->f_low, prm->f_high, prm->fps, o->c10, &o->warmup) < 0) {
        free(o);
        return NULL;
    }
    for (int c = 0; c < o->nch; c++) {
        o->orig[c] = calloc(o->nu, 4);
        o->filt[c] = calloc(o->nu, 4);
        o->ampd[c] = calloc(o->nu, 4);
        o->ll[c] = calloc(o->P, 4);
        o->chan[c] = calloc(o->P, 4);
        o->recon[c] = calloc(o->P, 4);
        if (prm->use_iir) o->state[c] = calloc(o->nu * 4, 4);
        else o->ring[c] = calloc(o->nu * (size_t)prm->window, 4);
    }
    return o;
}

void orc_pipe_close(void *h) {
    OrcPipe *o = h;
    if (!o) return;
    for (int c = 0; c < o->nch; c++) {
        free(o->orig[c]); free(o->filt[c]); free(o->ampd[c]); free(o->ll[c]);
        free(o->chan[c]); free(o->recon[c]); free(o->state[c]); free(o->ring[c]);
    }
    free(o);
}

int orc_pipe_warmup(void *h) { return ((OrcPipe *)h)->warmup; }

const float *orc_pipe_buffer(void *h, int which, int ch) {
    OrcPipe *o = h;
    if (ch >= o->nch) return NULL;
    return which == 0 ? o->orig[ch] : which == 1 ? o->filt[ch] : o->ampd[ch];
}

const float *orc_pipe_state(void *h, int ch) {
    OrcPipe *o = h;
    return ch < o->nch ? o->state[ch] : NULL;
}

static int level_of(const OrcPipe *o, size_t i) {
    int l = 0;
    while (i >= o->off[l + 1]) l++;
    return l;
}

/* Sliding-window delta (src/temporal_filter.c:170-319) for the position whose samples,
 * oldest first, are s[0..ws-1]. */
static float window_delta(const float *s, int ws, int bilateral, const float *tw) {
    const int center = ws / 2;
    const float inv_ws = 1.0f / (float)ws;
    float mean = 0.0f;
    for (int t = 0; t < ws; t++) mean += s[t];
    const float cur = s[center];
    mean *= inv_ws;
    if (!bilateral) return s[center] - mean;
    float mad = 0.0f;
    for (int t = 0; t < ws; t++) {
        float d = s[t] - cur;
        mad += d > 0.0f ? d : -d;
    }
    float sigma = fmaxf(mad * inv_ws * 0.5f, 0.001f);
    float inv2 = 1.0f / (2.0f * sigma * sigma);
    float wsum = 0.0f, acc = 0.0f;
    float wt[256];
    for (int t = 0; t < ws; t++) {
        float d = s[t] - cur;
        wt[t] = expf(-(d * d) * inv2) * tw[t];
        wsum += wt[t];
    }
    if (!(wsum > 1e-8f)) return s[center] - mean;
    float inv = 1.0f / wsum;
    for (int t = 0; t < ws; t++) acc = fmaf(wt[t], s[t] - mean, acc);
    return acc * inv;
}

int orc_pipe_frame(void *h, const uint8_t *rgb, uint8_t *out_u8, float *out_f32,
                   const float *thr_override, float *thr_used) {
    FP_REFERENCE_MODE();
    OrcPipe *o = h;
    const OrcParams *p = &o->p;
    const int W = p->width, H = p->height, L = o->L;
    const size_t P = o->P;
    float thr[3] = {NAN, NAN, NAN};
    int ws = p->window;
    int t_frame = o->frames;
    int passthrough = 0;

    for (int c = 0; c < o->nch; c++) {
        orc_ingest_f32(rgb, P, p->color ? c : 0, o->chan[c]);
        float *planes = p->use_iir ? o->orig[c] : o->ring[c] + (size_t)(t_frame % ws) * o->nu;
        orc_dwt_fwd_f32(o->chan[c], W, H, L, planes, o->ll[c]);
        if (!p->use_iir) memcpy(o->orig[c], planes, o->nu * 4);
    }
    if (!p->use_iir && t_frame + 1 < ws) passthrough = 1;

    float amp = p->amp;
    if (p->use_iir) {
        if (t_frame == 0)
            for (int c = 0; c < o->nch; c++)
                orc_iir_preseed_f32(o->c10, o->orig[c], o->nu, o->warmup, o->state[c]);
        amp = orc_fade_amp(p->amp, t_frame);
    }
    o->frames++;
    if (passthrough) {
        for (size_t i = 0; i < 3 * P; i++) {
            float v = u8f(rgb[i]);
            if (out_f32) out_f32[i] = v;
            if (out_u8) out_u8[i] = egress_u8(v);
        }
        if (thr_used) memcpy(thr_used, thr, sizeof(thr));
        return 0;
    }

    float tw[256];
    if (!p->use_iir && p->bilateral) {
        const float ts = (float)ws * 0.15f;
        const float inv = 1.0f / (2.0f * ts * ts);
        for (int t = 0; t < ws; t++) {
            float d = (float)(t - ws / 2);
            tw[t] = expf(-(d * d) * inv);
        }
    }

    for (int c = 0; c < o->nch; c++) {
        const float *cc = o->orig[c];
        if (p->use_iir) {
            orc_iir_step_f32(o->c10, cc, o->nu, o->state[c], o->filt[c]);
        } else {
            float s[256];
            for (size_t i = 0; i < o->nu; i++) {
                for (int t = 0; t < ws; t++)
                    s[t] = o->ring[c][(size_t)((t_frame + 1 + t) % ws) * o->nu + i];
                o->filt[c][i] = window_delta(s, ws, p->bilateral, tw);
            }
        }
        uint64_t acc = 0;
        for (size_t i = 0; i < o->nu; i++) acc += orc_abs_fixed(cc[i]);
        thr[c] = thr_override ? thr_override[c] : orc_threshold_from_fixed(acc, o->nu);
        float a_ch = c == 0 ? amp : amp * 0.5f;
        for (size_t i = 0; i < o->nu; i++) {
            int l = level_of(o, i);
            float al = a_ch * (l == 0 ? 0.5f : 1.0f);
            float v = cc[i];
            o->ampd[c][i] = (fabsf(v) <= thr[c]) ? v : fmaf(al, o->filt[c][i], v);
        }
        orc_dwt_inv_f32(o->ampd[c], o->ll[c], W, H, L, o->recon[c]);
    }
    if (thr_used) memcpy(thr_used, thr, sizeof(thr));

    if (p->color) {
        for (size_t i = 0; i < P; i++) {
            /* src/wema.c:80-102 */
            float yv = o->recon[0][i], cbv = o->recon[1][i] - 0.5f, crv = o->recon[2][i] - 0.5f;
            float rgbv[3];
            rgbv[0] = fmaf(1.402f, crv, yv);
            rgbv[1] = yv - fmaf(0.344f, cbv, 0.714f * crv);
            rgbv[2] = fmaf(1.772f, cbv, yv);
            for (int k = 0; k < 3; k++) {
                float v = clamp01(rgbv[k]);
                if (out_f32) out_f32[3 * i + k] = v;
                if (out_u8) out_u8[3 * i + k] = egress_u8(v);
            }
        }
        return 0;
    }

    /* Luma: guided (or Gaussian) smoothing of the luma delta, src/wema.c:289-335 */
    float *I = o->chan[0], *d = malloc(P * 4), *q = malloc(P * 4);
    for (size_t i = 0; i < P; i++) d[i] = o->recon[0][i] - I[i];
    if (p->edge_aware) orc_guided_f32(I, d, q, W, H, p->box_prefix);
    else orc_gauss5_f32(d, q, W, H);
    for (size_t i = 0; i < P; i++)
        for (int k = 0; k < 3; k++) {
            float v = clamp01(u8f(rgb[3 * i + k]) + q[i]);
            if (out_f32) out_f32[3 * i + k] = v;
            if (out_u8) out_u8[3 * i + k] = egress_u8(v);
        }
    free(d);
    free(q);
    return 0;
}
