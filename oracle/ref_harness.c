/*
 * Harness around the reference WEMA pipeline (test infrastructure, not product code).
 *
 * Compiled together with the reference sources into oracle/_ref/libwema_ref.so by
 * oracle/Makefile. It drives the reference's public C API (wema.h:253-283, dwt.h,
 * iir_filter.h, temporal_filter.h) with in-memory frames instead of FFmpeg pipes, so the
 * parity tests and the CPU baseline can call it through ctypes.
 *
 * u8 <-> float conversion restates ffio_read_frame / ffio_write_frame
 * (src/ffmpeg_io.c:292-294, 304-310) and is compiled with the same flags as the reference.
 */
#include "wema.h"
#include "dwt.h"
#include "iir_filter.h"
#include "temporal_filter.h"
#include "phase_amp.h"
#include "alloc.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Defined with external linkage in src/wema.c:289 but not declared in any header. */
void frame_gray_to_rgb_ex(const float *gray, const Frame *orig, Frame *out, float *delta_buf,
                          float *smooth_buf, float *guide_buf, float *work1, float *work2,
                          float *work3, float *work4, bool edge_aware);

int refh_max_levels(void) { return WEMA_MAX_LEVELS; }

/* dwt_compute_levels (src/dwt.c:588-596) with the shipped cap of 4 levels. */
int refh_auto_levels(int width, int height) {
    int min_dim = (width < height) ? width : height;
    int levels = 0;
    while (min_dim >= 16 && levels < 4) {
        min_dim /= 2;
        levels++;
    }
    return levels > 0 ? levels : 1;
}

/* src/ffmpeg_io.c:292-294 */
void refh_u8_to_f32(const uint8_t *in, float *out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = (float)in[i] / 255.0f;
}

/* src/ffmpeg_io.c:304-310 */
void refh_f32_to_u8(const float *in, uint8_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) {
        float val = in[i];
        if (val < 0.0f) val = 0.0f;
        if (val > 1.0f) val = 1.0f;
        out[i] = (uint8_t)(val * 255.0f + 0.5f);
    }
}

/* Same loop shape as the threshold in apply_amplified_delta (src/wema.c:718-723),
 * apply_delta_to_coeffs (src/wema_batch.c:271-276) and coeff_amplify (src/phase_amp.c:25-29). */
float refh_threshold(const float *orig, size_t total_pos) {
    float mag_sum = 0.0f;
    for (size_t i = 0; i < total_pos; i++) {
        mag_sum += fabsf(orig[i]);
    }
    return mag_sum / (float)total_pos * 0.2f;
}

void refh_rgb_to_gray(const float *rgb, int w, int h, float *gray) {
    Frame f = {(float *)rgb, w, h, 3};
    frame_rgb_to_gray(&f, gray);
}

static void flatten(const DWTCoeffs *c, float *flat) {
    size_t pos = 0;
    for (int lev = 0; lev < c->num_levels; lev++)
        for (int o = 0; o < WEMA_NUM_ORIENTATIONS; o++) {
            size_t n = (size_t)c->subbands[lev][o].width * c->subbands[lev][o].height;
            memcpy(flat + pos, c->subbands[lev][o].coeffs, n * sizeof(float));
            pos += n;
        }
}

static void unflatten(DWTCoeffs *c, const float *flat) {
    size_t pos = 0;
    for (int lev = 0; lev < c->num_levels; lev++)
        for (int o = 0; o < WEMA_NUM_ORIENTATIONS; o++) {
            size_t n = (size_t)c->subbands[lev][o].width * c->subbands[lev][o].height;
            memcpy(c->subbands[lev][o].coeffs, flat + pos, n * sizeof(float));
            pos += n;
        }
}

/* Forward DWT: flat6 receives the N6 reference-order coefficients (levels x 6 slots),
 * lowpass the final LL band. Returns N6, or -1. */
long refh_dwt_forward(const float *img, int w, int h, int levels, float *flat6, float *lowpass) {
    DWTCoeffs c;
    if (dwt_init(&c, w, h, levels) < 0) return -1;
    dwt_forward(img, w, h, &c);
    size_t n6 = dwt_num_positions(&c);
    if (flat6) flatten(&c, flat6);
    if (lowpass) memcpy(lowpass, c.lowpass, (size_t)c.lowpass_w * c.lowpass_h * sizeof(float));
    dwt_free(&c);
    return (long)n6;
}

long refh_dwt_inverse(const float *flat6, const float *lowpass, int w, int h, int levels,
                      float *img) {
    DWTCoeffs c;
    if (dwt_init(&c, w, h, levels) < 0) return -1;
    unflatten(&c, flat6);
    memcpy(c.lowpass, lowpass, (size_t)c.lowpass_w * c.lowpass_h * sizeof(float));
    dwt_inverse(&c, img);
    size_t n6 = dwt_num_positions(&c);
    dwt_free(&c);
    return (long)n6;
}

/* Band-pass design as iir_temporal_init does it: c[0..4] = HP b0,b1,b2,a1,a2;
 * c[5..9] = LP b0,b1,b2,a1,a2. */
int refh_iir_design(float f_low, float f_high, float fs, float *c, int *warmup) {
    IIRTemporalFilter f;
    if (iir_temporal_init(&f, 16, f_low, f_high, fs) < 0) return -1;
    for (int s = 0; s < 2; s++) {
        c[5 * s + 0] = f.coeffs.sections[s].b0;
        c[5 * s + 1] = f.coeffs.sections[s].b1;
        c[5 * s + 2] = f.coeffs.sections[s].b2;
        c[5 * s + 3] = f.coeffs.sections[s].a1;
        c[5 * s + 4] = f.coeffs.sections[s].a2;
    }
    if (warmup) *warmup = f.warmup_frames;
    iir_temporal_free(&f);
    return 0;
}

/* Runs nf frames of n positions through the reference IIR (preseed from frame 0 when
 * `preseed`, then per-frame or batch processing). out: [nf][n]; state: [4][n]
 * (HP w1, HP w2, LP w1, LP w2). */
int refh_iir_run(float f_low, float f_high, float fs, const float *in, int nf, size_t n,
                 int preseed, int batch, float *out, float *state) {
    IIRTemporalFilter f;
    if (iir_temporal_init(&f, n, f_low, f_high, fs) < 0) return -1;
    if (preseed) iir_temporal_preseed(&f, in);
    if (batch) {
        iir_temporal_process_batch(&f, in, out, nf);
    } else {
        for (int t = 0; t < nf; t++) iir_temporal_process(&f, in + (size_t)t * n, out + (size_t)t * n);
    }
    if (state) {
        for (int s = 0; s < 2; s++) {
            memcpy(state + (size_t)(2 * s) * n, f.state_soa[s].w1, n * sizeof(float));
            memcpy(state + (size_t)(2 * s + 1) * n, f.state_soa[s].w2, n * sizeof(float));
        }
    }
    iir_temporal_free(&f);
    return 0;
}

/*============================================================================
 * Whole pipeline
 *===========================================================================*/

typedef struct {
    WemaContext ctx;
    WemaBatchContext batch;
    int has_batch;
    int w, h, levels, use_iir, color;
    Frame fin, fout;
    Frame *bin, *bout;
    int bcap;
    float thr[3];   /* thresholds of the last processed frame, recomputed */
} RefPipe;

static int reinit_levels(WemaContext *c, int levels) {
    dwt_free(c->coeffs);
    if (dwt_init(c->coeffs, c->width, c->height, levels) < 0) return -1;
    c->num_levels = levels;
    if (c->temporal_buf) {
        temporal_buffer_free(c->temporal_buf);
        if (temporal_buffer_init(c->temporal_buf, c->coeffs, c->temporal_window) < 0) return -1;
    }
    mem_free(c->delta_phi);
    c->delta_phi = mem_alloc(dwt_num_positions(c->coeffs) * sizeof(float));
    return c->delta_phi ? 0 : -1;
}

/* Follows main.c:280-322 (init order, field writes), with the level count forced. */
void *refh_open(int w, int h, float fps, float amp, float f_low, float f_high, int window,
                int levels, int use_iir, int color, int edge_aware, int bilateral) {
    RefPipe *p = calloc(1, sizeof(RefPipe));
    if (!p) return NULL;
    if (wema_init(&p->ctx, w, h, fps, amp, f_low, f_high, window) < 0) { free(p); return NULL; }
    if (levels <= 0) levels = refh_auto_levels(w, h);
    if (levels > WEMA_MAX_LEVELS || reinit_levels(&p->ctx, levels) < 0) goto fail;
    p->ctx.edge_aware = edge_aware != 0;
    p->ctx.bilateral_temp = bilateral != 0;
    p->ctx.color_mode = color != 0;
    if (color && wema_init_color(&p->ctx) < 0) goto fail;
    if (use_iir) {
        if (wema_init_iir(&p->ctx) < 0) goto fail;
        if (color && wema_init_iir_color(&p->ctx) < 0) goto fail;
    }
    p->w = w; p->h = h; p->levels = levels; p->use_iir = use_iir; p->color = color;
    if (frame_alloc(&p->fin, w, h, 3) < 0 || frame_alloc(&p->fout, w, h, 3) < 0) goto fail;
    return p;
fail:
    frame_free(&p->fin);
    frame_free(&p->fout);
    wema_free(&p->ctx);
    free(p);
    return NULL;
}

void refh_close(void *hp) {
    RefPipe *p = hp;
    if (!p) return;
    if (p->has_batch) wema_batch_free(&p->batch);
    for (int i = 0; i < p->bcap; i++) { frame_free(&p->bin[i]); frame_free(&p->bout[i]); }
    free(p->bin);
    free(p->bout);
    frame_free(&p->fin);
    frame_free(&p->fout);
    wema_free(&p->ctx);
    free(p);
}

long refh_num_positions(void *hp) { return (long)dwt_num_positions(((RefPipe *)hp)->ctx.coeffs); }
int refh_warmup(void *hp) {
    RefPipe *p = hp;
    return p->ctx.iir_filt ? p->ctx.iir_filt->warmup_frames : 0;
}

static void record_thresholds(RefPipe *p) {
    WemaContext *c = &p->ctx;
    size_t n6 = dwt_num_positions(c->coeffs);
    p->thr[0] = p->thr[1] = p->thr[2] = NAN;
    if (c->use_iir) {
        p->thr[0] = refh_threshold(c->prev_coeffs, n6);
        if (p->color) {
            p->thr[1] = refh_threshold(c->prev_coeffs_cb, n6);
            p->thr[2] = refh_threshold(c->prev_coeffs_cr, n6);
        }
    } else if (temporal_buffer_ready(c->temporal_buf)) {
        p->thr[0] = refh_threshold(c->temporal_buf->amplitude, n6);
        if (p->color) {
            p->thr[1] = refh_threshold(c->temporal_buf_cb->amplitude, n6);
            p->thr[2] = refh_threshold(c->temporal_buf_cr->amplitude, n6);
        }
    }
}

/* One frame through wema_process_frame (src/wema.c:746). rgb: u8 HWC. Either output may
 * be NULL. thr (3 floats, NaN where unused) receives the thresholds that frame used. */
int refh_frame(void *hp, const uint8_t *rgb, uint8_t *out_u8, float *out_f32, float *thr) {
    RefPipe *p = hp;
    size_t n = (size_t)p->w * p->h * 3;
    refh_u8_to_f32(rgb, p->fin.data, n);
    if (wema_process_frame(&p->ctx, &p->fin, &p->fout) != 0) return -1;
    record_thresholds(p);
    if (thr) memcpy(thr, p->thr, sizeof(p->thr));
    if (out_f32) memcpy(out_f32, p->fout.data, n * sizeof(float));
    if (out_u8) refh_f32_to_u8(p->fout.data, out_u8, n);
    return 0;
}

/* Internal buffers of the last frame for stage-level parity: which = 0 orig coefficients
 * (flat N6), 1 IIR output / window delta (flat N6); ch = 0 Y, 1 Cb, 2 Cr. */
const float *refh_buffer(void *hp, int which, int ch) {
    WemaContext *c = &((RefPipe *)hp)->ctx;
    if (c->use_iir) {
        if (which == 0) return ch == 0 ? c->prev_coeffs : ch == 1 ? c->prev_coeffs_cb : c->prev_coeffs_cr;
        return ch == 0 ? c->delta_phi : ch == 1 ? c->delta_cb : c->delta_cr;
    }
    if (which == 0) {
        TemporalBuffer *b = ch == 0 ? c->temporal_buf : ch == 1 ? c->temporal_buf_cb : c->temporal_buf_cr;
        return b ? b->amplitude : NULL;
    }
    return ch == 0 ? c->delta_phi : ch == 1 ? c->delta_cb : c->delta_cr;
}

/* Amplified coefficients of the last frame (the DWT store after amplification, slots
 * 0..5 flattened): out has N6 floats. */
int refh_amplified(void *hp, int ch, float *out) {
    WemaContext *c = &((RefPipe *)hp)->ctx;
    DWTCoeffs *d = ch == 0 ? c->coeffs : ch == 1 ? c->coeffs_cb : c->coeffs_cr;
    if (!d) return -1;
    flatten(d, out);
    return 0;
}

/* frame_gray_to_rgb_ex (src/wema.c:289-335) on caller data: gray (P), rgb (3P floats). */
int refh_gray_to_rgb(const float *gray, const float *rgb, int w, int h, int edge_aware, float *out) {
    size_t n = (size_t)w * h;
    float *buf = malloc(7 * n * sizeof(float));
    if (!buf) return -1;
    Frame fi = {(float *)rgb, w, h, 3}, fo = {out, w, h, 3};
    frame_gray_to_rgb_ex(gray, &fi, &fo, buf, buf + n, buf + 2 * n, buf + 3 * n, buf + 4 * n,
                         buf + 5 * n, buf + 6 * n, edge_aware != 0);
    free(buf);
    return 0;
}

/* IIR state of channel ch: [4][n] (HP w1, HP w2, LP w1, LP w2). */
int refh_iir_state(void *hp, int ch, float *state) {
    WemaContext *c = &((RefPipe *)hp)->ctx;
    IIRTemporalFilter *f = ch == 0 ? c->iir_filt : ch == 1 ? c->iir_filt_cb : c->iir_filt_cr;
    if (!f || !f->state_soa) return -1;
    size_t n = f->num_positions;
    for (int s = 0; s < 2; s++) {
        memcpy(state + (size_t)(2 * s) * n, f->state_soa[s].w1, n * sizeof(float));
        memcpy(state + (size_t)(2 * s + 1) * n, f->state_soa[s].w2, n * sizeof(float));
    }
    return 0;
}

/* Temporal batch path (wema_batch_process, src/wema_batch.c:352): n frames u8 in, u8 out.
 * thr (n floats, may be NULL) receives each frame's threshold. threads = 0: OpenMP default. */
int refh_batch(void *hp, int batch_size, int threads, const uint8_t *rgb, int n,
               uint8_t *out_u8, float *thr) {
    RefPipe *p = hp;
    if (!p->has_batch) {
        if (wema_batch_init(&p->batch, &p->ctx, batch_size, threads) < 0) return -1;
        p->has_batch = 1;
        p->bcap = p->batch.batch_size;
        p->bin = calloc((size_t)p->bcap, sizeof(Frame));
        p->bout = calloc((size_t)p->bcap, sizeof(Frame));
        if (!p->bin || !p->bout) return -1;
        for (int i = 0; i < p->bcap; i++)
            if (frame_alloc(&p->bin[i], p->w, p->h, 3) < 0 || frame_alloc(&p->bout[i], p->w, p->h, 3) < 0)
                return -1;
    }
    if (n > p->bcap) return -1;
    size_t fs = (size_t)p->w * p->h * 3;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) refh_u8_to_f32(rgb + (size_t)i * fs, p->bin[i].data, fs);
    int got = wema_batch_process(&p->ctx, &p->batch, p->bin, p->bout, n);
    if (got != n) return -1;
    if (thr) {
        for (int i = 0; i < n; i++)
            thr[i] = refh_threshold(p->batch.coeffs_batch + (size_t)i * p->batch.num_positions,
                                    p->batch.num_positions);
    }
    if (out_u8) {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) refh_f32_to_u8(p->bout[i].data, out_u8 + (size_t)i * fs, fs);
    }
    return n;
}

int refh_omp_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
