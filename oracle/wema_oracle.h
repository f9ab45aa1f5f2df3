/*
 * CPU oracle for the WEMA per-frame amplification pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load this library, and only as the checker. The product (wema_amd/) never calls it.
 *
 * Two restatements of the reference algorithm (curioustorvald/wema):
 *   *_f32  FP32, following the reference's operation order and FMA contraction
 *          (the reference is built with -Ofast, which contracts a*b+c into fma);
 *   *_f64  FP64 "exact" version of the same mathematics.
 * Each function names the reference file:line it restates.
 *
 * Coefficient layout ("planes"): for level l = 0..L-1 three planes LH, HH, HL (the
 * reference's slots 0, 1, 2, src/dwt.c:680-682) of sw_l x sh_l floats each, row-major,
 * levels concatenated; sw_l = ceil(w_l / 2), w_0 = W. The reference's 6-slot store repeats
 * slots 2, 1, 0 as slots 3, 4, 5 (src/dwt.c:683-686); orc_planes_to_flat6 builds that
 * order for comparisons.
 */
#ifndef WEMA_ORACLE_H
#define WEMA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int    orc_level_dims(int w, int h, int levels, int *lw, int *lh);
size_t orc_num_unique(int w, int h, int levels);

void orc_fwd1d_f32(float *x, int n);
void orc_inv1d_f32(float *x, int n);
void orc_fwd1d_f64(double *x, int n);
void orc_inv1d_f64(double *x, int n);

void orc_dwt_fwd_f32(const float *img, int w, int h, int levels, float *planes, float *ll);
void orc_dwt_inv_f32(const float *planes, const float *ll, int w, int h, int levels, float *img);
void orc_dwt_fwd_f64(const double *img, int w, int h, int levels, double *planes, double *ll);
void orc_dwt_inv_f64(const double *planes, const double *ll, int w, int h, int levels, double *img);
void orc_planes_to_flat6(const float *planes, int w, int h, int levels, float *flat6);

void orc_ingest_f32(const uint8_t *rgb, size_t npx, int channel, float *out);

int  orc_iir_design(float f_low, float f_high, float fs, float *c10, int *warmup);
int  orc_iir_design_f64(double f_low, double f_high, double fs, double *c10);
void orc_iir_preseed_f32(const float *c10, const float *x, size_t n, int warmup, float *st);
void orc_iir_step_f32(const float *c10, const float *x, size_t n, float *st, float *y);
void orc_iir_preseed_f64(const double *c10, const double *x, size_t n, int warmup, double *st);
void orc_iir_step_f64(const double *c10, const double *x, size_t n, double *st, double *y);

uint64_t orc_abs_fixed(float v);
float    orc_threshold_from_fixed(uint64_t sum_unique, size_t n_unique);
float    orc_fade_amp(float amp, int global_frame);

int orc_box_f32(const float *in, float *out, int w, int h, int r, int prefix);
void orc_box_f64(const double *in, double *out, int w, int h, int r);
void orc_guided_f32(const float *I, const float *p, float *q, int w, int h, int prefix);
void orc_guided_f64(const double *I, const double *p, double *q, int w, int h);
void orc_gauss5_f32(const float *in, float *out, int w, int h);

typedef struct {
    int width, height;
    float fps, amp, f_low, f_high;
    int window;       /* temporal window (window mode) */
    int levels;       /* 0: reference auto rule */
    int use_iir;      /* 1: IIR band-pass (default); 0: sliding window ("--haar") */
    int color;
    int edge_aware;   /* 1: guided filter; 0: 5x5 binomial */
    int bilateral;    /* window mode only */
    int box_prefix;   /* 1: prefix-sum box filter (reference order); 0: direct 9-tap sums */
} OrcParams;

void *orc_pipe_open(const OrcParams *p);
void  orc_pipe_close(void *h);
/* One frame. thr_override (3 floats, NULL = compute exactly) replaces the thresholds;
 * thr_used receives them. Returns 0, or -1. out_f32 / out_u8 may be NULL. */
int   orc_pipe_frame(void *h, const uint8_t *rgb, uint8_t *out_u8, float *out_f32,
                     const float *thr_override, float *thr_used);
/* Coefficients of the last frame: which = 0 original planes, 1 temporal filter output
 * (IIR y or window delta), 2 amplified planes; ch = 0 Y, 1 Cb, 2 Cr. */
const float *orc_pipe_buffer(void *h, int which, int ch);
const float *orc_pipe_state(void *h, int ch);   /* IIR state [n_unique][4] */
int   orc_pipe_warmup(void *h);

#ifdef __cplusplus
}
#endif
#endif
