/*
 * wema.h - the frame-in/frame-out pipeline API of curioustorvald/wema (src/wema.h:253-283),
 * kept signature for signature, implemented in C11 over the MI355X C ABI (wema_dev.h).
 *
 * A program written against the reference's wema.h (its main.c and async_io.c) builds
 * against this header and links wema_amd/libwema.so instead of the reference's wema.c,
 * wema_batch.c, dwt.c, iir_filter.c, temporal_filter.c and phase_amp.c. Frames stay host
 * float RGB in [0,1] (Frame, src/wema.h:28-33); the pyramids, IIR state and window ring
 * live on the GPU for the life of the context.
 *
 * Differences, all additive:
 *   - num_levels may be changed between wema_init and the first frame (the reference
 *     always uses its automatic rule, src/wema.c:384); up to 12 levels;
 *   - frame_alloc allocates pinned host memory, so frames copy to and from the GPU at full
 *     link speed, while the process's pinned frame bytes stay within a cap
 *     (WEMA_PINNED_MAX_BYTES, default a quarter of physical memory up to 16 GiB); past it,
 *     or where pinning fails, frames are 64-byte-aligned pageable memory as in the
 *     reference; frame_free releases either (wema_frame_pinned tells which);
 *   - wema_batch_process streams its frames through the GPU in sub-batches, so the copies
 *     of one sub-batch overlap the processing of the next;
 *   - the device session is opened at the first wema_process_frame / wema_batch_process
 *     call, when the fields the caller writes after wema_init (src/main.c:288-290) are
 *     final; option errors (an invalid band or window) are reported there as -1;
 *   - one context runs either the per-frame or the batch path (the reference shares the
 *     IIR state between them; mixing them here returns -1);
 *   - wema_last_error() returns the text of the last error (the reference only prints).
 * frame_rgb_to_gray and frame_gray_to_rgb (src/wema.h:281-283) are CPU utilities outside
 * the per-frame path and are not provided.
 */
#ifndef WEMA_H
#define WEMA_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WEMA_TEMPORAL_WINDOW_DEF 32   /* src/wema.h:21 */
#define WEMA_TEMPORAL_WINDOW_MAX 256  /* src/wema.h:22 */
#define WEMA_BATCH_SIZE_DEF 64        /* src/wema.h:218 */
#define WEMA_BATCH_SIZE_MAX 256       /* src/wema.h:219 */

/* src/wema.h:28-33 */
typedef struct {
    float *data;      /* pixel data [height * width * channels] */
    int    width;
    int    height;
    int    channels;  /* 3 = RGB (the pipeline's input and output) */
} Frame;

struct wema_dev;

/* The public fields of WemaContext (src/wema.h:102-160) with the same names and meaning;
 * the reference's CPU buffers are replaced by one device session. */
typedef struct {
    int    width;
    int    height;
    float  fps;
    float  amp_factor;
    float  f_low;
    float  f_high;
    int    num_levels;
    int    temporal_window;
    bool   edge_aware;
    bool   bilateral_temp;
    bool   color_mode;
    bool   use_iir;
    float  f_low_norm;
    float  f_high_norm;

    /* device state */
    struct wema_dev *dev;
    int    dev_batch;      /* 0 not open, 1 per-frame path, >1 batch path of that size */
    bool   color_ready;    /* wema_init_color done */
    bool   iir_color_ready;
    long   frames;         /* frames processed */
    int    warmup;         /* IIR warm-up iterations (0 in window mode) */
} WemaContext;

/* src/wema.h:221-247 (public fields); the batch frames live in the device session. */
typedef struct {
    int     batch_size;
    int     num_threads;      /* accepted and ignored: the GPU does the work */
    size_t  num_positions;    /* the reference's position count N6 (dwt.c:769-778) */
    int     width;
    int     height;
    int     frames_in_batch;
    int     warmup_frames;
} WemaBatchContext;

/* src/wema.c:359-440: sets the fields (edge-aware and bilateral on, colour off, the
 * automatic level count), checks the window (src/temporal_filter.c); 0 or -1. */
int  wema_init(WemaContext *ctx, int width, int height, float fps,
               float amp_factor, float f_low, float f_high,
               int temporal_window);
int  wema_init_color(WemaContext *ctx);      /* src/wema.c:443-489 */
int  wema_init_iir(WemaContext *ctx);        /* src/wema.c:492-532: checks the band */
int  wema_init_iir_color(WemaContext *ctx);  /* src/wema.c:535-589 */
void wema_free(WemaContext *ctx);            /* src/wema.c:591-685 */
/* src/wema.c:746-952: one frame in, one frame out (the input itself while a window-mode
 * window fills); 0 or -1. */
int  wema_process_frame(WemaContext *ctx, const Frame *in, Frame *out);
/* src/wema.c:954-960: nothing is held back, so always -1. */
int  wema_flush(WemaContext *ctx, Frame *out);
/* src/wema.c:687-692: IIR: warmed up (after the first frame, whose preseed runs the
 * warm-up); window mode: the window is full. */
bool wema_ready(const WemaContext *ctx);

/* src/wema_batch.c:138-217, 219-234, 352-457: frame_count frames of the stream in, the same
 * number out (guided filter always on, src/wema_batch.c:452); frame_count is clamped to the
 * batch size (src/wema_batch.c:355-357); returns that count or -1. */
int  wema_batch_init(WemaBatchContext *batch, const WemaContext *wema,
                     int batch_size, int num_threads);
void wema_batch_free(WemaBatchContext *batch);
int  wema_batch_process(WemaContext *wema, WemaBatchContext *batch,
                        const Frame *frames_in, Frame *frames_out,
                        int frame_count);

/* src/wema.c:29-40 (pinned host memory here, within the cap above) */
int  frame_alloc(Frame *frame, int width, int height, int channels);
void frame_free(Frame *frame);
/* 1 if frame_alloc pinned the frame, 0 if it is pageable, -1 if frame_alloc did not
 * allocate it (no counterpart in the reference). */
int  wema_frame_pinned(const Frame *frame);

/* Types of the CLI and its FFmpeg pipes, which live in the same header in the reference
 * (src/wema.h:166-212); carried with the same fields so that its main.c, ffmpeg_io.c and
 * async_io.c build against this header unchanged. The pipeline does not use them. */
typedef struct {
    int      in_pid, in_fd;
    int      out_pid, out_fd;
    int      width, height;
    float    fps;
    int64_t  total_frames;
    char    *input_path;
    uint8_t *raw_buffer;
    size_t   buffer_size;
} FFmpegIO;

typedef struct {
    char  *input_path;
    char  *output_path;
    float  amp_factor;
    float  f_low, f_high;
    int    temporal_window;
    char  *ff_codec;
    char  *ff_options;
    bool   verbose;
    bool   edge_aware;
    bool   bilateral_temp;
    bool   color_mode;
    bool   use_iir;
    int    batch_size;
    int    num_threads;
} WemaConfig;

/* Text of the last error (no counterpart in the reference). */
const char *wema_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
