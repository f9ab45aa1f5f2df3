/*
 * C11 host of the frame-in/frame-out API (include/wema.h) over the MI355X C ABI
 * (include/wema_dev.h). Mirrors src/wema.c and src/wema_batch.c of curioustorvald/wema
 * (context set-up, option checks, the per-frame and batch entry points); all pixel work
 * runs in the device session.
 */
#include "wema.h"

#include <stdarg.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "wema_dev.h"

static _Thread_local char g_err[256];

static void set_err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *wema_last_error(void) { return g_err; }

/* Frames are pinned (wema_dev_host_alloc) while the process's pinned frame bytes stay
 * within a cap, so they copy to and from the GPU at full link speed; past the cap, or where
 * pinning fails (no GPU, a locked-memory limit), they are 64-byte-aligned pageable memory,
 * as the reference's frames are (mem_alloc, src/alloc.c:9-19). A 64-byte header before the
 * pixels records which, for frame_free. The cap is WEMA_PINNED_MAX_BYTES if set, else a
 * quarter of physical memory up to 16 GiB (the reference's CLI allocates 2 x batch x 2
 * frames, src/async_io.c:215-233: 25 GB at 4K with its default batch of 64). */
#define FRAME_HDR 64
#define FRAME_MAGIC_PINNED 0x57454D4150494E4EULL   /* "WEMAPINN" */
#define FRAME_MAGIC_PAGED 0x57454D4150414745ULL    /* "WEMAPAGE" */

typedef struct {
    uint64_t magic;
    uint64_t bytes;     /* whole allocation, header included */
} FrameHeader;

static _Atomic size_t g_pinned_bytes;

static size_t pinned_cap(void) {
    const char *env = getenv("WEMA_PINNED_MAX_BYTES");
    if (env && *env) return (size_t)strtoull(env, NULL, 10);
    const long pages = sysconf(_SC_PHYS_PAGES), psz = sysconf(_SC_PAGESIZE);
    size_t cap = pages > 0 && psz > 0 ? (size_t)pages * (size_t)psz / 4 : ((size_t)1 << 30);
    if (cap > ((size_t)16 << 30)) cap = (size_t)16 << 30;
    return cap;
}

static FrameHeader *frame_header(const Frame *frame) {
    return (FrameHeader *)((char *)frame->data - FRAME_HDR);
}

int frame_alloc(Frame *frame, int width, int height, int channels) {
    frame->width = width;
    frame->height = height;
    frame->channels = channels;
    frame->data = NULL;
    if (width < 1 || height < 1 || channels < 1) {
        set_err("invalid frame size %dx%dx%d", width, height, channels);
        return -1;
    }
    const size_t bytes = (size_t)width * height * channels * sizeof(float) + FRAME_HDR;
    char *base = NULL;
    uint64_t magic = FRAME_MAGIC_PINNED;
    const size_t before = atomic_fetch_add(&g_pinned_bytes, bytes);
    if (before + bytes <= pinned_cap()) base = (char *)wema_dev_host_alloc(bytes);
    if (!base) {
        atomic_fetch_sub(&g_pinned_bytes, bytes);
        magic = FRAME_MAGIC_PAGED;
        void *p = NULL;
        if (posix_memalign(&p, FRAME_HDR, bytes) != 0) {
            set_err("out of host memory for a %dx%dx%d frame", width, height, channels);
            return -1;
        }
        base = (char *)p;
    }
    memset(base, 0, bytes);
    FrameHeader *h = (FrameHeader *)base;
    h->magic = magic;
    h->bytes = bytes;
    frame->data = (float *)(base + FRAME_HDR);
    return 0;
}

void frame_free(Frame *frame) {
    if (!frame->data) return;
    FrameHeader *h = frame_header(frame);
    if (h->magic == FRAME_MAGIC_PINNED) {
        atomic_fetch_sub(&g_pinned_bytes, (size_t)h->bytes);
        h->magic = 0;
        wema_dev_host_free(h);
    } else if (h->magic == FRAME_MAGIC_PAGED) {
        h->magic = 0;
        free(h);
    }
    frame->data = NULL;
}

int wema_frame_pinned(const Frame *frame) {
    if (!frame || !frame->data) return -1;
    const uint64_t m = frame_header(frame)->magic;
    return m == FRAME_MAGIC_PINNED ? 1 : (m == FRAME_MAGIC_PAGED ? 0 : -1);
}

int wema_init(WemaContext *ctx, int width, int height, float fps, float amp_factor, float f_low,
              float f_high, int temporal_window) {
    memset(ctx, 0, sizeof(*ctx));
    if (width < 1 || height < 1 || !(fps > 0.0f)) {
        set_err("invalid video %dx%d at %g fps", width, height, fps);
        return -1;
    }
    ctx->width = width;
    ctx->height = height;
    ctx->fps = fps;
    ctx->amp_factor = amp_factor;
    ctx->f_low = f_low;
    ctx->f_high = f_high;
    ctx->temporal_window = temporal_window;
    /* defaults of src/wema.c:371-373 */
    ctx->edge_aware = true;
    ctx->bilateral_temp = true;
    ctx->color_mode = false;
    /* src/wema.c:375-381 */
    ctx->f_low_norm = f_low / (fps / 2.0f);
    ctx->f_high_norm = f_high / (fps / 2.0f);
    if (ctx->f_low_norm < 0.0f) ctx->f_low_norm = 0.0f;
    if (ctx->f_high_norm > 1.0f) ctx->f_high_norm = 1.0f;
    ctx->num_levels = wema_dev_auto_levels(width, height);
    /* the window mode's ring needs at least 2 frames (the CLI asks for 4..256,
     * src/main.c:135-143) */
    if (temporal_window < 2 || temporal_window > WEMA_TEMPORAL_WINDOW_MAX) {
        set_err("temporal window %d out of range", temporal_window);
        return -1;
    }
    return 0;
}

int wema_init_color(WemaContext *ctx) {
    if (ctx->dev) {
        set_err("colour mode must be set up before the first frame");
        return -1;
    }
    ctx->color_ready = true;
    return 0;
}

int wema_init_iir(WemaContext *ctx) {
    /* iir_temporal_init designs the band-pass and fails on an invalid band
     * (src/iir_filter.c:74, 103-154) */
    float c[10];
    int warmup = 0;
    if (wema_dev_design_bandpass(ctx->f_low, ctx->f_high, ctx->fps, c, &warmup) < 0) {
        set_err("%s", wema_dev_last_error());
        return -1;
    }
    if (ctx->dev) {
        set_err("IIR mode must be set up before the first frame");
        return -1;
    }
    ctx->use_iir = true;
    ctx->warmup = warmup;
    return 0;
}

int wema_init_iir_color(WemaContext *ctx) {
    if (!ctx->color_ready) return -1; /* src/wema.c:536 */
    if (!ctx->use_iir && wema_init_iir(ctx) < 0) return -1;
    ctx->iir_color_ready = true;
    return 0;
}

void wema_free(WemaContext *ctx) {
    if (ctx->dev) wema_dev_close(ctx->dev);
    ctx->dev = NULL;
    ctx->dev_batch = 0;
}

/* Opens the device session with the context's final options: batch = 1 for the per-frame
 * path, the batch size for the batch path (guided filter always on there). */
static int open_session(WemaContext *ctx, int batch) {
    if (ctx->dev) {
        if ((batch > 1) != (ctx->dev_batch > 1) || batch > ctx->dev_batch) {
            set_err("a context runs either the per-frame or the batch path");
            return -1;
        }
        return 0;
    }
    wema_dev_params p;
    wema_dev_params_default(&p, ctx->width, ctx->height, ctx->fps);
    p.amp = ctx->amp_factor;
    p.f_low = ctx->f_low;
    p.f_high = ctx->f_high;
    p.window = ctx->temporal_window;
    p.levels = ctx->num_levels;
    p.use_iir = ctx->use_iir;
    /* colour needs wema_init_color (and wema_init_iir_color in IIR mode), as the
     * reference's buffers do (src/wema.c:807, 862) */
    p.color = ctx->color_mode && ctx->color_ready && (!ctx->use_iir || ctx->iir_color_ready);
    if (ctx->color_mode && !p.color) {
        set_err("colour mode needs wema_init_color%s", ctx->use_iir ? " and wema_init_iir_color" : "");
        return -1;
    }
    p.edge_aware = ctx->edge_aware;
    p.bilateral = ctx->bilateral_temp;
    p.batch_path = batch > 1;
    p.max_batch = batch;
    p.device = 0;
    if (wema_dev_open(&ctx->dev, &p) < 0) {
        set_err("%s", wema_dev_last_error());
        ctx->dev = NULL;
        return -1;
    }
    ctx->dev_batch = batch;
    return 0;
}

static int check_frame(const WemaContext *ctx, const Frame *f) {
    if (!f || !f->data || f->width != ctx->width || f->height != ctx->height || f->channels != 3) {
        set_err("frame does not match the context (%dx%d RGB)", ctx->width, ctx->height);
        return -1;
    }
    return 0;
}

int wema_process_frame(WemaContext *ctx, const Frame *in, Frame *out) {
    if (check_frame(ctx, in) < 0 || check_frame(ctx, out) < 0) return -1;
    if (open_session(ctx, 1) < 0) return -1;
    const float *src = in->data;
    float *dst = out->data;
    if (wema_dev_process_host_f32(ctx->dev, &src, &dst, 1) < 0) {
        set_err("%s", wema_dev_last_error());
        return -1;
    }
    ctx->frames++;
    return 0;
}

int wema_flush(WemaContext *ctx, Frame *out) {
    (void)ctx;
    (void)out;
    return -1;
}

bool wema_ready(const WemaContext *ctx) {
    if (ctx->use_iir) return ctx->frames > 0;
    return ctx->frames >= ctx->temporal_window;
}

int wema_batch_init(WemaBatchContext *batch, const WemaContext *wema, int batch_size, int num_threads) {
    memset(batch, 0, sizeof(*batch));
    /* src/wema_batch.c:142-143 */
    if (batch_size < 1) batch_size = WEMA_BATCH_SIZE_DEF;
    if (batch_size > WEMA_BATCH_SIZE_MAX) batch_size = WEMA_BATCH_SIZE_MAX;
    batch->batch_size = batch_size;
    batch->num_threads = num_threads;
    batch->width = wema->width;
    batch->height = wema->height;
    /* N6 = 6 * sum of the subband sizes (dwt_num_positions, src/dwt.c:769-778) */
    size_t s = 0;
    int w = wema->width, h = wema->height;
    for (int l = 0; l < wema->num_levels; l++) {
        w = (w + 1) / 2;
        h = (h + 1) / 2;
        s += (size_t)w * h;
    }
    batch->num_positions = 6 * s;
    batch->warmup_frames = wema->warmup;
    return 0;
}

void wema_batch_free(WemaBatchContext *batch) { memset(batch, 0, sizeof(*batch)); }

/* Frames of a batch call go through the device in sub-batches of this many: each is
 * uploaded, processed and downloaded on its own streams (wema_dev_submit_f32), two on the
 * device at a time, so the copies of one overlap the processing of the next. */
#ifndef WEMA_HOST_SUB_BATCH
#define WEMA_HOST_SUB_BATCH 8
#endif

int wema_batch_process(WemaContext *wema, WemaBatchContext *batch, const Frame *frames_in,
                       Frame *frames_out, int frame_count) {
    if (frame_count < 1) return 0;
    /* src/wema_batch.c:355-357: at most batch_size frames, and that count is returned */
    if (frame_count > batch->batch_size) frame_count = batch->batch_size;
    const int sub = batch->batch_size < WEMA_HOST_SUB_BATCH ? (batch->batch_size > 1 ? batch->batch_size : 2)
                                                            : WEMA_HOST_SUB_BATCH;
    if (open_session(wema, sub) < 0) return -1;
    for (int i = 0; i < frame_count; i++)
        if (check_frame(wema, &frames_in[i]) < 0 || check_frame(wema, &frames_out[i]) < 0) return -1;
    int queued = 0, rc = 0;
    for (int i = 0; i < frame_count && rc == 0; i++) {
        if (wema_dev_submit_f32(wema->dev, frames_in[i].data, frames_out[i].data, (int64_t)i) < 0) rc = -1;
        else queued++;
    }
    /* every queued frame is collected (in order), also after an error */
    for (int i = 0; i < queued; i++) {
        int64_t pts = -1;
        float *o = NULL;
        if (wema_dev_collect_f32(wema->dev, &pts, &o) < 0 || pts != i || o != frames_out[i].data) rc = -1;
    }
    if (rc < 0) {
        set_err("%s", wema_dev_last_error());
        return -1;
    }
    wema->frames += frame_count;
    batch->frames_in_batch = frame_count;
    return frame_count;
}
