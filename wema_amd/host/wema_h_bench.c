/*
 * Times the drop-in wema.h path (libwema.so) the way the reference's CLI drives it
 * (src/main.c:343-470): host float RGB frames from frame_alloc, allocated as async_io does
 * (2 slots x batch x (in, out), src/async_io.c:215-233), processed by wema_batch_process
 * (luma IIR, src/main.c:355) or frame by frame by wema_process_frame (colour, which never
 * batches). Frames are the deterministic synthetic clip, converted u8 -> float as
 * ffio_read_frame does (src/ffmpeg_io.c:292-294) outside the timed region; the timed region
 * is the wema_* calls only (the reference's reader and writer threads convert concurrently).
 * The first batch (IIR preseed, tile tuning) is not timed.
 *
 * usage: wema_h_bench W H fps levels amp f_low f_high color preset batch batches
 * prints one JSON line.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "wema.h"
#include "wema_dev.h"

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char **argv) {
    if (argc != 12) {
        fprintf(stderr, "usage: %s W H fps levels amp f_low f_high color preset batch batches\n", argv[0]);
        return 2;
    }
    const int W = atoi(argv[1]), H = atoi(argv[2]), levels = atoi(argv[4]), color = atoi(argv[8]);
    const float fps = (float)atof(argv[3]), amp = (float)atof(argv[5]);
    const float fl = (float)atof(argv[6]), fh = (float)atof(argv[7]);
    const int preset = atoi(argv[9]), B = atoi(argv[10]), batches = atoi(argv[11]);
    if (W < 1 || H < 1 || B < 1 || B > WEMA_BATCH_SIZE_MAX || batches < 2) return 2;
    const size_t P = (size_t)W * H;

    WemaContext ctx;
    if (wema_init(&ctx, W, H, fps, amp, fl, fh, WEMA_TEMPORAL_WINDOW_DEF) < 0) {
        fprintf(stderr, "wema_init: %s\n", wema_last_error());
        return 1;
    }
    ctx.color_mode = color;
    if (levels > 0) ctx.num_levels = levels;
    if ((color && wema_init_color(&ctx) < 0) || wema_init_iir(&ctx) < 0 ||
        (color && wema_init_iir_color(&ctx) < 0)) {
        fprintf(stderr, "init: %s\n", wema_last_error());
        return 1;
    }
    WemaBatchContext bc;
    if (!color && wema_batch_init(&bc, &ctx, B, 0) < 0) return 1;

    /* 2 slots x B frames in and out */
    const int NF = 2 * B;
    Frame *in = calloc((size_t)NF, sizeof(Frame)), *out = calloc((size_t)NF, sizeof(Frame));
    int pinned = 0;
    for (int i = 0; i < NF; i++) {
        if (frame_alloc(&in[i], W, H, 3) < 0 || frame_alloc(&out[i], W, H, 3) < 0) {
            fprintf(stderr, "frame_alloc: %s\n", wema_last_error());
            return 1;
        }
        pinned += wema_frame_pinned(&in[i]) + wema_frame_pinned(&out[i]);
    }
    wema_synth_params sp;
    wema_synth_preset(&sp, preset, W, H, fps, 0);
    uint16_t *tex = malloc(P * sizeof(uint16_t));
    uint8_t *rgb = malloc(P * 3);
    wema_synth_texture_host(&sp, tex);
    for (int i = 0; i < NF; i++) {
        wema_synth_frame_host(&sp, tex, i, rgb);
        for (size_t j = 0; j < P * 3; j++) in[i].data[j] = (float)rgb[j] * (1.0f / 255.0f);
    }

    double t_first = 0.0, t_timed = 0.0;
    long timed = 0;
    for (int k = 0; k < batches; k++) {
        Frame *fi = in + (size_t)(k % 2) * B, *fo = out + (size_t)(k % 2) * B;
        const double t0 = now_s();
        if (!color) {
            if (wema_batch_process(&ctx, &bc, fi, fo, B) != B) {
                fprintf(stderr, "wema_batch_process: %s\n", wema_last_error());
                return 1;
            }
        } else {
            for (int i = 0; i < B; i++)
                if (wema_process_frame(&ctx, &fi[i], &fo[i]) < 0) {
                    fprintf(stderr, "wema_process_frame: %s\n", wema_last_error());
                    return 1;
                }
        }
        const double dt = now_s() - t0;
        if (k == 0) {
            t_first = dt;
        } else {
            t_timed += dt;
            timed += B;
        }
    }
    const double mpix = (double)timed * (double)P / 1e6;
    printf("{\"path\": \"%s\", \"width\": %d, \"height\": %d, \"frames_per_call\": %d, \"frames\": %ld, "
           "\"seconds\": %.4f, \"first_batch_s\": %.4f, \"fps\": %.2f, \"mpix_per_s\": %.2f, "
           "\"link_GBs\": %.2f, \"pinned_frames\": %d, \"frames_allocated\": %d}\n",
           color ? "wema_process_frame" : "wema_batch_process", W, H, color ? 1 : B, timed, t_timed,
           t_first, (double)timed / t_timed, mpix / t_timed, 2.0 * 12.0 * (double)timed * (double)P / t_timed / 1e9,
           pinned, 2 * NF);
    for (int i = 0; i < NF; i++) {
        frame_free(&in[i]);
        frame_free(&out[i]);
    }
    if (!color) wema_batch_free(&bc);
    wema_free(&ctx);
    free(in);
    free(out);
    free(tex);
    free(rgb);
    return 0;
}
