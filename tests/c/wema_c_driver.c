/*
 * Drives the C11 host API (include/wema.h) the way the reference's CLI drives its own
 * (src/main.c:277-320 set-up, 343-425 batch loop, 430-470 per-frame loop): raw RGB24 frames
 * in, raw RGB24 frames out, u8 <-> float as src/ffmpeg_io.c:292-294 and 304-310 do.
 *
 * usage: wema_c_driver in.rgb out.rgb W H N fps amp f_low f_high window levels
 *                      use_iir color edge_aware bilateral batch [over]
 * over > 0: every full batch call passes `over` frames beyond the batch size, which
 * wema_batch_process must clamp away (src/wema_batch.c:355-357), returning the batch size.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "wema.h"

static float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

int main(int argc, char **argv) {
    if (argc != 17 && argc != 18) {
        fprintf(stderr, "usage: %s in out W H N fps amp fl fh window levels iir color edge bil batch [over]\n", argv[0]);
        return 2;
    }
    const int W = atoi(argv[3]), H = atoi(argv[4]), N = atoi(argv[5]);
    const float fps = (float)atof(argv[6]), amp = (float)atof(argv[7]);
    const float fl = (float)atof(argv[8]), fh = (float)atof(argv[9]);
    const int window = atoi(argv[10]), levels = atoi(argv[11]), use_iir = atoi(argv[12]);
    const int color = atoi(argv[13]), edge = atoi(argv[14]), bil = atoi(argv[15]);
    const int batch = atoi(argv[16]);
    const int over = argc == 18 ? atoi(argv[17]) : 0;
    FILE *fi = fopen(argv[1], "rb"), *fo = fopen(argv[2], "wb");
    if (!fi || !fo) return 2;
    const size_t P = (size_t)W * H;
    unsigned char *raw = malloc(P * 3);

    WemaContext ctx;
    if (wema_init(&ctx, W, H, fps, amp, fl, fh, window) < 0) {
        fprintf(stderr, "wema_init: %s\n", wema_last_error());
        return 1;
    }
    ctx.edge_aware = edge;
    ctx.bilateral_temp = bil;
    ctx.color_mode = color;
    if (levels > 0) ctx.num_levels = levels;
    if (color && wema_init_color(&ctx) < 0) return 1;
    if (use_iir) {
        if (wema_init_iir(&ctx) < 0) {
            fprintf(stderr, "wema_init_iir: %s\n", wema_last_error());
            return 1;
        }
        if (color && wema_init_iir_color(&ctx) < 0) return 1;
    }

    const int B = batch > 1 ? batch : 1;
    const int NB = B + over;
    Frame *in = calloc((size_t)NB, sizeof(Frame)), *out = calloc((size_t)NB, sizeof(Frame));
    for (int i = 0; i < NB; i++)
        if (frame_alloc(&in[i], W, H, 3) < 0 || frame_alloc(&out[i], W, H, 3) < 0) return 1;
    WemaBatchContext bc;
    if (batch > 1 && wema_batch_init(&bc, &ctx, batch, 0) < 0) return 1;

    int done = 0, rc = 0;
    while (done < N && rc == 0) {
        const int k = (N - done) < B ? (N - done) : B;
        for (int i = 0; i < k; i++) {
            if (fread(raw, 1, P * 3, fi) != P * 3) return 1;
            for (size_t j = 0; j < P * 3; j++) in[i].data[j] = (float)raw[j] * (1.0f / 255.0f);
        }
        if (batch > 1) {
            const int pass = k == B ? k + over : k;
            const int got = wema_batch_process(&ctx, &bc, in, out, pass);
            if (got != k) {
                fprintf(stderr, "wema_batch_process(%d frames) returned %d, expected %d\n", pass, got, k);
                rc = 1;
            }
        } else {
            if (wema_process_frame(&ctx, &in[0], &out[0]) < 0) rc = 1;
        }
        if (rc) {
            fprintf(stderr, "process: %s\n", wema_last_error());
            break;
        }
        for (int i = 0; i < k; i++) {
            for (size_t j = 0; j < P * 3; j++) raw[j] = (unsigned char)fmaf(clamp01(out[i].data[j]), 255.0f, 0.5f);
            fwrite(raw, 1, P * 3, fo);
        }
        done += k;
    }
    printf("pinned %d frames %d ready %d\n", wema_frame_pinned(&in[0]), done, (int)wema_ready(&ctx));
    if (batch > 1) wema_batch_free(&bc);
    for (int i = 0; i < NB; i++) {
        frame_free(&in[i]);
        frame_free(&out[i]);
    }
    wema_free(&ctx);
    fclose(fi);
    fclose(fo);
    free(raw);
    free(in);
    free(out);
    return rc;
}
