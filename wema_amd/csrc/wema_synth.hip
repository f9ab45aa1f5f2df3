// Synthetic video for benchmarks and tests (see synth.h): presets for the measured
// workloads, the static texture, and frames on the host or the device.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/wema_dev.h"
#include "synth.h"
#include "wema_host.h"

extern "C" hipError_t wema_launch_synth(const SynthFrame *, const unsigned short *, unsigned char *,
                                        int, int, int, hipStream_t);

extern "C" int wema_synth_preset(wema_synth_params *p, int preset, int width, int height,
                                 float fps, uint32_t seed_offset) {
    if (!p) { wema_set_error("null argument"); return -1; }
    memset(p, 0, sizeof(*p));
    p->width = width;
    p->height = height;
    p->fps = fps;
    p->seed = 0x57454D41u + seed_offset;
    p->noise_sigma = 2.0f;
    p->disc_contrast = 60.0f;
    p->disc_radius = 40.0f * (float)height / 240.0f;
    switch (preset) {
    case 1:   // 320x240@30: one disc of radius 40 px moving 0.2 px at 1.5 Hz
    case 4:   // the same scene at any size (multi-stream mix)
        p->kind = WEMA_SYNTH_TEXTURE_DISCS;
        p->ndiscs = 1;
        p->disc_cx[0] = 0.5f; p->disc_cy[0] = 0.5f;
        p->disc_amp[0] = 0.2f; p->disc_freq[0] = 1.5f;
        if (preset == 1) p->disc_radius = 40.0f;
        break;
    case 2:   // 1920x1080@30: three discs at 0.8 / 1.5 / 2.5 Hz, 0.1 / 0.2 / 0.3 px, tinted
        p->kind = WEMA_SYNTH_TEXTURE_DISCS;
        p->ndiscs = 3;
        p->color = 1;
        p->tint_r = 30.0f;
        p->tint_b = -15.0f;
        for (int d = 0; d < 3; d++) {
            p->disc_cx[d] = 0.25f * (float)(d + 1);
            p->disc_cy[d] = 0.5f;
        }
        p->disc_freq[0] = 0.8f; p->disc_freq[1] = 1.5f; p->disc_freq[2] = 2.5f;
        p->disc_amp[0] = 0.1f; p->disc_amp[1] = 0.2f; p->disc_amp[2] = 0.3f;
        break;
    case 3:   // 3840x2160@60: 8 px checkerboard (contrast 40), 0.15 px shear at 1.2 Hz
        p->kind = WEMA_SYNTH_CHECKER_SHEAR;
        p->cell = 8;
        p->checker_contrast = 40.0f;
        p->shear_amp = 0.15f;
        p->shear_freq = 1.2f;
        p->noise_sigma = 1.5f;
        break;
    default:
        wema_set_error("unknown synth preset %d", preset);
        return -1;
    }
    return 0;
}

static SynthFrame frame_params(const wema_synth_params *p, int frame) {
    SynthFrame f;
    memset(&f, 0, sizeof(f));
    f.width = p->width;
    f.height = p->height;
    f.seed = p->seed;
    f.frame = (uint32_t)frame;
    f.kind = p->kind;
    f.color = p->color;
    f.ndiscs = p->ndiscs < WEMA_SYNTH_MAX_DISCS ? p->ndiscs : WEMA_SYNTH_MAX_DISCS;
    const double t = (double)frame / (double)p->fps;
    for (int d = 0; d < f.ndiscs; d++) {
        const double cx = (double)p->disc_cx[d] * p->width +
                          (double)p->disc_amp[d] * sin(2.0 * M_PI * (double)p->disc_freq[d] * t);
        f.disc_cx_q16[d] = llround(cx * 65536.0);
        f.disc_cy_q16[d] = llround((double)p->disc_cy[d] * p->height * 65536.0);
    }
    f.disc_r_q16 = llround((double)p->disc_radius * 65536.0);
    const int base = (int)lround((double)p->disc_contrast * 256.0);
    f.disc_contrast_q8[0] = base + (p->color ? (int)lround((double)p->tint_r * 256.0) : 0);
    f.disc_contrast_q8[1] = base;
    f.disc_contrast_q8[2] = base + (p->color ? (int)lround((double)p->tint_b * 256.0) : 0);
    f.noise_sigma_q8 = (int)lround((double)p->noise_sigma * 256.0);
    f.shear_amp_q16 = (int)lround((double)p->shear_amp * 2.0 *
                                  sin(2.0 * M_PI * (double)p->shear_freq * t) * 65536.0);
    f.cell = p->cell > 0 ? p->cell : 8;
    const int half = (int)lround((double)p->checker_contrast * 128.0);
    f.checker_lo_q8 = 128 * 256 - half;
    f.checker_hi_q8 = 128 * 256 + half;
    return f;
}

// Static texture: uniform 0..255 noise blurred by a sampled Gaussian (sigma 3 px, radius 9)
// with integer weights summing to 2^16, edges clamped. Result in Q8 levels.
extern "C" int wema_synth_texture_host(const wema_synth_params *p, uint16_t *tex) {
    if (!p || !tex) { wema_set_error("null argument"); return -1; }
    const int W = p->width, H = p->height, R = WEMA_SYNTH_GAUSS_R;
    if (p->kind != WEMA_SYNTH_TEXTURE_DISCS) {
        memset(tex, 0, sizeof(uint16_t) * (size_t)W * H);
        return 0;
    }
    int wgt[2 * WEMA_SYNTH_GAUSS_R + 1];
    double gs[2 * WEMA_SYNTH_GAUSS_R + 1], tot = 0.0;
    for (int k = -R; k <= R; k++) tot += gs[k + R] = exp(-(double)(k * k) / (2.0 * 9.0));
    int sum = 0;
    for (int k = 0; k < 2 * R + 1; k++) sum += wgt[k] = (int)lround(gs[k] / tot * 65536.0);
    wgt[R] += 65536 - sum;
    std::vector<int32_t> u((size_t)W * H), hrow((size_t)W * H);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) u[(size_t)y * W + x] = synth_tex_noise(p->seed, x, y);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int32_t s = 0;
            for (int k = -R; k <= R; k++) {
                int xx = x + k < 0 ? 0 : (x + k >= W ? W - 1 : x + k);
                s += wgt[k + R] * u[(size_t)y * W + xx];
            }
            hrow[(size_t)y * W + x] = s;
        }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int64_t s = 0;
            for (int k = -R; k <= R; k++) {
                int yy = y + k < 0 ? 0 : (y + k >= H ? H - 1 : y + k);
                s += (int64_t)wgt[k + R] * hrow[(size_t)yy * W + x];
            }
            tex[(size_t)y * W + x] = (uint16_t)((s + (1ll << 23)) >> 24);
        }
    return 0;
}

extern "C" int wema_synth_frame_host(const wema_synth_params *p, const uint16_t *tex, int frame,
                                     uint8_t *rgb) {
    if (!p || !rgb) { wema_set_error("null argument"); return -1; }
    const SynthFrame f = frame_params(p, frame);
    const int W = p->width, H = p->height;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t i = (size_t)y * W + x;
            const int tq = tex ? (int)tex[i] : 0;
            uint8_t *o = rgb + i * 3;
            o[0] = synth_pixel(f, tq, x, y, 0);
            o[1] = f.color ? synth_pixel(f, tq, x, y, 1) : o[0];
            o[2] = f.color ? synth_pixel(f, tq, x, y, 2) : o[0];
        }
    return 0;
}

extern "C" int wema_synth_frames_device(const wema_synth_params *p, const uint16_t *d_tex,
                                        int frame0, int n, uint8_t *d_rgb, void *hip_stream) {
    if (!p || !d_rgb || n < 1) { wema_set_error("bad argument"); return -1; }
    hipStream_t st = (hipStream_t)hip_stream;
    std::vector<SynthFrame> fr((size_t)n);
    for (int i = 0; i < n; i++) fr[(size_t)i] = frame_params(p, frame0 + i);
    SynthFrame *d_fr = nullptr;
    if (hipMalloc(&d_fr, sizeof(SynthFrame) * n) != hipSuccess) {
        wema_set_error("hipMalloc failed");
        return -1;
    }
    hipError_t e = hipMemcpy(d_fr, fr.data(), sizeof(SynthFrame) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = wema_launch_synth(d_fr, (const unsigned short *)d_tex, d_rgb, p->width, p->height, n, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(d_fr);
    if (e != hipSuccess) {
        wema_set_error("synth failed: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}
