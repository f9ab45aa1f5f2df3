// Deterministic synthetic video generator (integer arithmetic only), shared by the host
// (CPU-side frames for the CPU baseline and tests) and the device (frames generated
// directly in HBM for benchmarks). Both sides evaluate the same inline functions, and every
// floating-point quantity (sin of the motion phase, Gaussian weights) is turned into an
// integer on the host before either side uses it, so host and device frames are identical
// byte for byte.
//
// Scene kinds:
//   WEMA_SYNTH_TEXTURE_DISCS: static texture (uniform 0..255 noise blurred by a Gaussian of
//       sigma 3 px, fixed-point weights) + up to 4 discs of +contrast levels, edge
//       anti-aliased by 4x4 supersampled area coverage, each translating horizontally by
//       amp*sin(2*pi*f*t) px, + per-frame Gaussian sensor noise (Irwin-Hall of 4 uniforms).
//   WEMA_SYNTH_CHECKER_SHEAR: checkerboard of cell px cells (contrast levels around 128),
//       horizontally sheared by shear_amp*(y/H-0.5)*2*sin(2*pi*f*t) px, box-filtered over
//       the pixel footprint, + Gaussian sensor noise.
#pragma once
#include <stdint.h>

#define WEMA_SYNTH_TEXTURE_DISCS 0
#define WEMA_SYNTH_CHECKER_SHEAR 1
#define WEMA_SYNTH_MAX_DISCS 4
#define WEMA_SYNTH_GAUSS_R 9

// Per-frame integer parameters, computed on the host by wema_synth_frame_params().
struct SynthFrame {
    int32_t width, height;
    uint32_t seed;
    uint32_t frame;
    int32_t kind;
    int32_t color;                          // 1: independent noise per channel + disc tint
    int32_t ndiscs;
    int64_t disc_cx_q16[WEMA_SYNTH_MAX_DISCS];  // disc centre (px, Q16) for this frame
    int64_t disc_cy_q16[WEMA_SYNTH_MAX_DISCS];
    int64_t disc_r_q16;
    int32_t disc_contrast_q8[3];            // per channel (R, G, B), Q8 levels
    int32_t noise_sigma_q8;                 // Q8 levels
    int32_t shear_amp_q16;                  // px, Q16, already multiplied by 2*sin(...)
    int32_t cell;                           // checkerboard cell size (px)
    int32_t checker_lo_q8, checker_hi_q8;
};

#define SYNTH_FN __host__ __device__ static inline

SYNTH_FN uint64_t synth_mix(uint64_t z) {
    // splitmix64 finaliser (Steele, Lea, Flood 2014)
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

SYNTH_FN uint64_t synth_hash(uint32_t seed, uint32_t purpose, uint32_t frame, uint32_t y,
                             uint32_t x, uint32_t ch) {
    uint64_t h = synth_mix(((uint64_t)seed << 32) | purpose);
    h = synth_mix(h ^ (((uint64_t)frame << 8) | ch));
    h = synth_mix(h ^ (((uint64_t)y << 32) | x));
    return h;
}

// Uniform 0..255 texture noise before blurring.
SYNTH_FN int32_t synth_tex_noise(uint32_t seed, int32_t x, int32_t y) {
    return (int32_t)(synth_hash(seed, 1u, 0u, (uint32_t)y, (uint32_t)x, 0u) & 255u);
}

// Zero-mean approximately Gaussian noise in Q8 levels with the given sigma (Q8).
SYNTH_FN int32_t synth_noise_q8(const SynthFrame &f, int32_t x, int32_t y, int32_t ch) {
    uint64_t h = synth_hash(f.seed, 2u, f.frame, (uint32_t)y, (uint32_t)x, (uint32_t)ch);
    int64_t s = (int64_t)(h & 0xFFFF) + (int64_t)((h >> 16) & 0xFFFF) +
                (int64_t)((h >> 32) & 0xFFFF) + (int64_t)((h >> 48) & 0xFFFF) - 131070;
    // std of the sum of 4 uniforms on [0, 65535] is 65536*sqrt(1/3) ~= 37837
    return (int32_t)((s * f.noise_sigma_q8) / 37837);
}

SYNTH_FN int64_t synth_floordiv(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
    return q;
}

// Coverage (0..16) of pixel (x, y) by disc d, from 4x4 supersampling.
SYNTH_FN int32_t synth_disc_cover(const SynthFrame &f, int d, int32_t x, int32_t y) {
    const int64_t r = f.disc_r_q16;
    int64_t px = ((int64_t)x << 16) + 32768 - f.disc_cx_q16[d];
    int64_t py = ((int64_t)y << 16) + 32768 - f.disc_cy_q16[d];
    int64_t d2 = px * px + py * py;
    int64_t ro = r + 65536, ri = r - 65536;
    if (d2 > ro * ro) return 0;
    if (ri > 0 && d2 < ri * ri) return 16;
    int32_t cnt = 0;
    for (int j = 0; j < 4; j++) {
        int64_t sy = ((int64_t)y << 16) + (2 * j + 1) * 8192 - f.disc_cy_q16[d];
        for (int i = 0; i < 4; i++) {
            int64_t sx = ((int64_t)x << 16) + (2 * i + 1) * 8192 - f.disc_cx_q16[d];
            cnt += (sx * sx + sy * sy <= r * r) ? 1 : 0;
        }
    }
    return cnt;
}

SYNTH_FN int32_t synth_checker_q8(const SynthFrame &f, int32_t x, int32_t y) {
    const int64_t H = f.height;
    int64_t s = ((int64_t)f.shear_amp_q16 * (2 * (int64_t)y + 1 - H)) / (2 * H);
    int64_t a = ((int64_t)x << 16) + s;
    int64_t cw = (int64_t)f.cell << 16;
    int64_t ca = synth_floordiv(a, cw);
    int64_t bnd = (ca + 1) * cw;
    int64_t frac = (a + 65536 > bnd) ? (a + 65536 - bnd) : 0;
    int64_t cy = y / f.cell;
    int64_t v0 = ((ca + cy) & 1) ? f.checker_hi_q8 : f.checker_lo_q8;
    int64_t v1 = ((ca + 1 + cy) & 1) ? f.checker_hi_q8 : f.checker_lo_q8;
    return (int32_t)((v0 * (65536 - frac) + v1 * frac) >> 16);
}

// One output byte. tex_q8: blurred texture value at (x, y) (unused by the checkerboard).
SYNTH_FN uint8_t synth_pixel(const SynthFrame &f, int32_t tex_q8, int32_t x, int32_t y,
                             int32_t ch) {
    int32_t v;
    if (f.kind == WEMA_SYNTH_CHECKER_SHEAR) {
        v = synth_checker_q8(f, x, y);
    } else {
        v = tex_q8;
        for (int d = 0; d < f.ndiscs; d++) {
            int32_t c = synth_disc_cover(f, d, x, y);
            v += (f.disc_contrast_q8[f.color ? ch : 1] * c) / 16;
        }
    }
    v += synth_noise_q8(f, x, y, f.color ? ch : 0);
    int32_t o = (int32_t)synth_floordiv((int64_t)v + 128, 256);
    return (uint8_t)(o < 0 ? 0 : (o > 255 ? 255 : o));
}
