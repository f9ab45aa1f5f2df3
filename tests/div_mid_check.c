/* Check of the covariance division of the guided-filter kernels (div_mid):
 *   y = fma(fma(-b, est, 1), est, est); q = a * y;
 *   q = fma(fma(-b, q, a), y, q); q = fma(fma(-b, q, a), y, q)
 * against the IEEE quotient a / b, for b in [2^-7, 2) and 2^-66 <= |a| < 2^64 or a = +0,
 * from each starting reciprocal est within an ulp of 1/b: RN(1/b) and its two neighbours.
 * Every significand of b in each of its 8 binades is taken, with random a of both signs
 * over the whole exponent range, the range's end points, 1 and +0. FTZ/DAZ are set, as on
 * the GPU.
 *
 * Result: from RN(1/b) the sequence is exact everywhere. From either neighbour it misses
 * only where b has the largest significand of its binade (b = 2^k (2 - 2^-23), 1/b just
 * above a rounding tie): the refined reciprocal rounds to the float below RN(1/b), and a
 * quotient on a tie (a = 1 here, or any power of two) rounds the wrong way. The program
 * fails if any other mismatch appears. So the kernels equal IEEE division provided the
 * hardware estimate (v_rcp_f32) is RN(1/b) at those 8 values of b, which
 * tests/test_gpu_pipeline.py checks on the GPU over every b in the range.
 * (For a = -0 the sequence gives +0 where IEEE division gives -0: a zero that the guided
 * filter only adds to other terms and to a source pixel >= +0, so no output byte can
 * depend on it; -0 is left out.) */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <xmmintrin.h>

static uint64_t rs = 0x9E3779B97F4A7C15ull;
static uint32_t rnd(void) {
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return (uint32_t)(rs >> 11);
}

static float bits(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static float div_mid(float a, float b, float est) {
    const float y = fmaf(fmaf(-b, est, 1.0f), est, est);
    float q = a * y;
    q = fmaf(fmaf(-b, q, a), y, q);
    return fmaf(fmaf(-b, q, a), y, q);
}

int main(void) {
    _mm_setcsr(_mm_getcsr() | 0x8040u);
    long total = 0, bad[3] = {0, 0, 0}, bad_other_b = 0;
    const float ends[4] = {bits((127u - 66u) << 23), bits(((127u + 64u) << 23) - 1u), 0.0f, 1.0f};
    for (uint32_t e = 127u - 7u; e < 127u + 1u; e++) {
        for (uint32_t m = 0; m < (1u << 23); m++) {
            const float b = bits((e << 23) | m);
            volatile float vb = b;
            const float r = 1.0f / vb;
            const float est[3] = {r, nextafterf(r, INFINITY), nextafterf(r, 0.0f)};
            float a[5];
            a[0] = bits(((uint32_t)(127 - 66 + (int)(rnd() % 130u)) << 23) | (rnd() & 0x7FFFFFu));
            a[1] = -bits(((uint32_t)(127 - 66 + (int)(rnd() % 130u)) << 23) | (rnd() & 0x7FFFFFu));
            a[2] = ends[m & 3u];
            a[3] = -ends[(m >> 2) & 1u];
            a[4] = 1.0f;
            for (int i = 0; i < 5; i++) {
                volatile float va = a[i];
                const float want = va / vb;
                for (int k = 0; k < 3; k++) {
                    const float got = div_mid(a[i], b, est[k]);
                    total++;
                    if (memcmp(&got, &want, 4) != 0) {
                        if (bad[0] + bad[1] + bad[2] < 5)
                            printf("a=%a b=%a est=%a want=%a got=%a\n", a[i], b, est[k], want, got);
                        bad[k]++;
                        if (m != 0x7FFFFFu || (a[i] != 1.0f && a[i] != 0.0f)) bad_other_b++;
                    }
                }
            }
        }
    }
    printf("%ld quotients; mismatches from RN(1/b) %ld, from the float above %ld, from the float "
           "below %ld (%ld at other b)\n", total, bad[0], bad[1], bad[2], bad_other_b);
    return (bad[0] || bad_other_b) ? 1 : 0;
}
