/* Exhaustive check of the short conversion of |v| to a multiple of 2^-32 (round half up)
 * used by the forward-DWT kernel for |v| < 2^30: the integer part hi is exact by
 * truncation, the fraction a - hi is exact, x = fraction * 2^32 is exact and below 2^32;
 * the low word is x truncated, plus one where x - floor(x) >= 0.5 (the kernel's
 * v_fract_f32, exact for x >= 0). It must equal the general conversion (exponent /
 * mantissa shifts) for every float below 2^30 (subnormals count as 0, as the kernels
 * flush them; from 2^30 on the general conversion saturates). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint64_t q32_general(float v) {
    uint32_t u;
    memcpy(&u, &v, 4);
    const uint32_t e = (u >> 23) & 0xFFu;
    if (e == 0) return 0;
    const uint64_t m = (u & 0x7FFFFFu) | 0x800000u;
    const int sh = (int)e - 150 + 32;
    if (sh >= 0) return sh > 38 ? (1ull << 62) : (m << sh);
    if (sh < -25) return 0;
    return (m + (1ull << (-sh - 1))) >> (-sh);
}

static uint64_t q32_short(float v) {
    uint32_t u;
    memcpy(&u, &v, 4);
    if (((u >> 23) & 0xFFu) == 0) return 0;   /* flushed subnormal */
    const float a = fabsf(v);
    const uint32_t hi = (uint32_t)a;          /* integer part (truncation) */
    const float x = (a - (float)hi) * 4294967296.0f;
    const float fr = x - floorf(x);
    const uint32_t lo = (uint32_t)x + (fr >= 0.5f ? 1u : 0u);
    return ((uint64_t)hi << 32) + lo;
}

int main(void) {
    long bad = 0, total = 0;
    for (uint32_t u = 0; u < 0x4E800000u; u++) {   /* every float below 2^30 */
        float v;
        memcpy(&v, &u, 4);
        for (int s = 0; s < 2; s++) {
            const float x = s ? -v : v;
            total++;
            if (q32_general(x) != q32_short(x)) {
                if (bad < 5) printf("v=%a general=%llu short=%llu\n", x, (unsigned long long)q32_general(x),
                                    (unsigned long long)q32_short(x));
                bad++;
            }
        }
    }
    printf("%ld values, %ld mismatches\n", total, bad);
    return bad != 0;
}
