/* Exhaustive check of the window-count division used by the guided-filter kernels:
 * q = a * k, k = RN(1/c), q' = fma(fma(-q, c, a), k, q) must equal the IEEE quotient a / c
 * for c = 1..9. The rounding of a / c depends only on the significand of a (for normal a
 * and quotient), so one whole binade of a, both signs, covers every normal a. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

int main(void) {
    long bad = 0, total = 0;
    for (int c = 1; c <= 9; c++) {
        volatile float cf = (float)c;
        const float k = 1.0f / cf;
        for (uint32_t m = 0; m < (1u << 23); m++) {
            const uint32_t u = (127u << 23) | m;
            float a;
            memcpy(&a, &u, 4);
            for (int s = 0; s < 2; s++) {
                const float x = s ? -a : a;
                volatile float want = x / cf;
                const float q = x * k;
                const float got = fmaf(fmaf(-q, cf, x), k, q);
                const float w = want;
                total++;
                if (memcmp(&got, &w, 4) != 0) {
                    if (bad < 5) printf("c=%d a=%a want=%a got=%a\n", c, x, w, got);
                    bad++;
                }
            }
        }
    }
    printf("%ld quotients, %ld mismatches\n", total, bad);
    return bad != 0;
}
