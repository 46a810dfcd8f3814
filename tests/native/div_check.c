/* Checks that div_reuse of fesomtra_types.hpp (the reciprocal-corrected quotient used by the
 * flux divergence of one-column-per-wavefront kernels) equals the IEEE quotient a / b bit for
 * bit: random operands over a wide exponent range, divisor significands next to 1 and to 2,
 * signed zeros and tiny numerators.  Prints the number of mismatches; exit status 1 if any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st = 0x9e3779b97f4a7c15ull;
static uint64_t rnd(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }
static double mk(uint64_t mant, int e, int neg) {
    uint64_t b = ((uint64_t)(1023 + e) << 52) | (mant & ((1ull << 52) - 1)) | ((uint64_t)neg << 63);
    double d;
    memcpy(&d, &b, 8);
    return d;
}
/* the same expression as the device function */
static double div_reuse(double a, double b, double y) {
    const double q0 = a * y;
    if (a == 0.0) return q0;
    const double aq = fabs(q0);
    if (aq > 0x1p-960 && aq < 0x1p960) return fma(fma(-q0, b, a), y, q0);
    return a / b;
}
static int same(double x, double y) { return memcmp(&x, &y, 8) == 0; }

int main(int argc, char **argv) {
    const long n = argc > 1 ? atol(argv[1]) : 10000000L;
    long bad = 0;
    for (long it = 0; it < n; ++it) {
        const uint64_t r1 = rnd(), r2 = rnd();
        uint64_t mb = r2;
        if ((it & 3) == 1) mb = (1ull << 52) - 1 - (r2 & 0xff); /* divisor significand next to 2 */
        if ((it & 3) == 2) mb = r2 & 0xff;                      /* next to 1 */
        const double a = mk(r1, (int)((r1 >> 53) % 1400) - 700, (int)(r1 >> 63));
        const double b = mk(mb, (int)((r2 >> 53) % 400) - 200, (int)((r2 >> 52) & 1));
        if (!same(div_reuse(a, b, 1.0 / b), a / b)) ++bad;
    }
    const double as[6] = {0.0, -0.0, 0x1p-1074, -0x1p-1000, 0x1p1000, -0x1.fffffffffffffp1023};
    const double bs[3] = {3.0e9, -7.5, 0x1p-20};
    for (int i = 0; i < 6; ++i)
        for (int k = 0; k < 3; ++k)
            if (!same(div_reuse(as[i], bs[k], 1.0 / bs[k]), as[i] / bs[k])) ++bad;
    printf("%ld\n", bad);
    return bad != 0;
}
