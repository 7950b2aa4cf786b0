/* Checks two division forms of the FP64 stage kernel against IEEE division,
 * bit for bit.
 * 1. Division of a double by 3 or 6 as
 *      q = fma(x, z, x*z')   (z = RN(1/d), z' = RN(1/d - z)),
 *    signed zeros included: random doubles with 2^-900 <= |x| <= 2^1000,
 *    integers, and x = d*y +- a few ulp (quotients next to rounding
 *    boundaries).
 * 2. Division a/b by a reciprocal refined with two Newton steps,
 *      r = fma(r, fma(-b, r, 1), r) (twice);  q = a*r;  q = fma(fma(-b, q, a), r, q),
 *    from a starting reciprocal with a relative error up to 2^-22 (the
 *    accuracy of the hardware reciprocal), for a, b > 0 in the ranges the
 *    WENO5 weights take: b in [1e-4, 2^812], a/b normal, a >= 2^-969.
 * Prints the number of mismatches; argument: samples. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t next(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }

static double fastdiv(double x, double z, double zl) { return fma(x, z, x * zl); }

/* starting reciprocal: 1/b with a random relative error in [-2^-22, 2^-22] */
static double rough_rcp(double b, uint64_t u) {
    double e = ((double)(int64_t)(u >> 11) / 4503599627370496.0 - 1.0) * 0x1p-22;
    return (1.0 / b) * (1.0 + e);
}

static double nrdiv(double a, double b, double r) {
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    double q = a * r;
    return fma(fma(-b, q, a), r, q);
}

static double rand_in(uint64_t b, int e_lo, int e_hi) {   /* positive, exponent in [e_lo, e_hi] */
    uint64_t e = (uint64_t)(1023 + e_lo + (int)((b >> 12) % (uint64_t)(e_hi - e_lo + 1)));
    return bits((b & 0x000FFFFFFFFFFFFFull) | (e << 52));
}

int main(int argc, char** argv) {
    long n = argc > 1 ? atol(argv[1]) : 10000000;
    const double ds[2] = {3.0, 6.0}, zls[2] = {0x1.5555555555555p-56, 0x1.5555555555555p-57};
    long bad = 0;
    for (int k = 0; k < 2; ++k) {
        const double d = ds[k], z = 1.0 / d, zl = zls[k];
        for (long i = 0; i < n; ++i) {
            double x;
            uint64_t b = next();
            switch (i & 3) {
                case 0: {   /* any double with 2^-900 <= |x| < 2^1001 */
                    uint64_t e = 123 + (b >> 11) % (2024 - 123);
                    x = bits((b & 0x800FFFFFFFFFFFFFull) | (e << 52));
                    break;
                }
                case 1: x = (double)(int64_t)(b >> 11) - (double)(1ll << 52); break;
                case 2: {   /* d*y rounded, then +-3 ulp */
                    double y = bits((b & 0x000FFFFFFFFFFFFFull) | ((uint64_t)(1023 + (int)(b >> 60) - 8) << 52));
                    x = d * y;
                    int m = (int)((b >> 52) & 7) - 3;
                    for (; m > 0; --m) x = nextafter(x, INFINITY);
                    for (; m < 0; ++m) x = nextafter(x, -INFINITY);
                    break;
                }
                default: x = bits((b & 0x800FFFFFFFFFFFFFull) | (0x3FFull << 52)) * 8.0; break;
            }
            double a = x / d, f = fastdiv(x, z, zl);
            if (memcmp(&a, &f, 8) != 0) {
                if (bad < 10) printf("mismatch d=%g x=%a: %a vs %a\n", d, x, a, f);
                ++bad;
            }
        }
        double zz[2] = {0.0, -0.0};
        for (int t = 0; t < 2; ++t) {
            double a = zz[t] / d, f = fastdiv(zz[t], z, zl);
            if (memcmp(&a, &f, 8) != 0) ++bad;
        }
    }
    /* variable divisor, refined reciprocal */
    for (long i = 0; i < n; ++i) {
        uint64_t u1 = next(), u2 = next(), u3 = next();
        double b = (i & 1) ? rand_in(u1, -14, 812) : rand_in(u1, -14, 20);
        double a = (i & 2) ? rand_in(u2, -969, 14) : rand_in(u2, -10, 14);
        double ex = a / b;
        if (!(ex >= 0x1p-1022) || !isfinite(ex)) continue;
        double f = nrdiv(a, b, rough_rcp(b, u3));
        if (memcmp(&ex, &f, 8) != 0) {
            if (bad < 10) printf("mismatch a=%a b=%a: %a vs %a\n", a, b, ex, f);
            ++bad;
        }
    }
    printf("%ld\n", bad);
    return bad != 0;
}
