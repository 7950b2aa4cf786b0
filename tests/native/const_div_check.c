/* Checks that division of a double by 3 or 6 computed as
 *   q0 = x*z;  r = fma(-q0, d, x);  q = fma(r, z, q0)   (z = RN(1/d))
 * with q0 returned when it is zero (keeps the sign of zero)
 * equals the IEEE quotient x/d bit for bit.  Random doubles over the normal
 * range, integers, and x = d*y +- a few ulp (quotients next to rounding
 * boundaries).  Prints the number of mismatches; argument: samples. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t next(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }

static double fastdiv(double x, double d, double z) {
    double q0 = x * z;
    double r = fma(-q0, d, x);
    double q = fma(r, z, q0);
    return q0 == 0.0 ? q0 : q;
}

int main(int argc, char** argv) {
    long n = argc > 1 ? atol(argv[1]) : 10000000;
    const double ds[2] = {3.0, 6.0};
    long bad = 0;
    for (int k = 0; k < 2; ++k) {
        const double d = ds[k], z = 1.0 / d;
        for (long i = 0; i < n; ++i) {
            double x;
            uint64_t b = next();
            switch (i & 3) {
                case 0: {   /* any normal double with |x| >= 2^-1000 */
                    uint64_t e = 23 + (b >> 11) % (2046 - 23);
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
            double a = x / d, f = fastdiv(x, d, z);
            if (memcmp(&a, &f, 8) != 0) {
                if (bad < 10) printf("mismatch d=%g x=%a: %a vs %a\n", d, x, a, f);
                ++bad;
            }
        }
        double zz[2] = {0.0, -0.0};
        for (int t = 0; t < 2; ++t) {
            double a = zz[t] / d, f = fastdiv(zz[t], d, z);
            if (memcmp(&a, &f, 8) != 0) ++bad;
        }
    }
    printf("%ld\n", bad);
    return bad != 0;
}
