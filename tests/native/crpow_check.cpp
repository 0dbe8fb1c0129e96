// Driver for tests/test_crpow.py: evaluates ros3::crpow::pow_cr (the pow of
// the device's SMVGEAR step control, compiled here for the host with the same
// IEEE operations) on pairs read from stdin and prints the results as hex
// floats, one per line; with "--sweep N seed" it compares N seeded random
// pairs against quad-precision powq rounded to double and against libm's pow
// and prints the mismatch counts.
#include <inttypes.h>
#include <math.h>
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ros3_pow.h"

static uint64_t g_state;
static double uniform(double a, double b) {   // splitmix64
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return a + (b - a) * ((double)(z >> 11) * 0x1p-53);
}

int main(int argc, char **argv) {
    if (argc >= 4 && strcmp(argv[1], "--sweep") == 0) {
        const long n = atol(argv[2]);
        g_state = (uint64_t)atoll(argv[3]);
        const float ys[5] = {1.0f / 2, 1.0f / 3, 1.0f / 4, 1.0f / 5, 1.0f / 6};
        long bad = 0, bad_libm = 0;
        for (long i = 0; i < n; ++i) {
            double x, y;
            switch (i % 4) {
            case 0: x = exp(uniform(-70.0, 70.0)); y = (double)ys[(i / 4) % 5]; break;
            case 1: x = exp(uniform(-40.0, 40.0)); y = uniform(-3.0, 3.0); break;
            case 2: x = 1.0 + uniform(-1e-3, 1e-3); y = uniform(-300.0, 300.0); break;
            default: x = ldexp(1.0 + uniform(-1e-9, 1e-9), (int)uniform(-600.0, 600.0)); y = uniform(-1.0, 1.0);
            }
            const double q = (double)powq((__float128)x, (__float128)y);
            if (ros3::crpow::pow_cr(x, y) != q) {
                if (bad < 5) printf("mismatch x=%a y=%a pow_cr=%a powq=%a\n", x, y, ros3::crpow::pow_cr(x, y), q);
                ++bad;
            }
            if (pow(x, y) != q) ++bad_libm;
        }
        printf("n %ld mismatch %ld libm_mismatch %ld\n", n, bad, bad_libm);
        return 0;
    }
    double x, y;
    while (scanf("%lf %lf", &x, &y) == 2) printf("%a\n", ros3::crpow::pow_cr(x, y));
    return 0;
}
