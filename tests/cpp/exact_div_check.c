/* Brute-force check of the division used for the corner gradients on the
 * device: x / c computed as q = x * rc, e = fma(-q, c, x), q' = fma(e, rc, q)
 * with rc = 1 / c correctly rounded, taken only for |q| in [2^-960, 2^960]
 * (elsewhere, and for zeros, infinities and NaN, the division itself).
 * Every result must equal x / c bit for bit.
 *   exact_div_check <samples> <seed> c1 [c2 ...]   (c = "rand": a random divisor per sample)
 * prints the number of mismatches (0) and exits 1 on any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t sm(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double div_by(double x, double c, double rc) {
    const double q = x * rc;
    const double e = fma(-q, c, x);
    const double q2 = fma(e, rc, q);
    const double aq = fabs(q);
    return (rc != 0.0 && aq > 0x1p-960 && aq < 0x1p+960) ? q2 : x / c;
}

static int same(double a, double b) { return memcmp(&a, &b, sizeof a) == 0; }

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    uint64_t s = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
    long bad = 0, tested = 0;
    for (int a = 3; a < argc || a == 3; a++) {
        /* "rand": a new divisor per sample, 2^-40 .. 2^40, either sign */
        const int rnd = a < argc && strcmp(argv[a], "rand") == 0;
        double c = a < argc && !rnd ? atof(argv[a]) : 0.1;
        double rc = 1.0 / c;
        for (long i = 0; i < n; i++) {
            if (rnd) {
                const uint64_t t = sm(&s);
                c = ldexp(1.0 + (double)(t >> 12) * 0x1p-52, (int)((t >> 1) % 81) - 40) * ((t & 1) ? -1.0 : 1.0);
                rc = 1.0 / c;
            }
            const uint64_t r = sm(&s);
            double x;
            switch (r & 3) {
            case 0: {  /* difference of two floats, as the gradients are */
                float f0, f1;
                uint32_t b0 = (uint32_t)(r >> 2), b1 = (uint32_t)(r >> 34);
                memcpy(&f0, &b0, 4);
                memcpy(&f1, &b1, 4);
                x = (double)f1 - (double)f0;
                break;
            }
            case 1: {  /* moderate floats */
                float f0 = (float)((double)(r >> 11) * 0x1p-53 * 8.0 - 4.0);
                float f1 = (float)((double)(sm(&s) >> 11) * 0x1p-53 * 8.0 - 4.0);
                x = (double)f1 - (double)f0;
                break;
            }
            case 2: {  /* any double bit pattern */
                memcpy(&x, &r, 8);
                break;
            }
            default: {  /* a float minus a double iso-like value */
                float f0 = (float)((double)(r >> 11) * 0x1p-53 * 2.0 - 1.0);
                x = (double)f0 - 0.3;
                break;
            }
            }
            const double want = x / c, got = div_by(x, c, rc);
            tested++;
            if (!same(want, got) && !(isnan(want) && isnan(got))) {
                if (bad < 5) fprintf(stderr, "mismatch x=%a c=%a want=%a got=%a\n", x, c, want, got);
                bad++;
            }
        }
    }
    /* special values */
    const double sp[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 0x1p-1074, -0x1p-1074, 0x1p-1022, 0x1.fffffffffffffp+1023};
    for (unsigned i = 0; i < sizeof sp / sizeof sp[0]; i++) {
        const double c = 0.049087385212340517, rc = 1.0 / c;
        const double want = sp[i] / c, got = div_by(sp[i], c, rc);
        tested++;
        if (!same(want, got) && !(isnan(want) && isnan(got))) bad++;
    }
    printf("%ld of %ld differ\n", bad, tested);
    return bad != 0;
}
