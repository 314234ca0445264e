/* Division through the correctly rounded reciprocal, as the ordered GPU kernels divide
 * by a distance or a mass: y = 1/b (IEEE division), q0 = a*y, then two corrections
 * q <- fma(fma(-q, b, a), y, q).  Counts quotients that differ from the IEEE a/b over
 * random operands and adversarial families.  Built with -ffp-contract=off; fma() is the
 * C library's fused multiply-add.  Prints "<cases> <mismatches>". */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t state = 0x2545F4914F6CDD1Dull;
static uint64_t next(void)
{
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return state;
}
static double from_bits(uint64_t b)
{
    double d;
    memcpy(&d, &b, 8);
    return d;
}
static double rnd(int emin, int emax)
{
    const int e = emin + (int)(next() % (uint64_t)(emax - emin + 1));
    const double d = from_bits(((uint64_t)(e + 1023) << 52) | (next() & ((1ull << 52) - 1)));
    return (next() & 1) ? -d : d;
}
static double div_rcp(double a, double b, double y)
{
    const double q0 = a * y;
    const double q1 = fma(fma(-q0, b, a), y, q0);
    return fma(fma(-q1, b, a), y, q1);
}

int main(int argc, char** argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 10000000L;
    long bad = 0;
    for (long k = 0; k < n; k++) {
        double a, b;
        switch (k % 4) {
        case 0:  /* random over 2^+-300 */
            a = rnd(-300, 300);
            b = rnd(-300, 300);
            break;
        case 1:  /* divisors with (nearly) all-ones significands */
            a = rnd(-200, 200);
            b = from_bits(((uint64_t)(1023 + (int)(next() % 400) - 200) << 52) | (((1ull << 52) - 1) - (next() & 0xff)));
            break;
        case 2:  /* dividends with (nearly) all-ones significands */
            b = rnd(-200, 200);
            a = from_bits(((uint64_t)(1023 + (int)(next() % 400) - 200) << 52) | (((1ull << 52) - 1) - (next() & 0xff)));
            break;
        default: /* exact quotients and their neighbours */
            b = rnd(-100, 100);
            a = (double)(int64_t)(next() % 2000001 - 1000000) * b;
            if (next() & 1) a = nextafter(a, INFINITY);
            break;
        }
        if (a == 0 || !isfinite(a)) continue;
        const double y = 1.0 / b;
        const double q = div_rcp(a, b, y);
        const double r = a / b;
        if (memcmp(&q, &r, 8) != 0) bad++;
    }
    printf("%ld %ld\n", n, bad);
    return 0;
}
