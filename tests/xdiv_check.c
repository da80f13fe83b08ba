/*
 * CPU check of the GPU step's division (sw2_amd/csrc/sw2_xdiv.h): ddiv(a, b)
 * and xdiv(a, b, RN(1/b)) must equal the IEEE correctly rounded a/b of this
 * CPU, bit for bit, for every sample.
 *
 * The candidate division inside the header is replaced here by a quotient
 * that is only faithful: with probability 1/2 it is the neighbour of the
 * correctly rounded quotient on the far side of a/b, the worst a faithful
 * hardware sequence can return.  Samples mix dividends and divisors over the
 * whole exponent range (subnormals, quotients that underflow, overflow or
 * sit at a binade edge, near-midpoint quotients, specials).
 *
 *   xdiv_check <seed> <n> [divisors.bin]
 *
 * divisors.bin (optional) holds positive divisors to test xdiv and ddiv with
 * besides random ones.  Prints "checked N mismatches M"; exit 1 on mismatch.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st;
static uint64_t next(void) { /* splitmix64 */
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* a faithful but not always nearest quotient */
static double faithful_div(double a, double b) {
    double q = a / b;
    if (!isfinite(q) || q == 0 || fabs(q) < 0x1p-1022) return q;
    double r = fma(-b, q, a);
    if (r == 0 || !(next() & 1)) return q;
    return nextafter(q, ((r > 0) == (b > 0)) ? INFINITY : -INFINITY);
}
#define SW2_HW_DIV(a, b) faithful_div((a), (b))
#include "../sw2_amd/csrc/sw2_xdiv.h"

static double bits(uint64_t u) {
    double d;
    memcpy(&d, &u, 8);
    return d;
}
static uint64_t ubits(double d) {
    uint64_t u;
    memcpy(&u, &d, 8);
    return u;
}
/* random double with unbiased exponent in [e0, e1] (subnormal below -1022) */
static double rnd(int e0, int e1, int sign) {
    int e = e0 + (int)(next() % (uint64_t)(e1 - e0 + 1));
    double m = bits((1023ull << 52) | (next() & ((1ull << 52) - 1))); /* [1, 2) */
    double v = ldexp(m, e);
    return (sign && (next() & 1)) ? -v : v;
}

static long nchk = 0, nbad = 0;
static void report(const char *what, double a, double b, double got, double want) {
    if (nbad < 12) fprintf(stderr, "%s mismatch a=%a b=%a got=%a ieee=%a\n", what, a, b, got, want);
    nbad++;
}
static int same(double x, double y) { return ubits(x) == ubits(y) || (isnan(x) && isnan(y)); }
static void check(double a, double b) {
    double t = a / b;
    nchk++;
    double q = ddiv(a, b);
    if (!same(q, t)) report("ddiv", a, b, q, t);
    if (b > 0 && b >= SW2_XDIV_BMIN && b <= SW2_XDIV_BMAX) {
        double x = xdiv(a, b, 1.0 / b);
        if (!same(x, t)) report("xdiv", a, b, x, t);
    }
}

static double dividend_for(double b, int kind) {
    switch (kind) {
    case 0: return rnd(-1074, 1023, 1);                                  /* anywhere */
    case 1: return rnd(-40, 40, 1);                                      /* physical */
    case 2: return rnd(-60, 60, 1) * b;                                  /* quotient near a double */
    case 3: return nextafter(rnd(-60, 60, 1) * b, next() & 1 ? INFINITY : -INFINITY);
    case 4: { /* quotient near a rounding midpoint */
        double q = rnd(-60, 60, 0);
        double h = (nextafter(q, INFINITY) - q) * 0.5;
        double a = fma(b, h, q * b);
        return (next() & 1) ? -a : a;
    }
    case 5: return ldexp(next() & 1 ? 0x1.fffffffffffffp0 : 0x1.0000000000001p0, (int)(next() % 80) - 40) * b;
    case 6: return ldexp(rnd(0, 0, 1), -1022 - (int)(next() % 60)) * (fabs(b) > 1 ? 1.0 : fabs(b)); /* tiny */
    case 7: return rnd(-1074, -1000, 1) * fabs(b);                         /* quotient in the subnormal range */
    case 8: return rnd(-1000, -900, 1);                                     /* near the fast-path edge */
    default: return rnd(1000, 1023, 1);                                     /* huge */
    }
}

static void run_divisor(double b, long n) {
    for (long k = 0; k < n; k++) check(dividend_for(b, (int)(k % 10)), b);
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    st = strtoull(argv[1], NULL, 10);
    long n = atol(argv[2]);
    const double specials[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 0x1p-1074, -0x1p-1074, 0x1p-1022, -0x1p-1022, 0x1p-950,
                               0x1.fffffffffffffp1023, -0x1.fffffffffffffp1023, 1e300, 1.0, -3.0, 0x1p-800, 0x1p800};
    const int ns = sizeof specials / sizeof *specials;
    for (int i = 0; i < ns; i++)
        for (int k = 0; k < ns; k++) check(specials[i], specials[k]);
    const double sdiv[] = {1.0, 3.0, -10.0, 0x1.fffffffffffffp0, 0x1.0000000000001p0, 7.1492e7, 111194.92664455873,
                           0x1p-60, 0x1p60, 0x1p-1074, 0x1.8p-1030, 1e-300, -1e300, 0x1p1023};
    for (unsigned i = 0; i < sizeof sdiv / sizeof *sdiv; i++) run_divisor(sdiv[i], n);
    for (int i = 0; i < 64; i++) run_divisor(rnd(-1074, 1023, 1), n / 4 + 1);
    for (int i = 0; i < 64; i++) run_divisor(rnd(-30, 40, 0), n / 2 + 1);
    if (argc > 3) {
        FILE *f = fopen(argv[3], "rb");
        if (!f) return 2;
        double b;
        while (fread(&b, sizeof b, 1, f) == 1) run_divisor(b, n / 16 + 1);
        fclose(f);
    }
    printf("checked %ld mismatches %ld\n", nchk, nbad);
    return nbad != 0;
}
