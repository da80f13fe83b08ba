/*
 * sw2_xdiv.h -- correctly rounded FP64 division for the GPU step.
 *
 * SW2's step divides in FP64 on the CPU, where every quotient is the IEEE
 * correctly rounded one.  The division sequence the GPU compiler emits for
 * gfx950 (reciprocal, two Newton steps, one residual correction) is faithful
 * -- within one ulp -- but not always the nearest double: near the bottom of
 * the exponent range it has been seen to return the other neighbour.  The
 * step therefore divides only through the two functions below.
 *
 * ddiv(a, b)    any a, b.  The hardware quotient q is a faithful candidate;
 *               the exact remainders a - b*q and a - b*q' (fma) of q and of
 *               its neighbour q' toward a/b decide which one is nearer, ties
 *               to even.  Operands or quotients outside [2^-800, 2^800] go to
 *               ddiv_rare: IEEE specials (zeros, infinities, NaN) are exact in
 *               the hardware sequence; other values are scaled into [0.5, 1)
 *               with frexp, divided there and scaled back, which is exact
 *               for a normal quotient; a subnormal quotient is rounded to the
 *               2^-1074 grid by exact remainders of the integer candidates.
 *
 * xdiv(a, b, y) a divisor b > 0 known ahead with y = RN(1/b) computed on the
 *               host (the per-row metric constants).  One product and two
 *               residual corrections:
 *                   q0 = RN(a*y)            within 1.5 ulp of a/b
 *                   q1 = RN(q0 + r0*y)      r0 = RN(a - b*q0): faithful
 *                   q2 = RN(q1 + r1*y)      r1 = a - b*q1 exact: = RN(a/b)
 *               by Markstein's theorem (y within 1/2 ulp of 1/b and q1
 *               within 1 ulp => RN(q1 + r1*y) = RN(a/b); Markstein 1990,
 *               Muller et al., Handbook of Floating-Point Arithmetic, ch. 4),
 *               valid while nothing underflows or overflows; quotients
 *               outside [2^-800, 2^800] take ddiv_rare.  Divisors must lie in
 *               [SW2_XDIV_BMIN, SW2_XDIV_BMAX]; sessions check their tables.
 *
 * Multiply-add appears only as explicit fma, exact by definition; no
 * expression of SW2's is contracted.  SW2_HW_DIV(a, b) is the candidate
 * division; a checker may define it to a deliberately perturbed faithful
 * quotient to exercise the correction logic on the CPU.
 */
#ifndef SW2_XDIV_H
#define SW2_XDIV_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define SW2_HD __host__ __device__ __forceinline__
#else
#define SW2_HD static inline
#endif

#ifndef SW2_HW_DIV
#define SW2_HW_DIV(a, b) ((a) / (b))
#endif

#define SW2_XDIV_BMIN 0x1p-60
#define SW2_XDIV_BMAX 0x1p+60

SW2_HD uint64_t sw2_bits(double x) {
    uint64_t u;
    memcpy(&u, &x, sizeof u);
    return u;
}
SW2_HD double sw2_from_bits(uint64_t u) {
    double x;
    memcpy(&x, &u, sizeof x);
    return x;
}

/* a, b, q finite with magnitudes in [2^-800, 2^800], q within one ulp of
 * a/b: the nearer of q and its neighbour toward a/b (ties to even).  No
 * branch: when q is exact (r == 0) the neighbour's remainder is nonzero and
 * q is kept by the comparison. */
SW2_HD double sw2_nearest(double a, double b, double q) {
    double r = fma(-b, q, a);
    int up = (r > 0) == (b > 0); /* a/b lies above q */
    uint64_t u = sw2_bits(q);
    uint64_t un = ((q > 0) == up) ? u + 1 : u - 1;
    double qn = sw2_from_bits(un);
    double rn = fma(-b, qn, a);
    double ar = fabs(r), arn = fabs(rn);
    if (arn < ar || (arn == ar && (un & 1) == 0)) return qn;
    return q;
}

/* every case ddiv's fast path does not take */
#if defined(__HIPCC__)
static __host__ __device__ __noinline__
#else
static
#endif
double sw2_ddiv_rare(double a, double b) {
    if (a == 0 || b == 0 || a != a || b != b || fabs(a) == INFINITY || fabs(b) == INFINITY) return SW2_HW_DIV(a, b);
    int ea, eb;
    double ma = frexp(a, &ea), mb = frexp(b, &eb); /* |ma|, |mb| in [0.5, 1) */
    int k = ea - eb;
    double t = sw2_nearest(ma, mb, SW2_HW_DIV(ma, mb)); /* RN(ma/mb), |t| in [0.5, 2] */
    int et;
    (void)frexp(t, &et);
    int e = et + k; /* |RN(a/b)| in [2^(e-1), 2^e) if normal */
    if (e > 1024) return t > 0 ? INFINITY : -INFINITY;
    if (e - 1 >= -1022) return ldexp(t, k);
    /* subnormal or zero: n = round-half-even(|a/b| * 2^1074) */
    const int s = k + 1074;
    const double A = fabs(ma), B = fabs(mb);
    double n;
    if (s < -1) {
        n = 0; /* |a/b| * 2^1074 < 1/2 */
    } else if (s == -1) {
        n = A > B ? 1.0 : 0.0; /* in (1/4, 1); exactly 1/2 when A == B: even */
    } else {
        const double As = ldexp(A, s); /* exact */
        const double n0 = rint(sw2_nearest(As, B, SW2_HW_DIV(As, B)));
        n = n0;
        double best = fabs(fma(-B, n0, As));
        for (int d = -1; d <= 1; d += 2) {
            double c = n0 + d;
            if (c < 0) continue;
            double rc = fabs(fma(-B, c, As)); /* exact: multiple of 2^-53 below 1 */
            if (rc < best || (rc == best && fmod(c, 2.0) == 0)) {
                best = rc;
                n = c;
            }
        }
    }
    double res = ldexp(n, -1074);
    return ((a < 0) != (b < 0)) ? -res : res;
}

/* correctly rounded a / b */
SW2_HD double ddiv(double a, double b) {
    double q = SW2_HW_DIV(a, b);
    double aa = fabs(a), ab = fabs(b), aq = fabs(q);
    if (aa >= 0x1p-800 && aa <= 0x1p+800 && ab >= 0x1p-800 && ab <= 0x1p+800 && aq >= 0x1p-800 && aq <= 0x1p+800)
        return sw2_nearest(a, b, q);
    if (a == 0) return q; /* +-0 (or NaN when b is 0 or NaN): exact in hardware */
    return sw2_ddiv_rare(a, b);
}

/* correctly rounded a / b for a divisor b known ahead, y = RN(1/b) */
SW2_HD double xdiv(double a, double b, double y) {
    double q = a * y;
    double aq = fabs(q);
    if (!(aq >= 0x1p-800 && aq <= 0x1p+800)) return (a == 0) ? q : sw2_ddiv_rare(a, b);
    double r = fma(-q, b, a);
    q = fma(r, y, q);
    r = fma(-q, b, a);
    return fma(r, y, q);
}

#endif /* SW2_XDIV_H */
