/*
 * wq3d_math.h - FP64 elementary functions with one definition for host and
 * device code.
 *
 * The wavefront kernels (HIP, gfx950) and the CPU checker evaluate sin, cos,
 * exp, log, log10, atan, atan2, asin, 10^x and a few complex operations
 * through these functions, so both evaluate exactly the same sequence of
 * IEEE-754 operations and agree bit for bit.  Only correctly rounded
 * primitives are used underneath (+ - * /, sqrt, fma, floor, fabs, ldexp,
 * frexp, copysign), which every conforming host and the AMDGPU back end
 * implement identically.
 *
 * Methods (all from the mathematics; constants derived with exact decimal
 * arithmetic, hex literals below):
 *  - exp: Cody-Waite reduction x = k ln2 + r with a 32-bit ln2 head
 *    (k ln2_hi exact for |k| < 2^21), |r| <= ln2/2, Taylor series of e^r to
 *    degree 13 in Horner form with fma (truncation < 2^-57), 2^k by ldexp.
 *  - log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 *    s = (m-1)/(m+1), odd series in s to s^23 arranged as f - s (f - 2T);
 *    e ln2 with split constant.
 *  - sin/cos: Cody-Waite reduction by pi/2 with a three-part 33-bit split
 *    (exact products for |k| < 2^20, i.e. |x| < 1.6e6), Taylor kernels on
 *    [-pi/4, pi/4] to degree 19 / 20, quadrant selection.
 *  - atan: reciprocal for |x| > 1, then atan t = atan(k/4) + atan u with
 *    u = (t - k/4) / (1 + t k/4), |u| <= 1/8, odd Taylor series to u^21;
 *    atan2 and asin from atan with quadrant rules.
 * Accuracy is checked against the host C library in tests/test_math.py
 * (a few ulp at most over the argument ranges the hot path uses).
 */
#ifndef WQ3D_MATH_H
#define WQ3D_MATH_H

#include <math.h>
#if defined(WQ3D_EMULATE) && !defined(__HIP_DEVICE_COMPILE__)
#include <stdio.h>
#include <stdlib.h>
#endif

#if defined(__HIPCC__) || defined(__HIP__)
#define WQM_HD __host__ __device__ inline
#else
#define WQM_HD inline
#endif

namespace wqm {

// ln 2 split: head with 32 significant bits, tail
constexpr double LN2_HI = 0x1.62e42ffp-1;
constexpr double LN2_LO = -0x1.718432a1b0e26p-35;
constexpr double INV_LN2 = 0x1.71547652b82fep+0;
// pi/2 in three 33-bit pieces, and pi, pi/2 as head + tail
constexpr double PIO2_1 = 0x1.921fb544p+0;
constexpr double PIO2_2 = 0x1.0b4611a6p-34;
constexpr double PIO2_3 = 0x1.3198a2e037073p-69;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double PIO2_HI = 0x1.921fb54442d18p+0;
constexpr double PIO2_LO = 0x1.1a62633145c07p-54;
constexpr double PI_HI = 0x1.921fb54442d18p+1;
constexpr double PI_LO = 0x1.1a62633145c07p-53;
constexpr double LN10_HI = 0x1.26bb1bbb55516p+1;
constexpr double LN10_LO = -0x1.f48ad494ea3e9p-53;
constexpr double LOG10E_HI = 0x1.bcb7b1526e50ep-2;
constexpr double LOG10E_LO = 0x1.95355baaafad3p-57;
constexpr double SQRT_HALF = 0x1.6a09e667f3bcdp-1;

WQM_HD bool isnan_(double x) { return x != x; }

// e^x
WQM_HD double exp(double x) {
    if (isnan_(x)) return x;
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -745.1332191019412) return 0.0;
    const double k = ::floor(x * INV_LN2 + 0.5);
    double r = ::fma(-k, LN2_HI, x);
    r = ::fma(-k, LN2_LO, r);
    double p = 0x1.6124613a86d09p-33;    // 1/13!
    p = ::fma(p, r, 0x1.1eed8eff8d898p-29);  // 1/12!
    p = ::fma(p, r, 0x1.ae64567f544e4p-26);  // 1/11!
    p = ::fma(p, r, 0x1.27e4fb7789f5cp-22);  // 1/10!
    p = ::fma(p, r, 0x1.71de3a556c734p-19);  // 1/9!
    p = ::fma(p, r, 0x1.a01a01a01a01ap-16);  // 1/8!
    p = ::fma(p, r, 0x1.a01a01a01a01ap-13);  // 1/7!
    p = ::fma(p, r, 0x1.6c16c16c16c17p-10);  // 1/6!
    p = ::fma(p, r, 0x1.1111111111111p-7);   // 1/5!
    p = ::fma(p, r, 0x1.5555555555555p-5);   // 1/4!
    p = ::fma(p, r, 0x1.5555555555555p-3);   // 1/3!
    p = ::fma(p, r, 0.5);
    p = ::fma(p, r, 1.0);
    p = ::fma(p, r, 1.0);
    return ::ldexp(p, (int)k);
}

// natural logarithm
WQM_HD double log(double x) {
    if (isnan_(x) || x < 0.0) return NAN;
    if (x == 0.0) return -HUGE_VAL;
    if (x == HUGE_VAL) return x;
    int e = 0;
    double m = ::frexp(x, &e);
    if (m < SQRT_HALF) {
        m = m * 2.0;
        e -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    // 2 atanh(s) = 2s (1 + z/3 + z^2/5 + ... + z^11/23)
    double p = 0x1.642c8590b2164p-5;        // 1/23
    p = ::fma(p, z, 0x1.8618618618618p-5);  // 1/21
    p = ::fma(p, z, 0x1.af286bca1af28p-5);  // 1/19
    p = ::fma(p, z, 0x1.e1e1e1e1e1e1ep-5);  // 1/17
    p = ::fma(p, z, 0x1.1111111111111p-4);  // 1/15
    p = ::fma(p, z, 0x1.3b13b13b13b14p-4);  // 1/13
    p = ::fma(p, z, 0x1.745d1745d1746p-4);  // 1/11
    p = ::fma(p, z, 0x1.c71c71c71c71cp-4);  // 1/9
    p = ::fma(p, z, 0x1.2492492492492p-3);  // 1/7
    p = ::fma(p, z, 0x1.999999999999ap-3);  // 1/5
    p = ::fma(p, z, 0x1.5555555555555p-2);  // 1/3
    // 2 atanh(s) = 2s + 2s z P(z) and 2s = f - s f, so log m = f - s (f - 2 z P(z)):
    // the exact leading term f carries the result, the rounded part is O(f^2)
    const double t = z * p;
    const double logm = ::fma(-s, f - (t + t), f);
    const double de = (double)e;
    return ::fma(de, LN2_HI, ::fma(de, LN2_LO, logm));
}

WQM_HD double log10(double x) {
    const double t = log(x);
    return ::fma(t, LOG10E_HI, t * LOG10E_LO);
}

// 10^x
WQM_HD double pow10(double x) {
    const double hi = x * LN10_HI;
    const double lo = ::fma(x, LN10_HI, -hi) + x * LN10_LO;
    const double e = exp(hi);
    return ::fma(e, lo, e);
}

// reduction by pi/2: returns r and the quadrant q = k mod 4
WQM_HD double reduce_pio2(double x, int* q) {
#if defined(WQ3D_EMULATE) && !defined(__HIP_DEVICE_COMPILE__)
    if (!(x > -1.0e6 && x < 1.0e6)) {
        fprintf(stderr, "wq3d emulation check failed in reduce_pio2: %.17g\n", x);
        abort();
    }
#endif
    const double k = ::floor(x * TWO_OVER_PI + 0.5);
    double r = ::fma(-k, PIO2_1, x);
    r = ::fma(-k, PIO2_2, r);
    r = ::fma(-k, PIO2_3, r);
    // k mod 4 in double (exact below 2^53), so the integer conversion is of
    // a value in [0, 4) for every finite k; arguments this large are far
    // outside the reduction's accurate range anyway
    const double k4 = k - 4.0 * ::floor(k * 0.25);
    *q = (k4 >= 0.0 && k4 < 4.0) ? (int)k4 : 0;
    return r;
}

WQM_HD double sin_kernel(double r) {
    // sin r = r + r z P(z), P = -1/3! + z/5! - z^2/7! + ... - z^8/19!
    const double z = r * r;
    double p = -0x1.2f49b46814157p-57;           // -1/19!
    p = ::fma(p, z, 0x1.952c77030ad4ap-49);      // 1/17!
    p = ::fma(p, z, -0x1.ae7f3e733b81fp-41);     // -1/15!
    p = ::fma(p, z, 0x1.6124613a86d09p-33);      // 1/13!
    p = ::fma(p, z, -0x1.ae64567f544e4p-26);     // -1/11!
    p = ::fma(p, z, 0x1.71de3a556c734p-19);      // 1/9!
    p = ::fma(p, z, -0x1.a01a01a01a01ap-13);     // -1/7!
    p = ::fma(p, z, 0x1.1111111111111p-7);       // 1/5!
    p = ::fma(p, z, -0x1.5555555555555p-3);      // -1/3!
    return ::fma(r * z, p, r);
}

WQM_HD double cos_kernel(double r) {
    const double z = r * r;
    double p = 0x1.e542ba4020225p-62;            // 1/20!
    p = ::fma(p, z, -0x1.6827863b97d97p-53);     // -1/18!
    p = ::fma(p, z, 0x1.ae7f3e733b81fp-45);      // 1/16!
    p = ::fma(p, z, -0x1.93974a8c07c9dp-37);     // -1/14!
    p = ::fma(p, z, 0x1.1eed8eff8d898p-29);      // 1/12!
    p = ::fma(p, z, -0x1.27e4fb7789f5cp-22);     // -1/10!
    p = ::fma(p, z, 0x1.a01a01a01a01ap-16);      // 1/8!
    p = ::fma(p, z, -0x1.6c16c16c16c17p-10);     // -1/6!
    p = ::fma(p, z, 0x1.5555555555555p-5);       // 1/4!
    p = ::fma(p, z, -0.5);                       // -1/2!
    return ::fma(z, p, 1.0);
}

WQM_HD double sin(double x) {
    if (isnan_(x) || x == HUGE_VAL || x == -HUGE_VAL) return NAN;
    int q;
    const double r = reduce_pio2(x, &q);
    switch (q) {
        case 0: return sin_kernel(r);
        case 1: return cos_kernel(r);
        case 2: return -sin_kernel(r);
        default: return -cos_kernel(r);
    }
}

WQM_HD double cos(double x) {
    if (isnan_(x) || x == HUGE_VAL || x == -HUGE_VAL) return NAN;
    int q;
    const double r = reduce_pio2(x, &q);
    switch (q) {
        case 0: return cos_kernel(r);
        case 1: return -sin_kernel(r);
        case 2: return -cos_kernel(r);
        default: return sin_kernel(r);
    }
}

// arctangent on [0, 1]: atan t = atan c + atan u, c = k/4 nearest to t,
// u = (t - c) / (1 + t c), |u| <= 1/8, odd Taylor series of atan u to u^21;
// atan(k/4) as head + tail computed in exact decimal arithmetic
WQM_HD double atan_unit(double t) {
    const double AT_HI[5] = {0.0, 0x1.f5b75f92c80ddp-3, 0x1.dac670561bb4fp-2,
                             0x1.4978fa3269ee1p-1, 0x1.921fb54442d18p-1};
    const double AT_LO[5] = {0.0, 0x1.8ab6e3cf7afbdp-57, 0x1.a2b7f222f65e2p-56,
                             0x1.2419a87f2a458p-56, 0x1.1a62633145c07p-55};
#if defined(WQ3D_EMULATE) && !defined(__HIP_DEVICE_COMPILE__)
    if (!(t >= 0.0 && t <= 1.0)) {
        fprintf(stderr, "wq3d emulation check failed in atan_unit: %.17g\n", t);
        abort();
    }
#endif
    // the table index in double first: t is in [0, 1] for every caller, and
    // a NaN t (atan2 of two infinities) yields NaN, never an index outside
    // the table
    if (!(t >= 0.0 && t <= 1.0)) return t - t;
    const int k = (int)::floor(t * 4.0 + 0.5);
    const double c = 0.25 * (double)k;
    const double u = (t - c) / ::fma(t, c, 1.0);
    const double z = u * u;
    double p = -0x1.8618618618618p-5;            // -1/21
    p = ::fma(p, z, 0x1.af286bca1af28p-5);       // 1/19
    p = ::fma(p, z, -0x1.e1e1e1e1e1e1ep-5);      // -1/17
    p = ::fma(p, z, 0x1.1111111111111p-4);       // 1/15
    p = ::fma(p, z, -0x1.3b13b13b13b14p-4);      // -1/13
    p = ::fma(p, z, 0x1.745d1745d1746p-4);       // 1/11
    p = ::fma(p, z, -0x1.c71c71c71c71cp-4);      // -1/9
    p = ::fma(p, z, 0x1.2492492492492p-3);       // 1/7
    p = ::fma(p, z, -0x1.999999999999ap-3);      // -1/5
    p = ::fma(p, z, 0x1.5555555555555p-2);       // 1/3
    const double a = ::fma(-u * z, p, u);
    return AT_HI[k] + (AT_LO[k] + a);
}

WQM_HD double atan(double x) {
    if (isnan_(x)) return x;
    const double ax = ::fabs(x);
    double a;
    if (ax > 1.0) {
        a = (PIO2_HI - atan_unit(1.0 / ax)) + PIO2_LO;
    } else {
        a = atan_unit(ax);
    }
    return ::copysign(a, x);
}

WQM_HD double atan2(double y, double x) {
    if (isnan_(x) || isnan_(y)) return x + y;
    const double ax = ::fabs(x), ay = ::fabs(y);
    const bool xneg = ::copysign(1.0, x) < 0.0;
    if (ay == 0.0) return xneg ? ::copysign(PI_HI, y) : ::copysign(0.0, y);
    if (ax == 0.0) return ::copysign(PIO2_HI, y);
    double t;
    if (ay <= ax) {
        t = atan_unit(ay / ax);
    } else {
        t = (PIO2_HI - atan_unit(ax / ay)) + PIO2_LO;
    }
    if (xneg) t = (PI_HI - t) + PI_LO;
    return ::copysign(t, y);
}

WQM_HD double asin(double x) {
    if (isnan_(x) || ::fabs(x) > 1.0) return NAN;
    return atan2(x, ::sqrt((1.0 - x) * (1.0 + x)));
}

WQM_HD double hypot(double x, double y) { return ::sqrt(x * x + y * y); }

// complex arithmetic used by the Rayleigh reflection coefficient
struct cplx {
    double re, im;
};
WQM_HD cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
WQM_HD cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
WQM_HD cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
WQM_HD cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }
WQM_HD cplx cdivs(cplx a, double s) { return {a.re / s, a.im / s}; }
// Smith's algorithm
WQM_HD cplx cdiv(cplx a, cplx b) {
    if (::fabs(b.re) >= ::fabs(b.im)) {
        const double r = b.im / b.re;
        const double d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im;
    const double d = b.im + b.re * r;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
// principal square root
WQM_HD cplx csqrt(cplx z) {
    if (z.re == 0.0 && z.im == 0.0) return {0.0, z.im};
    const double t = ::sqrt((::fabs(z.re) + hypot(z.re, z.im)) * 0.5);
    if (z.re >= 0.0) return {t, z.im / (2.0 * t)};
    return {::fabs(z.im) / (2.0 * t), ::copysign(t, z.im)};
}
WQM_HD double cabs(cplx z) { return hypot(z.re, z.im); }
WQM_HD double carg(cplx z) { return atan2(z.im, z.re); }

}  // namespace wqm

#endif /* WQ3D_MATH_H */
