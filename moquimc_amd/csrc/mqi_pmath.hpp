// Portable elementary functions for the bit-exact verification build.
//
// Every function here is evaluated with plain IEEE double additions,
// multiplications and divisions in a fixed order (compile with
// -ffp-contract=off), so a host build (g++, x86-64 SSE2) and a gfx950 build
// (hipcc) return the same bits for the same input.  The float entry points
// evaluate in double and round once, which makes them correctly rounded in
// all but rare cases.  The transport kernel's production build uses the
// device library instead; this file exists so that the GPU transport can be
// compared history by history with the CPU restatement.
//
// Methods (all standard textbook forms, coefficients derived from the
// series themselves):
//   log: x = 2^e m, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1),
//        odd series in s to s^23; ln2 split into a 32-bit head and a tail.
//   exp: x = k ln2 + r, |r| <= ln2/2, Taylor series to r^17, scale by 2^k.
//   sin/cos: x = k pi/2 + r (Cody-Waite, two-part pi/2), Taylor series to r^21.
//   asin: Maclaurin series on |y| <= 1/2; acos by the half-angle identities.
//   sqrt: IEEE double square root (correctly rounded on both targets).
#ifndef MQI_PMATH_HPP
#define MQI_PMATH_HPP

#include <cstdint>
#include <cstring>

#if defined(__HIP__) || defined(__HIPCC__)
#define MQI_HD __host__ __device__
#else
#define MQI_HD
#endif

namespace mqi_pm
{

MQI_HD inline uint64_t
d2u(double x) {
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    return u;
}
MQI_HD inline double
u2d(uint64_t u) {
    double x;
    __builtin_memcpy(&x, &u, 8);
    return x;
}

constexpr double kLn2Hi  = 6.93147180369123816490e-01;   // ln 2, upper 32 bits
constexpr double kLn2Lo  = 1.90821492927058770002e-10;   // ln 2 - kLn2Hi
constexpr double kInvLn2 = 1.44269504088896338700e+00;
constexpr double kPio2Hi = 1.57079632673412561417e+00;   // pi/2, upper 33 bits
constexpr double kPio2Lo = 6.07710050650619224932e-11;   // pi/2 - kPio2Hi
constexpr double kTwoOverPi = 6.36619772367581382433e-01;
constexpr double kPi     = 3.14159265358979311600e+00;
constexpr double kPio2   = 1.57079632679489655800e+00;
constexpr double kSqrt2  = 1.41421356237309514547e+00;

MQI_HD inline double
pm_sqrt(double x) {
    return __builtin_sqrt(x);
}

// Natural logarithm.
MQI_HD inline double
pm_log(double x) {
    if (x != x) return x;
    if (x <= 0.0) return (x == 0.0) ? -__builtin_huge_val() : __builtin_nan("");
    if (x == __builtin_huge_val()) return x;
    int      e0   = 0;
    uint64_t bits = d2u(x);
    if ((bits >> 52) == 0) {   // subnormal: scale into the normal range
        x *= 18014398509481984.0;   // 2^54
        bits = d2u(x);
        e0   = -54;
    }
    int    e = (int) ((bits >> 52) & 0x7ff) - 1023 + e0;
    double m = u2d((bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);   // [1,2)
    if (m > kSqrt2) {
        m = m * 0.5;
        e = e + 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    // sum_{k=0..11} z^k / (2k+1)
    double p = 1.0 / 23.0;
    p        = p * z + 1.0 / 21.0;
    p        = p * z + 1.0 / 19.0;
    p        = p * z + 1.0 / 17.0;
    p        = p * z + 1.0 / 15.0;
    p        = p * z + 1.0 / 13.0;
    p        = p * z + 1.0 / 11.0;
    p        = p * z + 1.0 / 9.0;
    p        = p * z + 1.0 / 7.0;
    p        = p * z + 1.0 / 5.0;
    p        = p * z + 1.0 / 3.0;
    p        = p * z + 1.0;
    const double r  = 2.0 * s * p;
    const double de = (double) e;
    return de * kLn2Hi + (de * kLn2Lo + r);
}

// Multiply by 2^k for k in [-1074-60, 1023+60] without overflow in the scale.
MQI_HD inline double
pm_ldexp(double y, int k) {
    while (k > 1000) {
        y = y * u2d((uint64_t) (1000 + 1023) << 52);
        k -= 1000;
    }
    while (k < -1000) {
        y = y * u2d((uint64_t) (-1000 + 1023) << 52);
        k += 1000;
    }
    return y * u2d((uint64_t) (k + 1023) << 52);
}

// Exponential.
MQI_HD inline double
pm_exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return __builtin_huge_val();
    if (x < -745.1332191019412) return 0.0;
    const double kf = __builtin_floor(x * kInvLn2 + 0.5);
    const int    k  = (int) kf;
    const double r  = (x - kf * kLn2Hi) - kf * kLn2Lo;
    // e^r = 1 + r(1 + r/2(1 + r/3(... (1 + r/17))))
    double p = 1.0 + r * (1.0 / 17.0);
    p        = 1.0 + r * (1.0 / 16.0) * p;
    p        = 1.0 + r * (1.0 / 15.0) * p;
    p        = 1.0 + r * (1.0 / 14.0) * p;
    p        = 1.0 + r * (1.0 / 13.0) * p;
    p        = 1.0 + r * (1.0 / 12.0) * p;
    p        = 1.0 + r * (1.0 / 11.0) * p;
    p        = 1.0 + r * (1.0 / 10.0) * p;
    p        = 1.0 + r * (1.0 / 9.0) * p;
    p        = 1.0 + r * (1.0 / 8.0) * p;
    p        = 1.0 + r * (1.0 / 7.0) * p;
    p        = 1.0 + r * (1.0 / 6.0) * p;
    p        = 1.0 + r * (1.0 / 5.0) * p;
    p        = 1.0 + r * (1.0 / 4.0) * p;
    p        = 1.0 + r * (1.0 / 3.0) * p;
    p        = 1.0 + r * (1.0 / 2.0) * p;
    p        = 1.0 + r * p;
    return pm_ldexp(p, k);
}

// x^y for the arguments the transport uses (x >= 0, finite y).
MQI_HD inline double
pm_pow(double x, double y) {
    if (y == 0.0) return 1.0;
    if (x != x || y != y) return x + y;
    if (x == 0.0) return (y < 0.0) ? __builtin_huge_val() : 0.0;
    if (x < 0.0) return __builtin_nan("");
    if (x == 1.0) return 1.0;
    return pm_exp(y * pm_log(x));
}

// sin and cos of |x| < ~1e5 together.
MQI_HD inline void
pm_sincos(double x, double* sn, double* cs) {
    if (x != x || x == __builtin_huge_val() || x == -__builtin_huge_val()) {
        *sn = __builtin_nan("");
        *cs = __builtin_nan("");
        return;
    }
    const double kf = __builtin_floor(x * kTwoOverPi + 0.5);
    const double r  = (x - kf * kPio2Hi) - kf * kPio2Lo;
    const double z  = r * r;
    // sin r = r(1 - z/(2*3)(1 - z/(4*5)(1 - ...)))
    double s = 1.0 - z * (1.0 / (20.0 * 21.0));
    s        = 1.0 - z * (1.0 / (18.0 * 19.0)) * s;
    s        = 1.0 - z * (1.0 / (16.0 * 17.0)) * s;
    s        = 1.0 - z * (1.0 / (14.0 * 15.0)) * s;
    s        = 1.0 - z * (1.0 / (12.0 * 13.0)) * s;
    s        = 1.0 - z * (1.0 / (10.0 * 11.0)) * s;
    s        = 1.0 - z * (1.0 / (8.0 * 9.0)) * s;
    s        = 1.0 - z * (1.0 / (6.0 * 7.0)) * s;
    s        = 1.0 - z * (1.0 / (4.0 * 5.0)) * s;
    s        = 1.0 - z * (1.0 / (2.0 * 3.0)) * s;
    s        = r * s;
    // cos r = 1 - z/(1*2)(1 - z/(3*4)(1 - ...))
    double c = 1.0 - z * (1.0 / (21.0 * 22.0));
    c        = 1.0 - z * (1.0 / (19.0 * 20.0)) * c;
    c        = 1.0 - z * (1.0 / (17.0 * 18.0)) * c;
    c        = 1.0 - z * (1.0 / (15.0 * 16.0)) * c;
    c        = 1.0 - z * (1.0 / (13.0 * 14.0)) * c;
    c        = 1.0 - z * (1.0 / (11.0 * 12.0)) * c;
    c        = 1.0 - z * (1.0 / (9.0 * 10.0)) * c;
    c        = 1.0 - z * (1.0 / (7.0 * 8.0)) * c;
    c        = 1.0 - z * (1.0 / (5.0 * 6.0)) * c;
    c        = 1.0 - z * (1.0 / (3.0 * 4.0)) * c;
    c        = 1.0 - z * (1.0 / (1.0 * 2.0)) * c;
    const int q = ((int) (kf - 4.0 * __builtin_floor(kf * 0.25))) & 3;
    if (q == 0) {
        *sn = s;
        *cs = c;
    } else if (q == 1) {
        *sn = c;
        *cs = -s;
    } else if (q == 2) {
        *sn = -s;
        *cs = -c;
    } else {
        *sn = -c;
        *cs = s;
    }
}

MQI_HD inline double
pm_sin(double x) {
    double s, c;
    pm_sincos(x, &s, &c);
    return s;
}
MQI_HD inline double
pm_cos(double x) {
    double s, c;
    pm_sincos(x, &s, &c);
    return c;
}

// asin on |y| <= 0.5: sum_n c_n y^(2n+1), c_n = (2n)! / (4^n (n!)^2 (2n+1)).
MQI_HD inline double
pm_asin_small(double y) {
    const double z = y * y;
    // a_n = ((2n-1)!!/(2n)!!) / (2n+1), n = 27 .. 0 (exact doubles of that recurrence)
    double p = 0x1.018f963c229bfp-9;
    p        = p * z + 0x1.1052bc5fa960ap-9;
    p        = p * z + 0x1.208d3570ae5a6p-9;
    p        = p * z + 0x1.3275586c5f2f0p-9;
    p        = p * z + 0x1.464c0950f7d47p-9;
    p        = p * z + 0x1.5c5f56efaaaabp-9;
    p        = p * z + 0x1.750de64d7d05fp-9;
    p        = p * z + 0x1.90cb77f60c7cep-9;
    p        = p * z + 0x1.b026f57b13b14p-9;
    p        = p * z + 0x1.d3d2a8e0dd67dp-9;
    p        = p * z + 0x1.fcaf8fb6db6dbp-9;
    p        = p * z + 0x1.15ee9d45d1746p-8;
    p        = p * z + 0x1.31683bdef7bdfp-8;
    p        = p * z + 0x1.51ba308d3dcb1p-8;
    p        = p * z + 0x1.782dda12f684cp-8;
    p        = p * z + 0x1.a6863d70a3d71p-8;
    p        = p * z + 0x1.df3bd37a6f4dfp-8;
    p        = p * z + 0x1.12ef3cf3cf3cfp-7;
    p        = p * z + 0x1.3fde50d79435ep-7;
    p        = p * z + 0x1.7a87878787878p-7;
    p        = p * z + 0x1.c99999999999ap-7;
    p        = p * z + 0x1.1c4ec4ec4ec4fp-6;
    p        = p * z + 0x1.6e8ba2e8ba2e9p-6;
    p        = p * z + 0x1.f1c71c71c71c7p-6;
    p        = p * z + 0x1.6db6db6db6db7p-5;
    p        = p * z + 0x1.3333333333333p-4;
    p        = p * z + 0x1.5555555555555p-3;
    p        = p * z + 0x1.0000000000000p+0;
    return y * p;
}

MQI_HD inline double
pm_acos(double x) {
    if (x != x || x > 1.0 || x < -1.0) return __builtin_nan("");
    if (x <= 0.5 && x >= -0.5) return kPio2 - pm_asin_small(x);
    if (x > 0.5) return 2.0 * pm_asin_small(pm_sqrt((1.0 - x) * 0.5));
    return kPi - 2.0 * pm_asin_small(pm_sqrt((1.0 + x) * 0.5));
}

// Float entry points: evaluate in double, round once.
MQI_HD inline float
pm_logf(float x) {
    return (float) pm_log((double) x);
}
MQI_HD inline float
pm_expf(float x) {
    return (float) pm_exp((double) x);
}
MQI_HD inline float
pm_powf(float x, float y) {
    return (float) pm_pow((double) x, (double) y);
}
MQI_HD inline float
pm_sinf(float x) {
    return (float) pm_sin((double) x);
}
MQI_HD inline float
pm_cosf(float x) {
    return (float) pm_cos((double) x);
}
MQI_HD inline float
pm_acosf(float x) {
    return (float) pm_acos((double) x);
}
MQI_HD inline float
pm_sqrtf(float x) {
    return (float) pm_sqrt((double) x);
}

}   // namespace mqi_pm

#endif
