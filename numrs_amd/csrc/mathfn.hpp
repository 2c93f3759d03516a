// Elementary functions in FP64 for the elementwise map: exp, exp2, expm1, log, log2, log10, log1p,
// tanh, sinh, cosh, asinh, acosh, atanh, atan, asin, acos, sin, cos, tan (the reference's
// ufuncs.rs:398-510, :542-590 and :619-700, which
// call Rust's f64 methods below 64
// elements and a SIMD library above). Each is written here from a textbook method, named per
// function; every constant is a mathematical constant (ln 2 and its relatives, split so that
// products with small integers are exact, the Taylor coefficients 1/n! and 2/(2j+1)), derived
// with 80-digit decimal arithmetic. Binary32 inputs are evaluated in FP64 and rounded once.
//
// Accuracy (measured against correctly rounded results, tests/test_gpu_ufuncs.py): within 1 ulp
// for exp, exp2, expm1, log, log2, log10 and log1p, within 2 ulp for tanh, sinh, cosh, asinh and
// atanh, 3 ulp for acosh (near 1 the result is ~sqrt(2t): the rounding of 2t + t^2, the square
// root, t + s and log1p add up); atan 1.5 ulp, asin and acos 2; sin and cos 1, tan 2.5 (a
// quotient of the two), for every finite argument; exact special values
// (NaN, +-inf, +-0, the overflow and underflow edges).
#pragma once

namespace nrk {
namespace fm {

// ln 2 = LN2_HI + LN2_LO, LN2_HI with 32 significant bits: k * LN2_HI is exact for |k| < 2^21.
constexpr double LN2 = 0x1.62e42fefa39efp-1;
constexpr double LN2_TAIL = 0x1.abc9e3b39803fp-56;  // ln 2 - LN2
constexpr double LN2_HI = 0x1.62e42fee00000p-1;
constexpr double LN2_LO = 0x1.a39ef35793c76p-33;
constexpr double INV_LN2 = 0x1.71547652b82fep+0;
constexpr double INV_LN2_HI = 0x1.7154765200000p+0;
constexpr double INV_LN2_LO = 0x1.705fc2eefa200p-33;
constexpr double INV_LN10_HI = 0x1.bcb7b15200000p-2;
constexpr double INV_LN10_LO = 0x1.b9438ca9aadd5p-36;
constexpr double LOG10_2_HI = 0x1.3441350800000p-2;  // log10(2), 32 bits
constexpr double LOG10_2_LO = 0x1.f79fef311f12bp-34;
constexpr double SQRT_HALF = 0x1.6a09e667f3bcdp-1;

__device__ __forceinline__ double dinf() { return __builtin_inf(); }

// e^r - 1 for |r| <= 0.35: Taylor series to degree 13 (the first omitted term is below 2^-58
// relative), as r + r^2 (1/2! + r/3! + ... + r^11/13!) so that the leading term is exact.
__device__ __forceinline__ double expm1_kernel(double r) {
    double q = 0x1.6124613a86d09p-33;  // 1/13!
    q = fma(q, r, 0x1.1eed8eff8d898p-29);  // 1/12!
    q = fma(q, r, 0x1.ae64567f544e4p-26);  // 1/11!
    q = fma(q, r, 0x1.27e4fb7789f5cp-22);  // 1/10!
    q = fma(q, r, 0x1.71de3a556c734p-19);  // 1/9!
    q = fma(q, r, 0x1.a01a01a01a01ap-16);  // 1/8!
    q = fma(q, r, 0x1.a01a01a01a01ap-13);  // 1/7!
    q = fma(q, r, 0x1.6c16c16c16c17p-10);  // 1/6!
    q = fma(q, r, 0x1.1111111111111p-7);   // 1/5!
    q = fma(q, r, 0x1.5555555555555p-5);   // 1/4!
    q = fma(q, r, 0x1.5555555555555p-3);   // 1/3!
    q = fma(q, r, 0.5);                    // 1/2!
    return fma(r * r, q, r);
}

// e^x - 1 for |x| <= 0.7: Taylor series to degree 17 (first omitted term below 2^-61 relative),
// as x + x^2 (1/2! + x/3! + ... + x^15/17!).
__device__ __forceinline__ double expm1_series(double x) {
    double q = 0x1.952c77030ad4ap-49;      // 1/17!
    q = fma(q, x, 0x1.ae7f3e733b81fp-45);  // 1/16!
    q = fma(q, x, 0x1.ae7f3e733b81fp-41);  // 1/15!
    q = fma(q, x, 0x1.93974a8c07c9dp-37);  // 1/14!
    q = fma(q, x, 0x1.6124613a86d09p-33);  // 1/13!
    q = fma(q, x, 0x1.1eed8eff8d898p-29);  // 1/12!
    q = fma(q, x, 0x1.ae64567f544e4p-26);  // 1/11!
    q = fma(q, x, 0x1.27e4fb7789f5cp-22);  // 1/10!
    q = fma(q, x, 0x1.71de3a556c734p-19);  // 1/9!
    q = fma(q, x, 0x1.a01a01a01a01ap-16);  // 1/8!
    q = fma(q, x, 0x1.a01a01a01a01ap-13);  // 1/7!
    q = fma(q, x, 0x1.6c16c16c16c17p-10);  // 1/6!
    q = fma(q, x, 0x1.1111111111111p-7);   // 1/5!
    q = fma(q, x, 0x1.5555555555555p-5);   // 1/4!
    q = fma(q, x, 0x1.5555555555555p-3);   // 1/3!
    q = fma(q, x, 0.5);                    // 1/2!
    return fma(x * x, q, x);
}

// y 2^k for k in [-1100, 1100] without an intermediate overflow or a premature underflow.
__device__ __forceinline__ double scale2(double y, int k) {
    const int h = k / 2;
    return ldexp(ldexp(y, h), k - h);
}

// e^(rh + rl) for |rh| <= 0.35, |rl| << ulp(rh): 1 + s + rl (1 + s) with s = e^rh - 1, the
// rounding error of 1 + s and the rl term added before the one final rounding.
__device__ __forceinline__ double exp_dd(double rh, double rl) {
    const double s = expm1_kernel(rh);
    const double h = 1.0 + s;
    const double e = (1.0 - h) + s;  // 1 + s = h + e exactly (|s| < 1)
    return h + fma(rl, h, e);
}

// e^x: Cody & Waite's reduction x = k ln2 + r (|r| <= ln2 / 2) through the split ln 2, with the
// rounding error of r carried as rl, then the scaling by 2^k.
__device__ inline double exp(double x) {
    if (x != x) return x;
    if (x > 710.0) return dinf();
    if (x < -746.0) return 0.0;
    const double k = rint(x * INV_LN2);
    const double r1 = fma(-k, LN2_HI, x);  // exact
    const double r = fma(-k, LN2_LO, r1);
    const double rl = fma(-k, LN2_LO, r1 - r);
    return scale2(exp_dd(r, rl), (int)k);
}

// 2^x = 2^k e^(f ln2), k = rint(x), f = x - k exact.
__device__ inline double exp2(double x) {
    if (x != x) return x;
    if (x > 1025.0) return dinf();
    if (x < -1076.0) return 0.0;
    const double k = rint(x);
    const double f = x - k;
    const double rh = f * LN2;
    const double rl = fma(f, LN2, -rh) + f * LN2_TAIL;
    return scale2(exp_dd(rh, rl), (int)k);
}

// e^x - 1: the series directly for |x| <= 0.7; otherwise 2^k (1 + s) - 1 with s = e^r - 1,
// formed as one FMA fma(2^k, s, 2^k - 1) where 2^k - 1 is exact (-53 <= k <= 53), else scaled
// and then 1 subtracted. (Below 0.7, k = 1 would double the error of s in 2 s + 1.)
__device__ inline double expm1(double x) {
    if (x != x || x == 0.0) return x;  // NaN, and +-0 keeping its sign
    if (x > 710.0) return dinf();
    if (x < -38.0) return -1.0;  // e^x < 2^-54: -1 + e^x rounds to -1
    if (__builtin_fabs(x) <= 0.7) return expm1_series(x);
    const double k = rint(x * INV_LN2);
    const double r1 = fma(-k, LN2_HI, x);  // exact
    const double r = fma(-k, LN2_LO, r1);
    const int ki = (int)k;
    if (ki >= -53 && ki <= 53) {
        const double p = ldexp(1.0, ki);
        return fma(p, expm1_kernel(r), p - 1.0);
    }
    // 2^k (h + e) - 1 with h + e = e^r exactly enough: 2^k h - 1 carried exactly, one rounding
    const double s = expm1_kernel(r);
    const double h = 1.0 + s;
    const double e = fma(fma(-k, LN2_LO, r1 - r), h, (1.0 - h) + s);
    if (ki > 1000) return scale2(h + e, ki);  // the 1 is far below an ulp
    const double ph = ldexp(h, ki), pe = ldexp(e, ki);  // exact
    const double sm = ph - 1.0;
    const double bb = sm - ph;
    return sm + (((ph - (sm - bb)) + (-1.0 - bb)) + pe);  // TwoSum(ph, -1) + pe
}

// log(m) for m in [sqrt(1/2), sqrt(2)): with f = m - 1 (exact) and s = f / (2 + f),
// log(m) = 2 atanh(s) = 2s + 2s^3/3 + 2s^5/5 + ..., rearranged as f - (f^2/2 - s (f^2/2 + R))
// with R = sum_{j>=1} 2 s^(2j) / (2j + 1) (|s| < 0.1716: 13 terms reach 2^-63 relative).
// Returned as hi + lo: the last subtraction is carried exactly (|f| > |f^2/2 - ...|).
__device__ __forceinline__ void log_mant(double f, double &hi, double &lo) {
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.2f684bda12f68p-4;  // 2/27
    R = fma(R, z, 0x1.47ae147ae147bp-4);  // 2/25
    R = fma(R, z, 0x1.642c8590b2164p-4);  // 2/23
    R = fma(R, z, 0x1.8618618618618p-4);  // 2/21
    R = fma(R, z, 0x1.af286bca1af28p-4);  // 2/19
    R = fma(R, z, 0x1.e1e1e1e1e1e1ep-4);  // 2/17
    R = fma(R, z, 0x1.1111111111111p-3);  // 2/15
    R = fma(R, z, 0x1.3b13b13b13b14p-3);  // 2/13
    R = fma(R, z, 0x1.745d1745d1746p-3);  // 2/11
    R = fma(R, z, 0x1.c71c71c71c71cp-3);  // 2/9
    R = fma(R, z, 0x1.2492492492492p-2);  // 2/7
    R = fma(R, z, 0x1.999999999999ap-2);  // 2/5
    R = fma(R, z, 0x1.5555555555555p-1);  // 2/3
    R *= z;
    const double hfsq = 0.5 * f * f;
    const double X = hfsq - s * (hfsq + R);
    hi = f - X;
    lo = (f - hi) - X;  // Fast2Sum: exact
}

// x = 2^e m with m in [sqrt(1/2), sqrt(2)); returns false (and the IEEE result in *special)
// for x <= 0, +inf and NaN. Subnormals are scaled by 2^54 first.
__device__ __forceinline__ bool log_split(double x, int &e, double &f, double &special) {
    if (!(x > 0.0) || x == dinf()) {
        special = x == 0.0 ? -dinf() : (x == dinf() ? x : __builtin_nan(""));
        if (x != x) special = x;
        return false;
    }
    int adj = 0;
    if (x < 0x1p-1022) {
        x *= 0x1p54;
        adj = 54;
    }
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < SQRT_HALF) {
        m *= 2.0;
        e -= 1;
    }
    e -= adj;
    f = m - 1.0;  // exact (Sterbenz)
    return true;
}

// a + b + c with a + b carried exactly (TwoSum), the small terms added last.
__device__ __forceinline__ double sum3(double a, double b, double c) {
    const double s = a + b;
    const double bb = s - a;
    const double t = (a - (s - bb)) + (b - bb);
    return s + (t + c);
}

// log x = e ln2 + log(m) (+ c, a small correction folded in before the last rounding).
__device__ __forceinline__ double log_plus(double x, double c, bool &special_case) {
    int e;
    double f, sp;
    special_case = !log_split(x, e, f, sp);
    if (special_case) return sp;
    double hi, lo;
    log_mant(f, hi, lo);
    const double de = (double)e;
    return sum3(de * LN2_HI, hi, lo + fma(de, LN2_LO, c));  // e LN2_HI is exact
}
__device__ inline double log(double x) {
    bool sp;
    return log_plus(x, 0.0, sp);
}

// log2 x = e + log(m) / ln2 and log10 x = e log10(2) + log(m) / ln10, the products with the
// constants carried in two parts.
__device__ inline double log2(double x) {
    int e;
    double f, sp;
    if (!log_split(x, e, f, sp)) return sp;
    double hi, lo;
    log_mant(f, hi, lo);
    const double p = hi * INV_LN2_HI;
    const double plo = fma(hi, INV_LN2_HI, -p) + fma(hi, INV_LN2_LO, lo * INV_LN2);
    return sum3((double)e, p, plo);
}
__device__ inline double log10(double x) {
    int e;
    double f, sp;
    if (!log_split(x, e, f, sp)) return sp;
    double hi, lo;
    log_mant(f, hi, lo);
    const double de = (double)e;
    const double b = hi * INV_LN10_HI;
    const double blo = fma(hi, INV_LN10_HI, -b) + fma(hi, INV_LN10_LO, fma(lo, INV_LN10_HI, de * LOG10_2_LO));
    return sum3(de * LOG10_2_HI, b, blo);  // e LOG10_2_HI is exact
}

// log(1 + x): u = 1 + x rounded, then log(u) plus the first-order correction for that rounding,
// (x - (u - 1)) / u (Goldberg, "What every computer scientist should know about floating-point
// arithmetic", 1991).
__device__ inline double log1p(double x) {
    if (x != x) return x;
    if (x == dinf()) return x;
    if (x < -1.0) return __builtin_nan("");
    if (x == -1.0) return -dinf();
    if (__builtin_fabs(x) < 0x1p-54) return x;
    const double u = 1.0 + x;
    if (u == 1.0) return x;
    const double c = u < 0x1p53 ? (x - (u - 1.0)) / u : 0.0;
    bool sp;
    return log_plus(u, c, sp);
}

// tanh x = t / (t + 2) with t = e^(2|x|) - 1, sign restored; +-1 beyond |x| = 22, x below 2^-28.
// The divisor t + 2 is carried as d + dl exactly and the quotient corrected once by its residual.
__device__ inline double tanh(double x) {
    if (x != x) return x;
    const double ax = __builtin_fabs(x);
    if (ax > 22.0) return __builtin_copysign(1.0, x);
    if (ax < 0x1p-28) return x;
    const double t = expm1(2.0 * ax);
    const double d = t + 2.0;
    const double dl = t < 2.0 ? (2.0 - d) + t : (t - d) + 2.0;  // Fast2Sum, larger operand first
    const double q = t / d;
    const double res = fma(-q, d, t) - q * dl;  // t - q (d + dl)
    return __builtin_copysign(q + res / d, x);
}

// Hyperbolic functions from the exponential and logarithm above (identities of the definitions):
//   sinh x = (t + t / (t + 1)) / 2, t = e^|x| - 1 (no cancellation), e^|x| / 2 beyond |x| = 22;
//   cosh x = 1 + t^2 / (2 (1 + t)) below |x| = ln2 / 2, else (e + 1/e) / 2;
//   atanh x = log1p(2x / (1 - x)) / 2;  asinh x = log1p(|x| + x^2 / (1 + sqrt(1 + x^2)));
//   acosh x = log1p(t + sqrt(2t + t^2)), t = x - 1 (x <= 2), log(2x - 1/(x + sqrt(x^2 - 1)))
//   below 2^28 and log(x) + ln2 above. Sign restored for the odd ones.
__device__ __forceinline__ double half_exp_big(double ax) {  // e^ax / 2 for ax >= 22
    if (ax <= 709.78) return 0.5 * exp(ax);
    if (ax <= 710.5) {
        const double w = exp(0.5 * ax);
        return (0.5 * w) * w;
    }
    return dinf();
}
__device__ inline double sinh(double x) {
    if (x != x) return x;
    const double ax = __builtin_fabs(x);
    if (ax < 0x1p-28) return x;
    double h;
    if (ax < 22.0) {
        const double t = expm1(ax);
        h = 0.5 * (t + t / (t + 1.0));
    } else {
        h = half_exp_big(ax);
    }
    return __builtin_copysign(h, x);
}
__device__ inline double cosh(double x) {
    if (x != x) return x;
    const double ax = __builtin_fabs(x);
    if (ax < 0.3465) {
        const double t = expm1(ax);
        return 1.0 + (t * t) / (2.0 * (1.0 + t));
    }
    if (ax < 22.0) {
        const double e = exp(ax);
        return 0.5 * e + 0.5 / e;
    }
    return half_exp_big(ax);
}
__device__ inline double atanh(double x) {
    if (x != x) return x;
    const double ax = __builtin_fabs(x);
    if (ax > 1.0) return __builtin_nan("");
    if (ax == 1.0) return __builtin_copysign(dinf(), x);
    if (ax < 0x1p-28) return x;
    double r;
    if (ax < 0.5) {
        const double t = ax + ax;
        r = 0.5 * log1p(t + t * ax / (1.0 - ax));
    } else {
        r = 0.5 * log1p((ax + ax) / (1.0 - ax));
    }
    return __builtin_copysign(r, x);
}
__device__ inline double asinh(double x) {
    if (x != x || x == dinf() || x == -dinf()) return x;
    const double ax = __builtin_fabs(x);
    if (ax < 0x1p-28) return x;
    double r;
    if (ax > 0x1p28) {
        r = log(ax) + LN2;
    } else if (ax > 2.0) {
        r = log(2.0 * ax + 1.0 / (__dsqrt_rn(x * x + 1.0) + ax));
    } else {
        const double t = x * x;
        r = log1p(ax + t / (1.0 + __dsqrt_rn(1.0 + t)));
    }
    return __builtin_copysign(r, x);
}
__device__ inline double acosh(double x) {
    if (x != x) return x;
    if (x < 1.0) return __builtin_nan("");
    if (x == 1.0) return 0.0;
    if (x == dinf()) return x;
    if (x > 0x1p28) return log(x) + LN2;
    if (x > 2.0) return log(2.0 * x - 1.0 / (x + __dsqrt_rn(x * x - 1.0)));
    const double t = x - 1.0;
    return log1p(t + __dsqrt_rn(fma(t, t, 2.0 * t)));  // 2t + t^2 rounded once
}

// ---- inverse trigonometric functions (ufuncs.rs:542-590) ----
// atan t for t in [0, 1]: with c = j/8 the eighth at or below t, atan t = atan(c) + atan(u),
// u = (t - c) / (1 + t c) in [0, 1/8] (no cancellation between the two terms); atan(u) by its odd
// Taylor series to u^21 (the first omitted term is below 2^-60 relative); atan(j/8) in two parts. tl is a correction to t
// (the rounding error of 1/x when t came from it); the division's residual is carried too.
constexpr double PIO2_HI = 0x1.921fb54442d18p+0, PIO2_LO = 0x1.1a62633145c07p-54;
__device__ __forceinline__ double atan_tab_hi(int j) {
    switch (j) {
        case 1: return 0x1.fd5ba9aac2f6ep-4;
        case 2: return 0x1.f5b75f92c80ddp-3;
        case 3: return 0x1.6f61941e4def1p-2;
        case 4: return 0x1.dac670561bb4fp-2;
        case 5: return 0x1.1e00babdefeb4p-1;
        case 6: return 0x1.4978fa3269ee1p-1;
        case 7: return 0x1.700a7c5784634p-1;
        case 8: return 0x1.921fb54442d18p-1;
        default: return 0.0;
    }
}
__device__ __forceinline__ double atan_tab_lo(int j) {
    switch (j) {
        case 1: return -0x1.cd37686760c17p-59;
        case 2: return 0x1.8ab6e3cf7afbdp-57;
        case 3: return -0x1.c63aae6f6e918p-56;
        case 4: return 0x1.a2b7f222f65e2p-56;
        case 5: return -0x1.928df287a668fp-58;
        case 6: return 0x1.2419a87f2a458p-56;
        case 7: return -0x1.8c34d25aadef6p-56;
        case 8: return 0x1.1a62633145c07p-55;
        default: return 0.0;
    }
}
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void atan_core(double t, double tl, double &hi, double &lo) {
    const int j = (int)(t * 8.0);
    double u = t, ul = tl;
    if (j != 0) {
        const double c = 0.125 * (double)j;
        const double num = t - c;  // exact (Sterbenz)
        const double den = fma(t, c, 1.0);
        u = num / den;
        ul = fma(-u, den, num) / den + tl * fma(c, c, 1.0) / (den * den);
    }
    const double z = u * u;
    double q = 0x1.8618618618618p-5;      // 1/21
    q = fma(q, z, -0x1.af286bca1af28p-5); // -1/19
    q = fma(q, z, 0x1.e1e1e1e1e1e1ep-5);  // 1/17
    q = fma(q, z, -0x1.1111111111111p-4); // -1/15
    q = fma(q, z, 0x1.3b13b13b13b14p-4);  // 1/13
    q = fma(q, z, -0x1.745d1745d1746p-4); // -1/11
    q = fma(q, z, 0x1.c71c71c71c71cp-4);  // 1/9
    q = fma(q, z, -0x1.2492492492492p-3); // -1/7
    q = fma(q, z, 0x1.999999999999ap-3);  // 1/5
    q = fma(q, z, -0x1.5555555555555p-2); // -1/3
    const double p = fma(u * z, q, u);
    double e;
    two_sum(atan_tab_hi(j), p, hi, e);
    lo = e + (atan_tab_lo(j) + ul / (1.0 + z));
}
// atan t for t >= 0 (+inf included) as hi + lo; above 1 through pi/2 - atan(1/t).
__device__ __forceinline__ void atan_dd(double t, double &hi, double &lo) {
    if (t == dinf()) {
        hi = PIO2_HI;
        lo = PIO2_LO;
        return;
    }
    if (t > 1.0) {
        const double r = 1.0 / t;
        const double rl = fma(-r, t, 1.0) / t;
        double h, l, e;
        atan_core(r, rl, h, l);
        two_sum(PIO2_HI, -h, hi, e);
        lo = e + (PIO2_LO - l);
        return;
    }
    atan_core(t, 0.0, hi, lo);
}
__device__ inline double atan(double x) {
    if (x != x || x == 0.0) return x;
    double h, l;
    atan_dd(__builtin_fabs(x), h, l);
    return __builtin_copysign(h + l, x);
}
// asin x = atan(x / sqrt(1 - x^2)), acos x = atan(sqrt(1 - x^2) / x) (pi minus it for x < 0):
// 1 - x^2 is one FMA (rounded once, no cancellation), so the argument of atan is within ~1.25 ulp
// and atan passes a relative error on at most unchanged.
__device__ inline double asin(double x) {
    if (x != x || x == 0.0) return x;
    const double ax = __builtin_fabs(x);
    if (ax > 1.0) return __builtin_nan("");
    const double s = __dsqrt_rn(fma(-ax, ax, 1.0));
    double h, l;
    atan_dd(s > 0.0 ? ax / s : dinf(), h, l);
    return __builtin_copysign(h + l, x);
}
__device__ inline double acos(double x) {
    if (x != x) return x;
    const double ax = __builtin_fabs(x);
    if (ax > 1.0) return __builtin_nan("");
    const double s = __dsqrt_rn(fma(-ax, ax, 1.0));
    double h, l;
    atan_dd(ax > 0.0 ? s / ax : dinf(), h, l);
    if (x >= 0.0) return h + l;
    double a, e;
    two_sum(2.0 * PIO2_HI, -h, a, e);
    return a + (e + (2.0 * PIO2_LO - l));
}

// ---- sin, cos, tan (ufuncs.rs:477-510) ----
// Reduction x = k pi/2 + r, |r| <= pi/4, r as rh + rl:
//   |x| < 2^20: Cody & Waite with pi/2 in three parts (PIO2_HI exact against k, the rest carried);
//   larger: Payne & Hanek, "Radian reduction for trigonometric functions" (1983): with x = m 2^E
//   (m the 53-bit significand), only the bits of 2/pi from 2^-(E-1) on matter mod 4; m times a
//   192-bit window of them (three 64-bit words) gives k mod 4 and a 128-bit fraction f, and
//   r = f pi/2. The window's first dropped bit is worth m 2^-190 < 2^-137.
// The bits of 2/pi, 1536 of them (bit 1 = 2^-1 is the top bit of the first word), from 500-digit
// arithmetic.
__device__ __forceinline__ unsigned long long two_over_pi_word(int w) {
    static constexpr unsigned long long T[24] = {
        0xa2f9836e4e441529ull, 0xfc2757d1f534ddc0ull, 0xdb6295993c439041ull,
        0xfe5163abdebbc561ull, 0xb7246e3a424dd2e0ull, 0x06492eea09d1921cull,
        0xfe1deb1cb129a73eull, 0xe88235f52ebb4484ull, 0xe99c7026b45f7e41ull,
        0x3991d639835339f4ull, 0x9c845f8bbdf9283bull, 0x1ff897ffde05980full,
        0xef2f118b5a0a6d1full, 0x6d367ecf27cb09b7ull, 0x4f463f669e5fea2dull,
        0x7527bac7ebe5f17bull, 0x3d0739f78a5292eaull, 0x6bfb5fb11f8d5d08ull,
        0x56033046fc7b6babull, 0xf0cfbc209af4361dull, 0xa9e391615ee61b08ull,
        0x6599855f14a06840ull, 0x8dffd8804d732731ull, 0x06061556ca73a8c9ull};
    return T[w];
}
// 64 bits of 2/pi starting at bit b >= 1.
__device__ __forceinline__ unsigned long long two_over_pi_bits(int b) {
    const int o = b - 1, w = o >> 6, sft = o & 63;
    const unsigned long long hi = two_over_pi_word(w);
    return sft ? (hi << sft) | (two_over_pi_word(w + 1) >> (64 - sft)) : hi;
}
constexpr double PIO2_3 = -0x1.f1976b7ed8fbcp-110;  // pi/2 = PIO2_HI + PIO2_LO + PIO2_3
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;

// ax >= 2^20 (finite): quadrant k mod 4 and r = rh + rl.
__device__ inline int reduce_payne_hanek(double ax, double &rh, double &rl) {
    int e;
    const double mf = frexp(ax, &e);
    const unsigned long long m = (unsigned long long)ldexp(mf, 53);  // ax = m 2^E
    const int E = e - 53;
    const int i0 = E - 1 > 1 ? E - 1 : 1;
    const int p = i0 + 191 - E;  // the binary point of m * window, in [190, 224]
    const unsigned long long w0 = two_over_pi_bits(i0), w1 = two_over_pi_bits(i0 + 64), w2 = two_over_pi_bits(i0 + 128);
    // 256-bit product m * (w0:w1:w2), little-endian words R[0..3]
    unsigned long long R[4];
    const unsigned long long lo2 = m * w2, hi2 = __umul64hi(m, w2);
    const unsigned long long lo1 = m * w1, hi1 = __umul64hi(m, w1);
    const unsigned long long lo0 = m * w0, hi0 = __umul64hi(m, w0);
    R[0] = lo2;
    R[1] = hi2 + lo1;
    const unsigned long long c1 = R[1] < hi2 ? 1ull : 0ull;
    const unsigned long long t2 = hi1 + lo0;
    const unsigned long long c2a = t2 < hi1 ? 1ull : 0ull;
    R[2] = t2 + c1;
    const unsigned long long c2 = c2a + (R[2] < t2 ? 1ull : 0ull);
    R[3] = hi0 + c2;
    auto chunk = [&](int b) -> unsigned long long {  // 64 bits of R from bit b, 62 <= b <= 160
        const int w = b >> 6, sft = b & 63;
        return sft ? (R[w] >> sft) | (R[w + 1] << (64 - sft)) : R[w];
    };
    const int pw = p >> 6, ps = p & 63;
    unsigned q = (unsigned)(R[pw] >> ps);
    if (ps == 63) q |= (unsigned)(R[pw + 1] << 1);  // pw <= 2 here (p <= 224)
    q &= 3u;
    unsigned long long fh = chunk(p - 64), fl = chunk(p - 128);
    double sign = 1.0;
    if (fh >> 63) {  // f >= 1/2: take f - 1 and the next quadrant
        q = (q + 1u) & 3u;
        fl = ~fl + 1ull;
        fh = ~fh + (fl == 0ull ? 1ull : 0ull);
        sign = -1.0;
    }
    // f = fh 2^-64 + fl 2^-128 as a double-double, then times pi/2
    const double fd_hi = __ull2double_rn(fh) * 0x1p-64;
    const long long rem = (long long)(fh - (unsigned long long)(fd_hi * 0x1p64));
    const double fd_lo = ((double)rem + __ull2double_rn(fl) * 0x1p-64) * 0x1p-64;
    const double h = fd_hi * PIO2_HI;
    const double l = fma(fd_hi, PIO2_HI, -h) + fma(fd_hi, PIO2_LO, fd_lo * PIO2_HI);
    rh = sign * h;
    rl = sign * l;
    return (int)q;
}
__device__ inline int reduce_pio2(double ax, double &rh, double &rl) {
    if (ax <= 0.785398163397448) {
        rh = ax;
        rl = 0.0;
        return 0;
    }
    if (ax < 0x1p20) {
        const double k = rint(ax * TWO_OVER_PI);
        const double r1 = fma(-k, PIO2_HI, ax);  // exact
        const double th = k * PIO2_LO, tl = fma(k, PIO2_LO, -th);
        double h, e;
        two_sum(r1, -th, h, e);
        const double l = (e - tl) - k * PIO2_3;
        rh = h + l;
        rl = l - (rh - h);
        return (int)k & 3;
    }
    return reduce_payne_hanek(ax, rh, rl);
}
// sin and cos on |r| <= pi/4 by Taylor series (to r^19 and r^18: first omitted term below
// 2^-60 relative), the tail rl folded in to first order.
__device__ __forceinline__ double sin_kernel(double rh, double rl) {
    const double z = rh * rh;
    double q = -0x1.2f49b46814157p-57;
    q = fma(q, z, 0x1.952c77030ad4ap-49);
    q = fma(q, z, -0x1.ae7f3e733b81fp-41);
    q = fma(q, z, 0x1.6124613a86d09p-33);
    q = fma(q, z, -0x1.ae64567f544e4p-26);
    q = fma(q, z, 0x1.71de3a556c734p-19);
    q = fma(q, z, -0x1.a01a01a01a01ap-13);
    q = fma(q, z, 0x1.1111111111111p-7);
    q = fma(q, z, -0x1.5555555555555p-3);  // -1/3!
    return rh + fma(rh * z, q, rl * (1.0 - 0.5 * z));
}
__device__ __forceinline__ double cos_kernel(double rh, double rl) {
    const double z = rh * rh, zl = fma(rh, rh, -z);
    double q = -0x1.6827863b97d97p-53;
    q = fma(q, z, 0x1.ae7f3e733b81fp-45);
    q = fma(q, z, -0x1.93974a8c07c9dp-37);
    q = fma(q, z, 0x1.1eed8eff8d898p-29);
    q = fma(q, z, -0x1.27e4fb7789f5cp-22);
    q = fma(q, z, 0x1.a01a01a01a01ap-16);
    q = fma(q, z, -0x1.6c16c16c16c17p-10);
    q = fma(q, z, 0x1.5555555555555p-5);  // 1/4!
    const double hz = 0.5 * z, w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * z * q - (0.5 * zl + rh * rl)));
}
__device__ inline double sin(double x) {
    if (x != x || x == dinf() || x == -dinf()) return __builtin_nan("");
    if (x == 0.0) return x;
    double rh, rl;
    const int q = reduce_pio2(__builtin_fabs(x), rh, rl);
    const double v = (q & 1) ? cos_kernel(rh, rl) : sin_kernel(rh, rl);
    return __builtin_copysign(1.0, x) * ((q & 2) ? -v : v);  // sin(-x) = -sin(x)
}
__device__ inline double cos(double x) {
    if (x != x || x == dinf() || x == -dinf()) return __builtin_nan("");
    double rh, rl;
    const int q = reduce_pio2(__builtin_fabs(x), rh, rl);
    const double v = (q & 1) ? sin_kernel(rh, rl) : cos_kernel(rh, rl);
    return (q == 1 || q == 2) ? -v : v;
}
__device__ inline double tan(double x) {
    if (x != x || x == dinf() || x == -dinf()) return __builtin_nan("");
    if (x == 0.0) return x;
    double rh, rl;
    const int q = reduce_pio2(__builtin_fabs(x), rh, rl);
    const double sn = sin_kernel(rh, rl), cs = cos_kernel(rh, rl);
    const double v = (q & 1) ? -cs / sn : sn / cs;
    return __builtin_copysign(1.0, x) * v;
}

// ---- binary: hypot, atan2, logaddexp (ufuncs.rs:587, :841, :916) ----
// hypot: the larger magnitude scaled to [1, 2), the sum of squares in double-double, its square
// root corrected once by the residual; C semantics for infinities (hypot(inf, NaN) = inf).
__device__ inline double hypot(double a, double b) {
    double ax = __builtin_fabs(a), ay = __builtin_fabs(b);
    if (ax == dinf() || ay == dinf()) return dinf();
    if (ax != ax || ay != ay) return __builtin_nan("");
    if (ax < ay) {
        const double t = ax;
        ax = ay;
        ay = t;
    }
    if (ay == 0.0) return ax;
    int e;
    (void)frexp(ax, &e);
    const double x = ldexp(ax, 1 - e), y = ldexp(ay, 1 - e);  // x in [1, 2)
    const double xh = x * x, xl = fma(x, x, -xh);
    const double yh = y * y, yl = fma(y, y, -yh);
    double sh, sl;
    two_sum(xh, yh, sh, sl);
    sl += xl + yl;
    const double r = __dsqrt_rn(sh);
    const double rc = r + (fma(-r, r, sh) + sl) / (2.0 * r);
    return ldexp(rc, e - 1);
}
// atan2(y, x) with the C / Rust special cases; otherwise the angle of |y|, |x| from the smaller
// over the larger magnitude (its division residual carried) and the quadrant in double-double.
__device__ inline double atan2(double y, double x) {
    if (x != x || y != y) return __builtin_nan("");
    const double PI_HI = 2.0 * PIO2_HI, PI_LO = 2.0 * PIO2_LO;
    const bool xneg = __builtin_signbit(x);
    if (y == 0.0) return xneg ? __builtin_copysign(PI_HI + PI_LO, y) : y;  // +-0 or +-pi
    if (x == 0.0) return __builtin_copysign(PIO2_HI + PIO2_LO, y);
    const double ax = __builtin_fabs(x), ay = __builtin_fabs(y);
    double h, l;
    if (ay == dinf()) {
        if (ax == dinf()) {  // +-pi/4 or +-3pi/4
            h = xneg ? 3.0 * (0.5 * PIO2_HI) : 0.5 * PIO2_HI;
            l = xneg ? 3.0 * (0.5 * PIO2_LO) : 0.5 * PIO2_LO;
            return __builtin_copysign(h + l, y);
        }
        return __builtin_copysign(PIO2_HI + PIO2_LO, y);
    }
    if (ax == dinf()) return xneg ? __builtin_copysign(PI_HI + PI_LO, y) : __builtin_copysign(0.0, y);
    if (ay <= ax) {
        const double t = ay / ax;
        atan_core(t, fma(-t, ax, ay) / ax, h, l);
    } else {
        const double t = ax / ay;
        double ch, cl, e;
        atan_core(t, fma(-t, ay, ax) / ay, ch, cl);
        two_sum(PIO2_HI, -ch, h, e);
        l = e + (PIO2_LO - cl);
    }
    if (xneg) {
        double e;
        const double hh = h;
        two_sum(PI_HI, -hh, h, e);
        l = e + (PI_LO - l);
    }
    return __builtin_copysign(h + l, y);
}
// logaddexp as the reference writes it: m + ln(e^-|a - m| + e^-|b - m|), m = max(a, b) with
// Rust's NaN-ignoring max (so logaddexp(inf, inf) is NaN, as there).
__device__ inline double logaddexp(double a, double b) {
    const double m = __builtin_fmax(a, b);
    return m + log(exp(-__builtin_fabs(a - m)) + exp(-__builtin_fabs(b - m)));
}

// ---- activations, special functions and steps, as the reference writes them (ufuncs.rs
// :520-1480). Each is the reference's f64 expression with one IEEE rounding per operation (the
// library is built without FP contraction) over the elementary functions above, so it differs
// from the reference only where those differ from correctly rounded values.

__device__ inline double relu(double x) { return x > 0.0 ? x : 0.0; }  // x.max(0.0); NaN -> 0
__device__ inline double elu(double x, double alpha) { return x > 0.0 ? x : alpha * (exp(x) - 1.0); }
__device__ inline double selu(double x) {
    const double lambda = 1.0507009873554805, alpha = 1.6732632423543772;  // the Rust literals, rounded
    return lambda * (x > 0.0 ? x : alpha * (exp(x) - 1.0));
}
__device__ inline double hardsigmoid(double x) {
    return x <= -3.0 ? 0.0 : (x >= 3.0 ? 1.0 : x / 6.0 + 0.5);
}
__device__ inline double hardswish(double x) {
    return x <= -3.0 ? 0.0 : (x >= 3.0 ? x : x * (x + 3.0) / 6.0);
}
__device__ inline double softplus(double x) {
    return x > 20.0 ? x : (x < -20.0 ? exp(x) : log(1.0 + exp(x)));
}
__device__ inline double mish(double x) { return x * tanh(x > 20.0 ? x : log(1.0 + exp(x))); }
__device__ inline double silu(double x) { return x / (1.0 + exp(-x)); }
__device__ inline double gelu(double x) {
    const double c = 0x1.9884533d43651p-1;  // (2.0 / PI).sqrt() in f64: two roundings
    return 0.5 * x * (1.0 + tanh(c * (x + 0.044715 * (x * x * x))));  // powi(3) = (x x) x
}
__device__ inline double logit(double x) { return log(x / (1.0 - x)); }
__device__ inline double sinc(double x) {
    if (x == 0.0) return 1.0;
    const double px = 0x1.921fb54442d18p+1 * x;  // PI * x, rounded
    return sin(px) / px;
}
__device__ inline double smoothstep_poly(double t) { return t * t * (3.0 - 2.0 * t); }
__device__ inline double smootherstep_poly(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

// erf_scalar (ufuncs.rs:1081): the Chebyshev fit of erfc from Numerical Recipes (fractional error
// below 1.2e-7), evaluated in the reference's operation order.
__device__ inline double erf_ref(double x) {
    const double t = 1.0 / (1.0 + 0.5 * __builtin_fabs(x));
    double p = 0.17087277;
    p = -0.82215223 + t * p;
    p = 1.48851587 + t * p;
    p = -1.13520398 + t * p;
    p = 0.27886807 + t * p;
    p = -0.18628806 + t * p;
    p = 0.09678418 + t * p;
    p = 0.37409196 + t * p;
    p = 1.00002368 + t * p;
    const double tau = t * exp(-x * x - 1.26551223 + t * p);
    return x >= 0.0 ? 1.0 - tau : tau - 1.0;
}

// digamma (ufuncs.rs:1198-1231): NaN for x <= 0; below 6 the recurrence
// psi(x) = psi(x + 1) - 1/x, each x + 1 rounded; from 6 on
// ln x - 1/(2x) - x^-2 (1/12 - x^-2 (1/120 - x^-2 / 252)).
__device__ inline double digamma_asym(double x) {
    const double inv_x = 1.0 / x;
    const double inv_x2 = inv_x * inv_x;
    return log(x) - 0.5 * inv_x - inv_x2 * (1.0 / 12.0 - inv_x2 * (1.0 / 120.0 - inv_x2 / 252.0));
}
__device__ inline double digamma(double x) {
    if (x <= 0.0) return __builtin_nan("");
    int k = 0;
    double xk = x;
    while (xk < 6.0) {  // at most 6 steps for x > 0
        xk = xk + 1.0;
        ++k;
    }
    double r = digamma_asym(xk);
    for (int i = k - 1; i >= 0; --i) {  // unwind: r = psi(x_i + 1) - 1/x_i
        double xi = x;
        for (int j = 0; j < i; ++j) xi = xi + 1.0;
        r = r - 1.0 / xi;
    }
    return r;
}


// cbrt (ufuncs.rs:367, x.cbrt()): |x| = m 2^(3q) with m in [0.5, 4); a binary32 start from the
// hardware exp2 / log2 (about 22 bits), two Newton steps y <- y - (y^3 - m) / (3 y^2), then one
// more with the residual y^3 - m formed in double-double, so the last step's own rounding is the
// only one left. Signed zeros, infinities and NaN return themselves.
__device__ inline double cbrt(double x) {
    if (x == 0.0 || x != x || __builtin_fabs(x) == dinf()) return x;
    int e;
    double m = frexp(__builtin_fabs(x), &e);  // [0.5, 1)
    int q = e / 3, r = e - 3 * q;
    if (r < 0) {
        r += 3;
        --q;
    }
    m = ldexp(m, r);  // [0.5, 4)
    const float mf = (float)m;
    double y = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf(mf) * (1.0f / 3.0f));
    y = y - (y * y * y - m) / (3.0 * y * y);
    y = y - (y * y * y - m) / (3.0 * y * y);
    const double y2h = y * y, y2l = fma(y, y, -y2h);
    const double y3h = y2h * y, y3l = fma(y2h, y, -y3h) + y2l * y;
    const double res = (y3h - m) + y3l;  // y3h - m is exact (Sterbenz: y^3 within 2^-40 of m)
    y = y - res / (3.0 * y2h);
    return __builtin_copysign(ldexp(y, q), x);
}

// ---- pow and the Lanczos gamma (ufuncs.rs:226, 305, 1125-1195) ----

// ln z for z = 2^k m in double-double, relative error near 2^-68 (what pow needs: |y ln x| up to
// 745 must keep its error below 2^-58). m is reduced to [0.6875, 1.4375) and split as
// m = c (1 + r) with c from 96 intervals of width 1/128: LOG_TAB[j] = {1/c, ln c hi, ln c lo},
// 1/c rounded to a double (so it is exactly some c'), ln c' to 106 bits (60-digit decimal
// arithmetic). The two intervals around 1 take c = 1, keeping ln z relative near z = 1.
// r = m / c - 1 is formed exactly as a TwoProd, and log1p(r) = r - r^2/2 + r^3 P(r) with the
// r^2 term in double-double and P to r^11 (|r| < 2^-7.4: truncation below 2^-90).
__device__ __forceinline__ void log_tab(int j, double &invc, double &lh, double &ll) {
    static constexpr double LOG_TAB[96][3] = {
        {0x1.724287f46debcp+0, -0x1.79e26687cfb3dp-2, -0x1.fe977e8bbc0dep-56},
        {0x1.6e1f76b4337c7p+0, -0x1.6e60ee6af1973p-2, 0x1.56a0f7749e5cdp-56},
        {0x1.6a13cd1537290p+0, -0x1.630030b3aac48p-2, -0x1.ee0c6728fffccp-56},
        {0x1.661ec6a5122f9p+0, -0x1.57bf753c8d1fbp-2, 0x1.2908d15f88b63p-57},
        {0x1.623fa77016240p+0, -0x1.4c9e09e172c3dp-2, 0x1.123615b147a5fp-58},
        {0x1.5e75bb8d015e7p+0, -0x1.419b423d5e8c6p-2, -0x1.5b7648704e721p-58},
        {0x1.5ac056b015ac0p+0, -0x1.36b6776be1116p-2, 0x1.324f0e8838590p-58},
        {0x1.571ed3c506b3ap+0, -0x1.2bef07cdc9355p-2, 0x1.22dad7fd86088p-56},
        {0x1.5390948f40febp+0, -0x1.214456d0eb8d5p-2, 0x1.50a2dca28b3edp-58},
        {0x1.5015015015015p+0, -0x1.16b5ccbacfb73p-2, -0x1.56fbd28b40935p-56},
        {0x1.4cab88725af6ep+0, -0x1.0c42d676162e2p-2, 0x1.5a74e18a8bb85p-56},
        {0x1.49539e3b2d067p+0, -0x1.01eae5626c691p-2, -0x1.d9f5bd0b5b348p-57},
        {0x1.460cbc7f5cf9ap+0, -0x1.ef5ade4dcffe5p-3, -0x1.7754d2238f75fp-58},
        {0x1.42d6625d51f87p+0, -0x1.db13db0d48941p-3, 0x1.8af715b0349a4p-57},
        {0x1.3fb013fb013fbp+0, -0x1.c6ffbc6f00f71p-3, 0x1.ae58b2c57a4a5p-57},
        {0x1.3c995a47babe7p+0, -0x1.b31d8575bce3bp-3, 0x1.0d4eace1aa537p-59},
        {0x1.3991c2c187f63p+0, -0x1.9f6c407089663p-3, 0x1.52979a7e86605p-57},
        {0x1.3698df3de0748p+0, -0x1.8beafeb38fe8fp-3, 0x1.54aae92cd0b87p-59},
        {0x1.33ae45b57bcb2p+0, -0x1.7898d85444c74p-3, -0x1.be3dbaf3ec804p-60},
        {0x1.30d190130d190p+0, -0x1.6574ebe8c1339p-3, -0x1.c5961e173bc82p-57},
        {0x1.2e025c04b8097p+0, -0x1.527e5e4a1b58dp-3, 0x1.b8d4b411cadffp-60},
        {0x1.2b404ad012b40p+0, -0x1.3fb45a59928cap-3, 0x1.d87e6a354d057p-57},
        {0x1.288b01288b013p+0, -0x1.2d1610c86813dp-3, -0x1.d997036941a6dp-60},
        {0x1.25e22708092f1p+0, -0x1.1aa2b7e23f729p-3, -0x1.6e44389934420p-57},
        {0x1.23456789abcdfp+0, -0x1.08598b59e3a07p-3, 0x1.fd7009902bf32p-57},
        {0x1.20b470c67c0d9p+0, -0x1.ec739830a1126p-4, -0x1.eea033743f95bp-58},
        {0x1.1e2ef3b3fb874p+0, -0x1.c885801bc4b20p-4, 0x1.5c734aa6598fcp-58},
        {0x1.1bb4a4046ed29p+0, -0x1.a4e7640b1bc38p-4, 0x1.9b5ca203e4259p-58},
        {0x1.19453808ca29cp+0, -0x1.8197e2f40e3f0p-4, 0x1.230690020895fp-59},
        {0x1.16e0689427379p+0, -0x1.5e95a4d9791cdp-4, 0x1.4c78ba3a3baf6p-58},
        {0x1.1485f0e0acd3bp+0, -0x1.3bdf5a7d1ee5ep-4, -0x1.f52eda76b68acp-60},
        {0x1.12358e75d3033p+0, -0x1.1973bd1465561p-4, 0x1.7aac1b3d35680p-58},
        {0x1.0fef010fef011p+0, -0x1.eea31c006b87cp-5, 0x1.7c9f9276f6cd8p-60},
        {0x1.0db20a88f4696p+0, -0x1.aaef2d0fb1108p-5, -0x1.68d4eed0b82aep-59},
        {0x1.0b7e6ec259dc8p+0, -0x1.67c94f2d4bb65p-5, -0x1.0413e6505e5f9p-59},
        {0x1.0953f39010954p+0, -0x1.252f32f8d1840p-5, -0x1.ae021b67a9ba8p-61},
        {0x1.073260a47f7c6p+0, -0x1.c63d2ec14aad7p-6, -0x1.8fe7acbca131dp-63},
        {0x1.05197f7d73404p+0, -0x1.432a925980cbcp-6, 0x1.8cdaf39004193p-60},
        {0x1.03091b51f5e1ap+0, -0x1.82448a388a283p-7, -0x1.04b16137f0970p-62},
        {0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0},
        {0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0},
        {0x1.fa11caa01fa12p-1, 0x1.7dc475f810a69p-7, 0x1.74944bc161072p-61},
        {0x1.f6310aca0dbb5p-1, 0x1.3cea44346a584p-6, -0x1.865ad48159d00p-61},
        {0x1.f25f644230ab5p-1, 0x1.b9fc027af919ap-6, -0x1.90ae69229dc86p-60},
        {0x1.ee9c7f8458e02p-1, 0x1.1b0d98923d97fp-5, -0x1.74d7444dd6241p-59},
        {0x1.eae807aba01ebp-1, 0x1.58a5bafc8e4d3p-5, -0x1.cab8569c56e40p-64},
        {0x1.e741aa59750e4p-1, 0x1.95c830ec8e3f2p-5, 0x1.eb41d00a417e9p-60},
        {0x1.e3a9179dc1a73p-1, 0x1.d276b8adb0b56p-5, 0x1.078f14c95ff53p-59},
        {0x1.e01e01e01e01ep-1, 0x1.075983598e471p-4, 0x1.006d2999e22dcp-58},
        {0x1.dca01dca01dcap-1, 0x1.253f62f0a1417p-4, 0x1.1f6d34e01d981p-61},
        {0x1.d92f2231e7f8ap-1, 0x1.42edcbea646eep-4, -0x1.511583653349bp-58},
        {0x1.d5cac807572b2p-1, 0x1.60658a93750c4p-4, -0x1.f108b1d8436d3p-59},
        {0x1.d272ca3fc5b1ap-1, 0x1.7da766d7b12d0p-4, 0x1.a2240644d7da2p-59},
        {0x1.cf26e5c44bfc6p-1, 0x1.9ab42462033aep-4, -0x1.a099e1c184e8ep-59},
        {0x1.cbe6d9601cbe7p-1, 0x1.b78c82bb0eda0p-4, -0x1.3ef0e61f9b03cp-58},
        {0x1.c8b265afb8a42p-1, 0x1.d4313d66cb35dp-4, 0x1.b90dd951d90fap-58},
        {0x1.c5894d10d4986p-1, 0x1.f0a30c01162a4p-4, 0x1.8be64b8b7759bp-59},
        {0x1.c26b5392ea01cp-1, 0x1.0671512ca596fp-3, -0x1.2f39b81479b67p-58},
        {0x1.bf583ee868d8bp-1, 0x1.14785846742acp-3, 0x1.94409f1d3f83ap-60},
        {0x1.bc4fd65883e7bp-1, 0x1.2266f190a5acdp-3, -0x1.dab840e7f6177p-57},
        {0x1.b951e2b18ff23p-1, 0x1.303d718e47fd5p-3, -0x1.b5ae71f658247p-57},
        {0x1.b65e2e3beee05p-1, 0x1.3dfc2b0ecc62ap-3, 0x1.ba62b8c13f7f4p-57},
        {0x1.b37484ad806cep-1, 0x1.4ba36f39a55e5p-3, -0x1.f767e433c98aap-57},
        {0x1.b094b31d922a4p-1, 0x1.59338d9982085p-3, 0x1.8d16eaaba9419p-57},
        {0x1.adbe87f94905ep-1, 0x1.66acd4272ad51p-3, -0x1.9201c9c3d5165p-59},
        {0x1.aaf1d2f87ebfdp-1, 0x1.740f8f54037a3p-3, 0x1.6d9bf9d57b326p-58},
        {0x1.a82e65130e159p-1, 0x1.815c0a14357e9p-3, 0x1.141b7f8c5fa9ep-58},
        {0x1.a574107688a4ap-1, 0x1.8e928de886d41p-3, 0x1.2589eb96a6240p-59},
        {0x1.a2c2a87c51ca0p-1, 0x1.9bb362e7dfb85p-3, -0x1.51439c1ff83e7p-58},
        {0x1.a01a01a01a01ap-1, 0x1.a8becfc882f19p-3, -0x1.a8c37918c39ebp-58},
        {0x1.9d79f176b682dp-1, 0x1.b5b519e8fb5a6p-3, -0x1.d5d8023e61e5fp-57},
        {0x1.9ae24ea5510dap-1, 0x1.c2968558c18c2p-3, 0x1.6108e3ae024acp-60},
        {0x1.9852f0d8ec0ffp-1, 0x1.cf6354e09c5ddp-3, 0x1.339a07d55b696p-57},
        {0x1.95cbb0be377aep-1, 0x1.dc1bca0abec7bp-3, 0x1.c698a33316dfbp-58},
        {0x1.934c67f9b2ce6p-1, 0x1.e8c0252aa5a60p-3, -0x1.dc074737f9135p-60},
        {0x1.90d4f120190d5p-1, 0x1.f550a564b7b37p-3, -0x1.13a09202fe73dp-57},
        {0x1.8e6527af1373fp-1, 0x1.00e6c45ad501dp-2, -0x1.3b9568ff6feadp-57},
        {0x1.8bfce8062ff3ap-1, 0x1.071b85fcd590dp-2, 0x1.08b83fcbdef40p-57},
        {0x1.899c0f601899cp-1, 0x1.0d46b579ab74bp-2, 0x1.21f640e1e5ec9p-56},
        {0x1.87427bcc092b9p-1, 0x1.136870293a8b0p-2, 0x1.86cc531dba494p-57},
        {0x1.84f00c2780614p-1, 0x1.1980d2dd4236fp-2, -0x1.02c2e4f1b2eb9p-56},
        {0x1.82a4a0182a4a0p-1, 0x1.1f8ff9e48a2f3p-2, -0x1.93fbf3418960dp-57},
        {0x1.8060180601806p-1, 0x1.2596010df763ap-2, -0x1.9eed8ae0ebd3cp-59},
        {0x1.7e225515a4f1dp-1, 0x1.2b9303ab89d25p-2, -0x1.85ad7f614ab51p-58},
        {0x1.7beb3922e017cp-1, 0x1.31871c9544185p-2, -0x1.ea3598981366fp-57},
        {0x1.79baa6bb6398bp-1, 0x1.3772662bfd85cp-2, 0x1.02a7589fba088p-57},
        {0x1.77908119ac60dp-1, 0x1.3d54fa5c1f710p-2, 0x1.53668e578d9cdp-58},
        {0x1.756cac201756dp-1, 0x1.432ef2a04e813p-2, -0x1.83262e2b59206p-57},
        {0x1.734f0c541fe8dp-1, 0x1.49006804009d0p-2, -0x1.bff0d07c5df6dp-59},
        {0x1.713786d9c7c09p-1, 0x1.4ec9732600269p-2, -0x1.1aa87d977dc5ep-56},
        {0x1.6f26016f26017p-1, 0x1.548a2c3add263p-2, -0x1.58ce7bf1846eep-56},
        {0x1.6d1a62681c861p-1, 0x1.5a42ab0f4cfe2p-2, -0x1.c6bcb7dee9a3dp-56},
        {0x1.6b1490aa31a3dp-1, 0x1.5ff3070a793d4p-2, -0x1.063077d7e37b7p-56},
        {0x1.691473a88d0c0p-1, 0x1.659b57303e1f2p-2, 0x1.db0af8efb83c7p-62},
        {0x1.6719f3601671ap-1, 0x1.6b3bb2235943dp-2, 0x1.957a93326784dp-56},
        {0x1.6524f853b4aa3p-1, 0x1.70d42e2789236p-2, 0x1.ee99bf7143954p-56},
    };
    invc = LOG_TAB[j][0];
    lh = LOG_TAB[j][1];
    ll = LOG_TAB[j][2];
}

// x > 0 finite (subnormals included).
__device__ inline void log_dd(double x, double &hi, double &lo) {
    int e;
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < 0.6875) {
        m *= 2.0;
        --e;
    }
    int j = (int)((m - 0.6875) * 128.0);  // m - 0.6875 is exact (Sterbenz)
    j = j < 0 ? 0 : (j > 95 ? 95 : j);
    double invc, lh, ll;
    log_tab(j, invc, lh, ll);
    const double ph = m * invc, pl = fma(m, invc, -ph);
    double rh, rl;
    two_sum(ph - 1.0, pl, rh, rl);  // ph - 1 exact (ph in [0.98, 1.02])
    const double sqh = rh * rh, sql = fma(rh, rh, -sqh);
    double p = -0x1.745d1745d1746p-4;    // -1/11
    p = fma(p, rh, 0x1.999999999999ap-4);  // 1/10
    p = fma(p, rh, -0x1.c71c71c71c71cp-4);  // -1/9
    p = fma(p, rh, 0x1.0000000000000p-3);   // 1/8
    p = fma(p, rh, -0x1.2492492492492p-3);  // -1/7
    p = fma(p, rh, 0x1.5555555555555p-3);   // 1/6
    p = fma(p, rh, -0x1.999999999999ap-3);  // -1/5
    p = fma(p, rh, 0x1.0000000000000p-2);   // 1/4
    p = fma(p, rh, -0x1.5555555555555p-2);  // -1/3
    p = -p * (sqh * rh);                    // r^3/3 - r^4/4 + ...
    const double k = (double)e;
    double s1, e1, s2, e2, s3, e3;
    two_sum(k * LN2_HI, lh, s1, e1);       // k LN2_HI exact
    two_sum(s1, rh, s2, e2);
    two_sum(s2, -0.5 * sqh, s3, e3);
    const double tail = e1 + e2 + e3 + ll + k * LN2_LO + rl - 0.5 * sql - rh * rl + p;
    hi = s3 + tail;
    lo = tail - (hi - s3);
}

// e^(xh + xl), |xl| <= 2^-50 |xh|: the reduction of exp with xl joined to the reduced tail.
__device__ inline double exp_tail(double xh, double xl) {
    if (xh > 710.0) return dinf();
    if (xh < -746.0) return 0.0;
    const double k = rint(xh * INV_LN2);
    const double r1 = fma(-k, LN2_HI, xh);  // exact
    const double r = fma(-k, LN2_LO, r1);
    const double rl = fma(-k, LN2_LO, r1 - r) + xl;
    const double rh2 = r + rl, rl2 = rl - (rh2 - r);
    return scale2(exp_dd(rh2, rl2), (int)k);
}

// x^y with C's (and Rust's powf) special cases: e^(y ln|x|) with ln|x| in double-double and the
// product y ln|x| carried as a TwoProd.
__device__ inline double pow(double x, double y) {
    if (y == 0.0 || x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    const double ax = __builtin_fabs(x), ay = __builtin_fabs(y);
    const bool yint = __builtin_trunc(y) == y;
    const bool yodd = yint && ay < 0x1p53 && __builtin_fmod(ay, 2.0) == 1.0;
    const bool neg = __builtin_signbit(x) && yodd;  // -0 included: pow(-0, 3) = -0
    if (ay == dinf()) {
        if (ax == 1.0) return 1.0;
        return ((ax < 1.0) == (y < 0.0)) ? dinf() : 0.0;
    }
    if (ax == dinf()) {
        const double v = y < 0.0 ? 0.0 : dinf();
        return neg ? -v : v;
    }
    if (ax == 0.0) {
        const double v = y < 0.0 ? dinf() : 0.0;
        return neg ? __builtin_copysign(v, x) : v;
    }
    if (x < 0.0 && !yint) return __builtin_nan("");
    double lh, ll;
    log_dd(ax, lh, ll);
    const double ph = y * lh;
    const double pl = fma(y, lh, -ph) + y * ll;
    const double v = exp_tail(ph, pl);
    return neg ? -v : v;
}

// The reference's Lanczos approximation (g = 7, nine coefficients), in its operation order.
__device__ inline double lanczos_ref(double x) {
    const double c[9] = {0.99999999999980993, 676.5203681218851, -1259.1392167224028, 771.32342877765313,
                         -176.61502916214059, 12.507343278686905, -0.13857109526572012, 9.9843695780195716e-6,
                         1.5056327351493116e-7};
    const double xm = x - 1.0;
    double y = c[0];
#pragma unroll
    for (int i = 1; i <= 8; ++i) y += c[i] / (xm + (double)i);
    const double t = xm + 7.0 + 0.5;
    const double SQRT_2PI = 0x1.40d931ff62705p+1;  // (2.0 * PI).sqrt(): two roundings
    return SQRT_2PI * pow(t, xm + 0.5) * exp(-t) * y;
}
__device__ inline double gamma_ref(double x) {
    if (x <= 0.0 && __builtin_floor(x) == x) return dinf();
    const double PI = 0x1.921fb54442d18p+1;
    if (x < 0.5) return PI / (sin(PI * x) * lanczos_ref(1.0 - x));
    return lanczos_ref(x);
}
__device__ inline double lgamma_ref(double x) { return log(__builtin_fabs(lanczos_ref(x))); }

// ---------------- binary32 inputs ----------------
// exp, tanh and pow of binary32 arguments, evaluated in binary64 only as far as binary32 needs:
// short Taylor series after the usual reductions, with a relative error below 2^-36 (exp), 2^-41
// (tanh) or 2^-30 (pow, |y| <= 2^10), then rounded once to binary32, so within 0.52 ulp of it -- the
// full-precision binary64 routines above cost 2-5x as much, which kept these maps below the HBM
// rate. Special values as in the binary64 routines.
namespace f32 {

// 1 / i!, folded at compile time
template <int I>
struct InvFact {
    static constexpr double v = InvFact<I - 1>::v / (double)I;
};
template <>
struct InvFact<0> {
    static constexpr double v = 1.0;
};
template <int D, int I = D>
__device__ __forceinline__ double exp_horner(double r, double p) {
    if constexpr (I == 0) return p;
    else return exp_horner<D, I - 1>(r, __builtin_fma(p, r, InvFact<I - 1>::v));
}
// e^r for |r| <= ln2/2 + a little: Taylor of degree D in Horner form (relative error below
// |r|^(D+1) / (D+1)! / e^-|r|).
template <int D>
__device__ __forceinline__ double exp_taylor(double r) {
    return exp_horner<D>(r, InvFact<D>::v);
}
// 2^k e^r, k = rint(x log2 e): for |x| <= 1100
template <int D>
__device__ __forceinline__ double exp_reduced(double x) {
    const double k = __builtin_rint(x * 0x1.71547652b82fep0);
    const double r = __builtin_fma(-k, 0x1.62e42fefa39efp-1, x);  // |k ln2 - fl(k ln2)| <= 2^-43
    return __builtin_ldexp(exp_taylor<D>(r), (int)k);
}

// 1 / d for the positive normal d here: the hardware reciprocal refined by one Newton step
// (relative error below 2^-50), where a correctly rounded quotient is not needed
__device__ __forceinline__ double recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    const double e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

// e^x unrounded, degree 9: (0.3466)^10 / 10! / 0.707 < 2^-36 relative (NaN propagates)
__device__ __forceinline__ double exp_d(float xf) {
    return exp_reduced<9>(__builtin_fmin(__builtin_fmax((double)xf, -110.0), 100.0));  // 0 / inf once rounded
}
__device__ inline float exp(float xf) {
    if (xf != xf) return xf + xf;
    return (float)exp_d(xf);
}

// tanh: x for |x| < 2^-12 (the cubic term is below half an ulp); else (1 - t) / (1 + t) with
// t = e^-2|x| to 2^-47: 1 - t cancels for small |x|, but its relative error stays below
// 2^-52 / (2|x|) <= 2^-41 there, well inside the 29 bits binary64 carries beyond binary32; the
// sign restored. No branches: every lane of a wavefront runs the same instructions.
__device__ inline float tanh(float xf) {
    const double x = (double)xf, a = __builtin_fmin(__builtin_fabs(x), 600.0);
    const double t = exp_reduced<11>(-2.0 * a);
    const double v = (1.0 - t) * recip(1.0 + t);
    return (a < 0x1p-12 || xf != xf) ? xf : (float)__builtin_copysign(v, x);  // NaN: fmin dropped it
}

// log2 of a positive normal binary64 (a binary32 value): x = 2^e m, m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, to s^13 (relative error < 2^-40)
__device__ __forceinline__ double log2_pos(double x) {
    int e;
    const double f = __builtin_frexp(x, &e);  // [0.5, 1)
    const bool lo = f < 0x1.6a09e667f3bcdp-1;
    const double m = lo ? 2.0 * f : f;
    const double s = (m - 1.0) * recip(m + 1.0), z = s * s;
    double p = 1.0 / 13.0;
#pragma unroll
    for (int k = 11; k >= 1; k -= 2) p = __builtin_fma(p, z, 1.0 / (double)k);
    return __builtin_fma(2.0 * s * p, 0x1.71547652b82fep0, (double)(lo ? e - 1 : e));
}

// The special cases of pow (ufuncs.rs:166 semantics, as the binary64 routine), for the arguments
// the fast path below does not take.
__device__ __noinline__ float pow_special(float xf, float yf) {
    const double x = (double)xf, y = (double)yf;
    if (y == 0.0 || x == 1.0) return 1.0f;
    if (x != x || y != y) return (float)(x + y);
    const double ax = __builtin_fabs(x), ay = __builtin_fabs(y);
    const bool yint = __builtin_trunc(y) == y;
    const bool yodd = yint && ay < 0x1p53 && __builtin_trunc(ay * 0.5) != ay * 0.5;  // ay / 2 is exact
    const bool neg = __builtin_signbit(x) && yodd;
    if (ay == __builtin_inf()) {
        if (ax == 1.0) return 1.0f;
        return ((ax < 1.0) == (y < 0.0)) ? __builtin_inff() : 0.0f;
    }
    if (ax == __builtin_inf()) {
        const float v = y < 0.0 ? 0.0f : __builtin_inff();
        return neg ? -v : v;
    }
    if (ax == 0.0) {
        const float v = y < 0.0 ? __builtin_inff() : 0.0f;
        return neg ? __builtin_copysignf(v, xf) : v;
    }
    if (x < 0.0 && !yint) return __builtin_nanf("");
    const double t = __builtin_fmin(__builtin_fmax(y * log2_pos(ax), -1100.0), 1100.0);
    const double k = __builtin_rint(t);
    const double v = __builtin_ldexp(exp_taylor<10>((t - k) * 0x1.62e42fefa39efp-1), (int)k);
    return (float)(neg ? -v : v);
}

// pow: 2^(y log2 x) for finite x > 0 and finite y, the exponent's error |y| 2^-40 + 2^-53 |y log2 x|
// carried into the result's relative error (degree 10 for 2^f: (0.347)^11 / 11! < 2^-42); every
// other argument through pow_special.
__device__ inline float pow(float xf, float yf) {
    const double x = (double)xf, y = (double)yf;
    if (!(x > 0.0 && x < __builtin_inf() && __builtin_fabs(y) < __builtin_inf()) || y == 0.0 || x == 1.0)
        return pow_special(xf, yf);
    const double t = __builtin_fmin(__builtin_fmax(y * log2_pos(x), -1100.0), 1100.0);
    const double k = __builtin_rint(t);
    return (float)__builtin_ldexp(exp_taylor<10>((t - k) * 0x1.62e42fefa39efp-1), (int)k);
}

// Logarithms of binary32 values: log2_pos (relative error below 2^-40 of the log2 of the mantissa,
// and the exponent added exactly) scaled once by a binary64 constant, then rounded to binary32 --
// within 0.5 ulp + 2^-15 ulp. log1p forms 1 + x exactly in binary64 (x has 24 significant bits);
// |x| < 2^-25 returns x (log1p(x) = x (1 - x/2 + ...) is within half an ulp of it).
__device__ __forceinline__ bool log_special(float xf, float &r) {
    if (xf > 0.0f && xf < __builtin_inff()) return false;
    r = xf != xf ? xf : (xf == 0.0f ? -__builtin_inff() : (xf < 0.0f ? __builtin_nanf("") : xf));
    return true;
}
__device__ inline float log(float xf) {
    float r;
    if (log_special(xf, r)) return r;
    return (float)(log2_pos((double)xf) * 0x1.62e42fefa39efp-1);  // ln 2
}
__device__ inline float log2(float xf) {
    float r;
    if (log_special(xf, r)) return r;
    return (float)log2_pos((double)xf);
}
__device__ inline float log10(float xf) {
    float r;
    if (log_special(xf, r)) return r;
    return (float)(log2_pos((double)xf) * 0x1.34413509f79ffp-2);  // log10 2
}
__device__ inline float log1p(float xf) {
    if (xf != xf || xf == __builtin_inff()) return xf;
    if (xf == -1.0f) return -__builtin_inff();
    if (xf < -1.0f) return __builtin_nanf("");
    if (__builtin_fabsf(xf) < 0x1p-25f) return xf;
    return (float)(log2_pos(1.0 + (double)xf) * 0x1.62e42fefa39efp-1);
}

// sin and cos of binary32 values with |x| < 2^20: k = rint(|x| 2/pi), r = |x| - k pi/2 from the
// two leading parts of pi/2 (the product with the first exact, the second rounded once: |r| is at
// least ~2^-30 for binary32 inputs there, so r keeps 2^-36 relative), then Taylor series to r^11
// and r^12 on |r| <= pi/4 (first omitted terms below 6e-12 and 5e-13). Larger, infinite and NaN
// arguments take the binary64 routine.
__device__ __forceinline__ double sin_r11(double r) {
    const double z = r * r;
    double q = -InvFact<11>::v;
    q = __builtin_fma(q, z, InvFact<9>::v);
    q = __builtin_fma(q, z, -InvFact<7>::v);
    q = __builtin_fma(q, z, InvFact<5>::v);
    q = __builtin_fma(q, z, -InvFact<3>::v);
    return __builtin_fma(r * z, q, r);
}
__device__ __forceinline__ double cos_r12(double r) {
    const double z = r * r;
    double q = InvFact<12>::v;
    q = __builtin_fma(q, z, -InvFact<10>::v);
    q = __builtin_fma(q, z, InvFact<8>::v);
    q = __builtin_fma(q, z, -InvFact<6>::v);
    q = __builtin_fma(q, z, InvFact<4>::v);
    q = __builtin_fma(q, z, -0.5);
    return __builtin_fma(z, q, 1.0);
}
__device__ __forceinline__ double reduce_f32(double ax, int &q) {
    const double k = __builtin_rint(ax * 0x1.45f306dc9c883p-1);  // 2/pi
    q = (int)k & 3;
    return __builtin_fma(-k, 0x1.1a62633145c07p-54, __builtin_fma(-k, 0x1.921fb54442d18p+0, ax));
}
__device__ inline float sin(float xf) {
    const double x = (double)xf, ax = __builtin_fabs(x);
    if (!(ax < 0x1p20)) return (float)nrk::fm::sin(x);
    int q;
    const double r = reduce_f32(ax, q);
    const double v = (q & 1) ? cos_r12(r) : sin_r11(r);
    return (float)(__builtin_copysign(1.0, x) * ((q & 2) ? -v : v));  // sin(-x) = -sin(x)
}
__device__ inline float cos(float xf) {
    const double x = (double)xf, ax = __builtin_fabs(x);
    if (!(ax < 0x1p20)) return (float)nrk::fm::cos(x);
    int q;
    const double r = reduce_f32(ax, q);
    const double v = (q & 1) ? sin_r11(r) : cos_r12(r);
    return (float)((q == 1 || q == 2) ? -v : v);
}

}  // namespace f32

// The elementary functions by argument type: binary64 at full precision, binary32 through the
// routines above. Every map kernel of the library that evaluates exp, tanh, pow, a logarithm, sin
// or cos of a T calls these.
__device__ __forceinline__ double exp_t(double x) { return exp(x); }
__device__ __forceinline__ float exp_t(float x) { return f32::exp(x); }
__device__ __forceinline__ double tanh_t(double x) { return tanh(x); }
__device__ __forceinline__ float tanh_t(float x) { return f32::tanh(x); }
__device__ __forceinline__ double pow_t(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float pow_t(float x, float y) { return f32::pow(x, y); }
__device__ __forceinline__ double log_t(double x) { return log(x); }
__device__ __forceinline__ float log_t(float x) { return f32::log(x); }
__device__ __forceinline__ double log2_t(double x) { return log2(x); }
__device__ __forceinline__ float log2_t(float x) { return f32::log2(x); }
__device__ __forceinline__ double log10_t(double x) { return log10(x); }
__device__ __forceinline__ float log10_t(float x) { return f32::log10(x); }
__device__ __forceinline__ double log1p_t(double x) { return log1p(x); }
__device__ __forceinline__ float log1p_t(float x) { return f32::log1p(x); }
__device__ __forceinline__ double sin_t(double x) { return sin(x); }
__device__ __forceinline__ float sin_t(float x) { return f32::sin(x); }
__device__ __forceinline__ double cos_t(double x) { return cos(x); }
__device__ __forceinline__ float cos_t(float x) { return f32::cos(x); }

}  // namespace fm
}  // namespace nrk
