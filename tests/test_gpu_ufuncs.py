"""Elementwise ufuncs (ufuncs.rs:319-1240) through the op-coded map, on the GPU.

The reference's exp / log / tanh ... are Rust's f64 methods below 64 elements and a SIMD library
above (absent here); neither is bit-reproducible, so the bar is the mathematics: the error
against the exact value, computed in 80-bit long double (64-bit significand, 11 bits more than
binary64) on the CPU, measured in units of the last place of the binary64 result. Bounds, written
here: exp, exp2, expm1, log, log1p <= 1 ulp; log2, log10 <= 1 ulp; tanh, sinh, cosh, asinh,
atanh <= 2 ulp; acosh <= 3 ulp (near 1 its result ~sqrt(2t) collects the roundings of
2t + t^2, the square root, t + s and log1p); atan <= 1.5 ulp, asin, acos <= 2 ulp; sin, cos <= 1 ulp and
tan <= 2.5 ulp for every finite argument; sigmoid
(the reference's 1 / (1 + exp(-x)): exp's relative error passes through two more IEEE roundings,
and a relative error converts to up to twice as many ulps at a binade edge) <= 3 ulp; binary32
inputs <= 1 ulp of binary32 (exp, tanh and pow through the binary32 routines, whose error before
the one rounding is below 2^-31 relative; the others the FP64 evaluation, rounded once). Special
values (NaN, +-inf, +-0, overflow and underflow edges) must match exactly.

Exact ops are bit-exact against an independent computation: floor / ceil / trunc (IEEE), round
(half away from zero: Python's decimal ROUND_HALF_UP), fract (math.modf; Rust's x - x.trunc()
gives NaN for +-inf), sign (ufuncs.rs:756), clamp, and the single IEEE steps of reciprocal,
rsqrt, deg2rad and rad2deg."""
import decimal
import math

import numpy as np
import pytest

from numrs_amd import kernels as K

pytestmark = pytest.mark.gpu
LD = np.longdouble


def ulp_err(got, exact_ld, dtype=np.float64):
    """|got - exact| in ulps of the correctly rounded result (0 where both are the same special)."""
    got = np.asarray(got, dtype)
    ref = exact_ld.astype(dtype)
    err = np.zeros(got.shape, np.float64)
    fin = np.isfinite(ref) & np.isfinite(got)
    tiny = np.finfo(dtype).smallest_subnormal
    sp = np.maximum(np.spacing(np.abs(ref[fin])).astype(LD), LD(tiny))
    err[fin] = (np.abs(got[fin].astype(LD) - exact_ld[fin]) / sp).astype(np.float64)
    bad = ~fin & ~((np.isnan(ref) & np.isnan(got)) | (ref == got))
    err[bad] = np.inf
    return err


def sample(lo, hi, n, seed):
    r = np.random.default_rng(seed)
    return r.uniform(lo, hi, n)


SPECIAL = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 1.0, -1.0, 5e-324, -5e-324, 2.2250738585072014e-308,
                    1e-300, -1e-300, 1e300, -1e300, 0.5, 2.0])

TRIG_HARD = np.array([6381956970095103 * 2.0 ** 797, 5.319372648326541e+255, 1e22, 2.0 ** 20, 2.0 ** 20 + 1,
                      2.0 ** 20 - 1, 1.5707963267948966, 3.141592653589793, 0.7853981633974483, 0.785398163397448,
                      1e300, 1.7976931348623157e308, -1.7976931348623157e308, 355.0, 103993.0])

CASES = {
    # name: (gpu fn, long double fn, inputs, bound)
    "exp": (K.exp, np.exp, np.concatenate([sample(-745.2, 709.8, 60000, 1), sample(-1e-3, 1e-3, 5000, 2),
                                          [709.782712893383, 709.7827128933841, -745.1332191019411, -745.14,
                                           -708.39, -720.5, 1e-18, -1e-18], SPECIAL]), 1.0),
    "exp2": (K.exp2, np.exp2, np.concatenate([sample(-1075.5, 1024.5, 60000, 3), sample(-1, 1, 5000, 4),
                                             [1023.9999, 1024.0, -1074.0, -1074.5, -1022.5, 0.25], SPECIAL]), 1.0),
    "expm1": (K.expm1, np.expm1, np.concatenate([sample(-45, 709.7, 40000, 5), sample(-0.4, 0.4, 20000, 6),
                                                sample(-1e-12, 1e-12, 5000, 7), [-36.7, -37.4, -38.0, 0.3465, -0.3466,
                                                                                 0.6931, 36.0, 37.0], SPECIAL]), 1.0),
    "log": (K.log, np.log, np.concatenate([np.exp(sample(-744, 709, 40000, 8)), 1.0 + sample(-1e-3, 1e-3, 10000, 9),
                                          sample(0.5, 2.0, 10000, 10), [3e-320, 1e-310, 0.7071067811865476,
                                                                       1.4142135623730951], SPECIAL]), 1.0),
    "log2": (K.log2, np.log2, np.concatenate([np.exp(sample(-744, 709, 40000, 11)), 1.0 + sample(-1e-3, 1e-3, 10000, 12),
                                             2.0 ** np.arange(-1074, 1024, 7.0), SPECIAL]), 1.0),
    "log10": (K.log10, np.log10, np.concatenate([np.exp(sample(-744, 709, 40000, 13)), 1.0 + sample(-1e-3, 1e-3, 10000, 14),
                                                10.0 ** np.arange(-300, 300, 3.0), SPECIAL]), 1.0),
    "log1p": (K.log1p, np.log1p, np.concatenate([sample(-1, 1, 30000, 15), np.exp(sample(-700, 700, 20000, 16)),
                                                sample(-1e-10, 1e-10, 5000, 17), [-0.999999999999, -0.5, 1e-17, -1e-17,
                                                                                  -2.0], SPECIAL]), 1.0),
    "tanh": (K.tanh, np.tanh, np.concatenate([sample(-25, 25, 40000, 18), sample(-1, 1, 20000, 19),
                                             sample(-1e-9, 1e-9, 5000, 20), [22.0, -22.0, 21.99, 0.5493, 3.7e-9], SPECIAL]), 2.0),
    "sinh": (K.sinh, np.sinh, np.concatenate([sample(-711, 711, 20000, 50), sample(-25, 25, 20000, 51),
                                             sample(-1, 1, 10000, 52), [709.78, 710.4, 710.5, -710.4, 22.0], SPECIAL]), 2.0),
    "cosh": (K.cosh, np.cosh, np.concatenate([sample(-711, 711, 20000, 53), sample(-25, 25, 20000, 54),
                                             sample(-1, 1, 10000, 55), [709.78, 710.4, 0.3465, 22.0], SPECIAL]), 2.0),
    "asinh": (K.arcsinh, np.arcsinh, np.concatenate([sample(-10, 10, 20000, 56), np.exp(sample(-30, 700, 20000, 57)),
                                                    -np.exp(sample(-30, 700, 5000, 58)), [2.0, 2.0000001, 2.0**28], SPECIAL]), 2.0),
    "acosh": (K.arccosh, np.arccosh, np.concatenate([1 + np.exp(sample(-40, 2, 20000, 59)), np.exp(sample(0, 700, 20000, 60)),
                                                    [1.0, 2.0, 2.0**28, 0.999], SPECIAL]), 3.0),
    "atanh": (K.arctanh, np.arctanh, np.concatenate([sample(-1, 1, 30000, 61), 1 - np.exp(sample(-36, 0, 10000, 62)),
                                                    sample(-1e-8, 1e-8, 5000, 63), [0.5, -0.5, 1.0, -1.0, 1.5]], ), 2.0),
    "atan": (K.arctan, np.arctan, np.concatenate([sample(-1, 1, 20000, 64), sample(-20, 20, 20000, 65), np.exp(sample(-30, 700, 10000, 66)),
                                                 -np.exp(sample(-30, 700, 5000, 67)), [1.0, -1.0, 0.0625, 0.1875, 0.9375], SPECIAL]), 1.5),
    "asin": (K.arcsin, np.arcsin, np.concatenate([sample(-1, 1, 30000, 68), 1 - np.exp(sample(-36, -1, 10000, 69)),
                                                 sample(-1e-9, 1e-9, 5000, 70), [0.5, -0.5, 1.0000000000000002]], ), 2.0),
    "acos": (K.arccos, np.arccos, np.concatenate([sample(-1, 1, 30000, 71), 1 - np.exp(sample(-36, -1, 10000, 72)),
                                                 -1 + np.exp(sample(-36, -1, 10000, 73)), [0.5, -0.5]], ), 2.0),
    # every finite argument: Cody-Waite below 2^20, Payne-Hanek above (incl. 6381956970095103 2^797,
    # the binary64 closest to a multiple of pi/2, and the largest double)
    "sin": (K.sin, np.sin, np.concatenate([sample(-10, 10, 20000, 74), sample(-1e6, 1e6, 10000, 75), np.exp(sample(14, 709, 15000, 76)),
                                         -np.exp(sample(14, 709, 5000, 77)), TRIG_HARD, SPECIAL]), 1.0),
    "cos": (K.cos, np.cos, np.concatenate([sample(-10, 10, 20000, 78), sample(-1e6, 1e6, 10000, 79), np.exp(sample(14, 709, 15000, 80)),
                                         -np.exp(sample(14, 709, 5000, 81)), TRIG_HARD, SPECIAL]), 1.0),
    "tan": (K.tan, np.tan, np.concatenate([sample(-10, 10, 20000, 82), sample(-1e6, 1e6, 10000, 83), np.exp(sample(14, 709, 15000, 84)),
                                         -np.exp(sample(14, 709, 5000, 85)), TRIG_HARD, SPECIAL]), 2.5),
}


@pytest.mark.parametrize("name", list(CASES))
def test_elementary_f64_within_ulp_bound(gpu, name):
    fn, ref_fn, x, bound = CASES[name]
    x = np.asarray(x, np.float64)
    with np.errstate(all="ignore"):
        exact = ref_fn(x.astype(LD))
    got = fn(x)
    err = ulp_err(got, exact)
    worst = int(np.argmax(err))
    print(f"{name}: max {err.max():.3f} ulp at x = {x[worst]!r} (got {got[worst]!r}, exact {exact[worst]!r})")
    assert err.max() <= bound, (name, x[worst], got[worst], exact[worst])
    # special values exactly, sign of zero included
    with np.errstate(all="ignore"):
        ref_sp = ref_fn(SPECIAL.astype(LD)).astype(np.float64)
    got_sp = fn(SPECIAL)
    same = (np.isnan(ref_sp) & np.isnan(got_sp)) | ((ref_sp == got_sp) & (np.signbit(ref_sp) == np.signbit(got_sp)))
    ok = same | (np.isfinite(ref_sp) & (ulp_err(got_sp, ref_fn(SPECIAL.astype(LD))) <= bound))
    assert np.all(ok), (name, SPECIAL[~ok], got_sp[~ok], ref_sp[~ok])


@pytest.mark.parametrize("name", list(CASES))
def test_elementary_f32_within_one_ulp(gpu, name):
    fn, ref_fn, x, _ = CASES[name]
    x32 = np.asarray(x, np.float64).astype(np.float32)
    with np.errstate(all="ignore"):
        exact = ref_fn(x32.astype(LD))
    got = fn(x32)
    assert got.dtype == np.float32
    err = ulp_err(got, exact, np.float32)
    print(f"{name} f32: max {err.max():.3f} ulp")
    assert err.max() <= 1.0


def test_sigmoid_within_three_ulp(gpu):
    """ufuncs.rs:1235 computes 1 / (1 + (-x).exp()) in f64: below x = -709.78 (-x).exp() overflows
    and the result is exactly 0, which the GPU reproduces; elsewhere within 3 ulp of the exact value."""
    x = np.concatenate([sample(-40, 40, 40000, 21), sample(-709.7, 745, 5000, 22), [0.0, -0.0, np.inf, np.nan]])
    with np.errstate(all="ignore"):
        exact = LD(1) / (LD(1) + np.exp(-x.astype(LD)))
    err = ulp_err(K.sigmoid(x), exact)
    print(f"sigmoid: max {err.max():.3f} ulp")
    assert err.max() <= 3.0
    under = np.array([-709.8, -720.0, -745.5, -1e300, -np.inf])
    np.testing.assert_array_equal(K.sigmoid(under), np.zeros(5))


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def assert_same(got, want):
    """Bit-identical, except that any NaN matches any NaN (the sign of a NaN produced by an
    invalid operation differs between x86 and the GPU)."""
    got, want = np.asarray(got), np.asarray(want)
    assert got.dtype == want.dtype
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_array_equal(bits(got[ok]), bits(want[ok]))


EXACT_X = np.concatenate([np.array([0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 0.49999999999999994, -0.49999999999999994,
                                    4503599627370495.5, -4503599627370495.5, 4503599627370497.0, 1e300, -1e300,
                                    0.0, -0.0, 5e-324, -5e-324, np.inf, -np.inf, np.nan, 3.0, -3.0, 0.3, -0.7]),
                          sample(-1e6, 1e6, 20000, 30), sample(-4, 4, 20000, 31)])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rounding_ops_bit_exact(gpu, dtype):
    x = EXACT_X.astype(dtype)
    assert_same(K.floor(x), np.floor(x))
    assert_same(K.ceil(x), np.ceil(x))
    assert_same(K.trunc(x), np.trunc(x))
    # Rust's round: half away from zero (decimal ROUND_HALF_UP), sign of zero kept
    want = np.array([v if not math.isfinite(v) else
                     math.copysign(float(decimal.Decimal(float(v)).to_integral_value(rounding=decimal.ROUND_HALF_UP)), v)
                     for v in x.astype(np.float64)]).astype(dtype)
    assert_same(K.round(x), want)
    # Rust's fract = x - x.trunc(): the C fractional part for finite x, but +0.0 (not modf's -0.0)
    # for integral x, as x - x is +0 in IEEE arithmetic; NaN for +-inf
    with np.errstate(invalid="ignore"):
        want = np.array([math.modf(float(v))[0] + 0.0 if math.isfinite(v) else math.nan
                         for v in x.astype(np.float64)]).astype(dtype)
    assert_same(K.fract(x), want)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sign_clip_and_ieee_steps_bit_exact(gpu, dtype):
    x = EXACT_X.astype(dtype)
    # ufuncs.rs:756: 0.0 for +-0, 1 above zero, -1 otherwise (NaN included)
    want = np.array([0.0 if v == 0 else (1.0 if v > 0 else -1.0) for v in x.astype(np.float64)]).astype(dtype)
    assert_same(K.sign(x), want)
    # clamp(-2.5, 7.0): NaN passes, -0.0 stays -0.0
    got = K.clip(x, -2.5, 7.0)
    want = np.array([v if v != v else (-2.5 if v < -2.5 else (7.0 if v > 7.0 else v)) for v in x.astype(np.float64)]).astype(dtype)
    assert_same(got, want)
    with pytest.raises(ValueError):
        K.clip(x, 3.0, 1.0)
    with np.errstate(all="ignore"):
        assert_same(K.reciprocal(x), dtype(1) / x)
        assert_same(K.rsqrt(x), dtype(1) / np.sqrt(x))
    assert_same(K.deg2rad(x), x * dtype(np.pi / 180.0))
    assert_same(K.rad2deg(x), x * dtype(180.0 / np.pi))


def test_elementary_steps_compose_in_one_chain(gpu):
    """log1p(exp(x)) as one two-step chain equals the two single-step maps, bit for bit."""
    x = sample(-30, 30, 10000, 40)
    assert_same(K.unary_dispatch(x, [("exp",), ("log1p",)]), K.log1p(K.exp(x)))


# ---- binary ufuncs on the zip kernel (ufuncs.rs:587, 784, 841, 916)

def _wide(n, seed):
    r = np.random.default_rng(seed)
    return r.choice([-1.0, 1.0], n) * np.exp(r.uniform(-690, 690, n))


def test_copysign_bit_exact(gpu):
    x = EXACT_X
    s = EXACT_X[::-1].copy()
    assert_same(K.copysign(x, s), np.copysign(x, s))


@pytest.mark.parametrize("name,bound", [("hypot", 1.0), ("atan2", 1.5), ("logaddexp", 2.0)])
def test_binary_elementary_within_ulp_bound(gpu, name, bound):
    r = np.random.default_rng(90)
    a = np.concatenate([_wide(20000, 91), r.uniform(-10, 10, 20000), r.uniform(-1e-3, 1e-3, 5000)])
    b = np.concatenate([_wide(20000, 92), r.uniform(-10, 10, 20000), r.uniform(-1e3, 1e3, 5000)])
    fn = {"hypot": K.hypot, "atan2": K.arctan2, "logaddexp": K.logaddexp}[name]
    al, bl = a.astype(LD), b.astype(LD)
    with np.errstate(all="ignore"):
        if name == "hypot":
            exact = np.hypot(al, bl)
        elif name == "atan2":
            exact = np.arctan2(al, bl)
        else:
            exact = None
    got = fn(a, b)
    if exact is None:
        # logaddexp is the reference's f64 expression m + ln(e^-|a-m| + e^-|b-m|) (ufuncs.rs:933-936),
        # whose sum absorbs the smaller term: compare it with that expression evaluated in f64 with
        # correctly rounded exp and ln. One ulp of exp can move the rounded sum s in [1, 2] by
        # 2^-52, so ln s by 2^-52, plus ln's own ulp (<= 2^-53 on [0, ln 2]) and the final rounding.
        m = np.fmax(a, b)
        s = (np.exp(-np.abs(a - m).astype(LD)).astype(np.float64) + np.exp(-np.abs(b - m).astype(LD)).astype(np.float64))
        want = m + np.log(s.astype(LD)).astype(np.float64)
        tol = 1.5 * 2.0**-52 + np.spacing(np.abs(want))
        dev = np.abs(got - want) / tol
        w = int(np.argmax(dev))
        print(f"{name}: max {dev.max():.3f} of the bound at ({a[w]!r}, {b[w]!r})")
        assert dev.max() <= 1.0
        err = ulp_err(got, np.asarray(want, LD))
        bound = np.inf
    else:
        err = ulp_err(got, exact)
    w = int(np.argmax(err))
    print(f"{name}: max {err.max():.3f} ulp at ({a[w]!r}, {b[w]!r})")
    assert err.max() <= bound
    # special values: C / Rust semantics (numpy's float64 hypot / arctan2 follow C), the
    # reference's formula for logaddexp
    sp = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 5e-324, 1e308, -2.5])
    A, B = np.meshgrid(sp, sp)
    A, B = A.ravel(), B.ravel()
    with np.errstate(all="ignore"):
        if name == "hypot":
            want = np.hypot(A, B)
        elif name == "atan2":
            want = np.arctan2(A, B)
        else:
            m = np.fmax(A, B)
            want = m + np.log(np.exp(-np.abs(A - m)) + np.exp(-np.abs(B - m)))
    got = fn(A, B)
    same = (np.isnan(want) & np.isnan(got)) | ((want == got) & (np.signbit(want) == np.signbit(got)))
    close = np.isfinite(want) & np.isfinite(got) & (np.abs(got - want) <= 2 * np.spacing(np.abs(want)))
    assert np.all(same | close), (name, A[~(same | close)], B[~(same | close)], got[~(same | close)], want[~(same | close)])


def test_binary_f32_within_one_ulp(gpu):
    r = np.random.default_rng(93)
    a = r.uniform(-100, 100, 20000).astype(np.float32)
    b = r.uniform(-100, 100, 20000).astype(np.float32)
    for fn, ref in ((K.hypot, np.hypot), (K.arctan2, np.arctan2)):
        got = fn(a, b)
        assert got.dtype == np.float32
        assert ulp_err(got, ref(a.astype(LD), b.astype(LD)), np.float32).max() <= 1.0


# ---- activations, special functions and steps (ufuncs.rs:520-1480)
#
# Parity: each is the reference's f64 expression, one IEEE rounding per operation, over exp / ln /
# tanh / sin. Those differ from the correctly rounded values by at most K ulps (K = 1 for exp, ln
# and sin, 2 for tanh, from their bounds above), and every expression is monotone in each of their
# values, so the result must lie between the expression's values with each function's correctly
# rounded output moved -K and +K ulps (all corners when several appear).

_K = {"exp": 1, "log": 1, "tanh": 2, "sin": 1}
_CR = {"exp": np.exp, "log": np.log, "tanh": np.tanh, "sin": np.sin}
_PI = 3.141592653589793
_GELU_C = np.sqrt(np.float64(2.0) / np.float64(_PI))


def _erf_expr(x, F):
    t = 1.0 / (1.0 + 0.5 * np.abs(x))
    p = np.float64(0.17087277)
    for c in (-0.82215223, 1.48851587, -1.13520398, 0.27886807, -0.18628806, 0.09678418, 0.37409196,
              1.00002368):
        p = c + t * p
    tau = t * F["exp"](-x * x - 1.26551223 + t * p)
    return np.where(x >= 0.0, 1.0 - tau, tau - 1.0)


def _digamma_expr(x, F):
    out = np.full_like(x, np.nan)
    for i, v in enumerate(x):
        if not v <= 0.0:
            xs = [v]
            while xs[-1] < 6.0:
                xs.append(xs[-1] + 1.0)
            ix = 1.0 / xs[-1]
            ix2 = ix * ix
            r = F["log"](np.array([xs[-1]]))[0] - 0.5 * ix - ix2 * (1.0 / 12.0 - ix2 * (1.0 / 120.0 - ix2 / 252.0))
            for xi in reversed(xs[:-1]):
                r = r - 1.0 / xi
            out[i] = r
    return out


_SELU_L, _SELU_A = 1.0507009873554805, 1.6732632423543772
ACT = {  # name: (kernel call, reference expression, functions it calls)
    "relu": (K.relu, lambda x, F: np.where(x > 0, x, 0.0), ()),
    "leaky_relu": (lambda x: K.leaky_relu(x, 0.2), lambda x, F: np.where(x > 0, x, 0.2 * x), ()),
    "elu": (lambda x: K.elu(x, 1.5), lambda x, F: np.where(x > 0, x, 1.5 * (F["exp"](x) - 1.0)), ("exp",)),
    "selu": (K.selu, lambda x, F: _SELU_L * np.where(x > 0, x, _SELU_A * (F["exp"](x) - 1.0)), ("exp",)),
    "hardsigmoid": (K.hardsigmoid, lambda x, F: np.where(x <= -3, 0.0, np.where(x >= 3, 1.0, x / 6.0 + 0.5)), ()),
    "hardswish": (K.hardswish, lambda x, F: np.where(x <= -3, 0.0, np.where(x >= 3, x, x * (x + 3.0) / 6.0)), ()),
    "softplus": (K.softplus, lambda x, F: np.where(x > 20, x, np.where(x < -20, F["exp"](x), F["log"](1.0 + F["exp"](x)))), ("exp", "log")),
    "mish": (K.mish, lambda x, F: x * F["tanh"](np.where(x > 20, x, F["log"](1.0 + F["exp"](x)))), ("exp", "log", "tanh")),
    "silu": (K.silu, lambda x, F: x / (1.0 + F["exp"](-x)), ("exp",)),
    "gelu": (K.gelu, lambda x, F: 0.5 * x * (1.0 + F["tanh"](_GELU_C * (x + 0.044715 * (x * x * x)))), ("tanh",)),
    "logit": (K.logit, lambda x, F: F["log"](x / (1.0 - x)), ("log",)),
    "sinc": (K.sinc, lambda x, F: np.where(x == 0, 1.0, F["sin"](_PI * x) / (_PI * x)), ("sin",)),
    "erf": (K.erf, _erf_expr, ("exp",)),
    "erfc": (K.erfc, lambda x, F: 1.0 - _erf_expr(x, F), ("exp",)),
    "digamma": (K.digamma, _digamma_expr, ("log",)),
}


def _fn(name, d):
    def f(*v):
        with np.errstate(all="ignore"):
            r = _CR[name](*[np.asarray(u, np.float64).astype(LD) for u in v]).astype(np.float64)
        for _ in range(abs(d)):
            r = np.nextafter(r, np.inf if d > 0 else -np.inf)
        return r
    return f


def _bracket(expr, fns, x):
    """(centre, lo, hi) of the expression over every corner of the +-K ulp box."""
    import itertools
    with np.errstate(all="ignore"):
        centre = expr(x, {n: _fn(n, 0) for n in _CR})
        vals = [expr(x, {**{n: _fn(n, 0) for n in _CR}, **{n: _fn(n, s * _K[n]) for n, s in zip(fns, signs)}})
                for signs in itertools.product((-1, 1), repeat=len(fns))] or [centre]
    v = np.stack(vals)
    return centre, v.min(axis=0), v.max(axis=0)


def _act_inputs(name):
    r = np.random.default_rng(hash(name) % 1000)
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 20.0, -20.0, 3.0, -3.0, 1e-310, -1e-310, 709.8,
                   -709.8, 745.5, -745.5, 0.5, 1.0, -1.0, 6.0, 1e-12, 1e300, -1e300])
    if name == "digamma":
        return np.concatenate([sp, r.uniform(0, 12, 3000), np.exp(r.uniform(-40, 40, 1000)), r.uniform(-5, 0, 200)])
    if name == "logit":
        return np.concatenate([sp, r.uniform(0, 1, 20000), r.uniform(-2, 3, 2000)])
    return np.concatenate([sp, r.uniform(-8, 8, 20000), r.uniform(-40, 40, 10000), _wide(5000, 95)])


@pytest.mark.parametrize("name", list(ACT))
def test_activation_matches_reference_expression(gpu, name):
    call, expr, fns = ACT[name]
    x = _act_inputs(name)
    got = call(x)
    centre, lo, hi = _bracket(expr, fns, x)
    ok = (np.isnan(centre) & np.isnan(got)) | ((got >= lo) & (got <= hi))
    assert np.all(ok), (name, x[~ok][:5], got[~ok][:5], centre[~ok][:5])
    if not fns:  # no elementary function: the expression itself, to the bit (zeros by value)
        assert np.array_equal(got, centre, equal_nan=True)
    exact_frac = np.mean((got == centre) | (np.isnan(got) & np.isnan(centre)))
    print(f"{name}: {exact_frac:.4f} equal to the correctly rounded evaluation")
    # binary32 arrays: the same f64 evaluation, rounded once
    x32 = x.astype(np.float32)
    c32, lo32, hi32 = _bracket(expr, fns, x32.astype(np.float64))
    g32 = call(x32)
    assert g32.dtype == np.float32
    ok32 = (np.isnan(c32) & np.isnan(g32)) | ((g32 >= lo32.astype(np.float32)) & (g32 <= hi32.astype(np.float32)))
    assert np.all(ok32), (name, x32[~ok32][:5], g32[~ok32][:5], c32[~ok32][:5])


def test_activations_against_their_mathematics(gpu):
    """Pins the expressions (kernel and test alike) to what they approximate: a dropped term, a
    wrong constant or sign fails here."""
    import math
    import scipy.special as sps
    r = np.random.default_rng(96)
    x = r.uniform(-15, 15, 20000)
    xl = x.astype(LD)
    np.testing.assert_array_equal(K.relu(x), np.maximum(x, 0.0))
    np.testing.assert_array_equal(K.hardsigmoid(x), np.clip(x / 6.0 + 0.5, 0.0, 1.0))
    np.testing.assert_allclose(K.hardswish(x), x * np.clip(x + 3.0, 0.0, 6.0) / 6.0, rtol=2**-51, atol=0)
    np.testing.assert_allclose(K.softplus(x), np.logaddexp(0.0, x), rtol=1e-7, atol=0)
    np.testing.assert_allclose(K.silu(x), x * sps.expit(x), rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(K.elu(x, 1.5), np.where(x > 0, x, 1.5 * np.expm1(np.minimum(x, 0))), rtol=0, atol=4e-16)
    np.testing.assert_allclose(K.selu(x), _SELU_L * np.where(x > 0, x, _SELU_A * np.expm1(np.minimum(x, 0))),
                               rtol=1e-15, atol=1e-15)
    mish = (xl * np.tanh(np.log1p(np.exp(xl)))).astype(np.float64)
    # 1 + e^x rounds away e^x's low bits: ln(1 + e^x) has an absolute error near 2^-53
    assert np.all(np.abs(K.mish(x) - mish) <= 1e-13 * np.abs(mish) + 4e-16 * np.abs(x))
    gelu = (0.5 * xl * (1 + np.tanh(np.sqrt(LD(2) / np.pi) * (xl + LD(0.044715) * xl**3)))).astype(np.float64)
    assert np.all(np.abs(K.gelu(x) - gelu) <= 1e-13 * np.abs(gelu) + 1e-15 * np.abs(x))
    s = x / 4.0
    np.testing.assert_allclose(K.sinc(s), np.sinc(s), rtol=0, atol=1e-15)
    p = r.uniform(1e-6, 1 - 1e-6, 20000)
    np.testing.assert_allclose(K.logit(p), np.log(p / (1 - p)), rtol=1e-14, atol=1e-15)
    e = np.array([math.erf(v) for v in x])
    assert np.max(np.abs(K.erf(x) - e)) <= 1.2e-7  # the fit's stated error
    ec = np.array([math.erfc(v) for v in x])
    assert np.all(np.abs(K.erfc(x) - ec) <= 1.2e-7 * ec + 2e-16)
    d = np.concatenate([r.uniform(1e-3, 30, 20000), [1e-10, 0.5, 1.0, 6.0]])
    # the series is cut after x^-6: the first dropped term is 1/(240 x^8), 2.5e-9 at x = 6
    assert np.all(np.abs(K.digamma(d) - sps.digamma(d)) <= 3e-9 + 1e-15 * np.abs(sps.digamma(d)))
    assert np.isnan(K.digamma(np.array([0.0, -1.0, -2.5]))).all()


def test_prelu_and_steps_exact(gpu):
    r = np.random.default_rng(97)
    a, al = r.normal(size=30001), r.uniform(0, 1, 30001)
    a[::97] = np.nan
    np.testing.assert_array_equal(K.prelu(a, al), np.where(a > 0, a, al * a))
    with pytest.raises(ValueError):
        K.prelu(a, al[:-1])
    t = np.array([0.0, 0.25, 0.5, 1.0, -1.0, 2.0, np.nan])
    np.testing.assert_array_equal(K.smoothstep(0.0, 1.0, t), [0.0, 0.15625, 0.5, 1.0, 0.0, 1.0, np.nan])
    np.testing.assert_array_equal(K.smootherstep(0.0, 1.0, t), [0.0, 0.103515625, 0.5, 1.0, 0.0, 1.0, np.nan])
    x = r.uniform(-3, 5, 50000)
    e0, e1 = -1.25, 3.5
    tt = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    np.testing.assert_array_equal(K.smoothstep(e0, e1, x), tt * tt * (3.0 - 2.0 * tt))
    np.testing.assert_array_equal(K.smootherstep(e0, e1, x), tt * tt * tt * (tt * (tt * 6.0 - 15.0) + 10.0))


# ---- power (ufuncs.rs:166, 226, 305) and the Lanczos gamma (:1122-1195)

def _pow_cases():
    r = np.random.default_rng(98)
    x = np.exp(r.uniform(-700, 700, 30000))
    y = r.uniform(-1, 1, 30000) * np.minimum(700 / np.abs(np.log(x)), 1e6)  # |y ln x| < 700
    x2 = 1.0 + r.uniform(-1, 1, 5000) * 2.0**-30  # near 1 with large y
    y2 = r.uniform(-1, 1, 5000) * 2.0**38
    x3 = -np.exp(r.uniform(-5, 5, 5000))  # negative base, integer exponent
    y3 = np.floor(r.uniform(-60, 60, 5000))
    x4 = r.uniform(0, 10, 5000)  # subnormal and near-overflow results
    y4 = np.where(r.random(5000) < 0.5, -1074.0, 1023.0) / np.maximum(np.log2(x4), 0.01) * r.uniform(0.98, 1.0, 5000)
    return np.concatenate([x, x2, x3, x4]), np.concatenate([y, y2, y3, y4])


def test_power_within_one_ulp(gpu):
    x, y = _pow_cases()
    got = K.power(x, y)
    with np.errstate(all="ignore"):
        exact = np.power(x.astype(LD), y.astype(LD))
    err = ulp_err(got, exact)
    w = int(np.argmax(err))
    print(f"pow: max {err.max():.3f} ulp at ({x[w]!r}, {y[w]!r})")
    assert err.max() <= 1.0
    np.testing.assert_array_equal(K.power_scalar(x[:1000], 2.5), K.power(x[:1000], np.full(1000, 2.5)))
    # C's (and Rust's powf's) special cases, to the bit
    sp = np.array([0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 0.5, -0.5, 3.0, -3.0, np.inf, -np.inf, np.nan, 5e-324, 1e308])
    A, B = np.meshgrid(sp, sp)
    A, B = A.ravel(), B.ravel()
    with np.errstate(all="ignore"):
        pg, pw = K.power(A, B), np.power(A, B)
    # zeros, infinities, NaN and +-1 to the bit; ordinary finite results (0.5^0.5) within 1 ulp
    near = np.isfinite(pw) & (pw != 0) & (np.abs(pg - pw) <= np.spacing(np.abs(pw)))
    bad = ~((np.isnan(pg) & np.isnan(pw)) | ((pg == pw) & (np.signbit(pg) == np.signbit(pw))) | near)
    assert not bad.any(), list(zip(A[bad], B[bad], pg[bad], pw[bad]))
    # broadcasting, as the power ufunc does
    np.testing.assert_array_equal(K.power(np.array([[2.0], [3.0]]), np.array([0.0, 1.0, 2.0])), [[1, 2, 4], [1, 3, 9]])
    # binary32 (the binary32 routine: exponent error |y| 2^-48 + 2^-53 |y log2 x|, then one
    # rounding): within 1 ulp, the bound of libm's powf that Rust's f32::powf calls
    x32, y32 = np.abs(x[:20000]).astype(np.float32), (y[:20000] / 50).astype(np.float32)
    g32 = K.power(x32, y32)
    with np.errstate(all="ignore"):
        e32 = np.power(x32.astype(LD), y32.astype(LD))
    fin = np.isfinite(e32.astype(np.float32)) & (np.abs(e32) > 1e-37)
    err32 = ulp_err(g32[fin], e32[fin], np.float32)
    print(f"power f32: max {err32.max():.3f} ulp")
    assert err32.max() <= 1.0
    # the scalar-exponent map and the elementwise zip give the same bits in binary32 too
    np.testing.assert_array_equal(K.power_scalar(x32[:1000], np.float32(2.5)), K.power(x32[:1000], np.full(1000, 2.5, np.float32)))


_SQRT_2PI = float(np.sqrt(np.float64(2.0) * np.float64(np.pi)))
_LANCZOS = [0.99999999999980993, 676.5203681218851, -1259.1392167224028, 771.32342877765313, -176.61502916214059,
            12.507343278686905, -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7]


def _lanczos_expr(x, F):
    xm = x - 1.0
    y = np.full_like(x, _LANCZOS[0])
    for i in range(1, 9):
        y = y + _LANCZOS[i] / (xm + float(i))
    t = xm + 7.0 + 0.5
    return _SQRT_2PI * F["pow"](t, xm + 0.5) * F["exp"](-t) * y


def _gamma_expr(x, F):
    with np.errstate(all="ignore"):
        refl = _PI / (F["sin"](_PI * x) * _lanczos_expr(1.0 - x, F))
        main = _lanczos_expr(x, F)
    out = np.where(x < 0.5, refl, main)
    return np.where((x <= 0) & (np.floor(x) == x), np.inf, out)


_K["pow"] = 1
_CR["pow"] = np.power
ACT["gamma"] = (K.gamma, _gamma_expr, ("pow", "exp", "sin"))
ACT["lgamma"] = (K.lgamma, lambda x, F: F["log"](np.abs(_lanczos_expr(x, F))), ("pow", "exp", "log"))


@pytest.mark.parametrize("name", ["gamma", "lgamma"])
def test_gamma_matches_reference_expression(gpu, name):
    r = np.random.default_rng(99)
    x = np.concatenate([[0.5, 1.0, 2.0, 0.0, -1.0, -2.5, 141.0, 150.0, 171.5, np.nan, np.inf, 1e-300],
                        r.uniform(0.5, 30, 5000), r.uniform(-20, 0.5, 3000), r.uniform(30, 180, 1000)])
    if name == "lgamma":
        x = x[(x > 0) | np.isnan(x)]
    _check_act(name, x)


def _check_act(name, x):
    call, expr, fns = ACT[name]
    got = call(x)
    centre, lo, hi = _bracket(expr, fns, x)
    ok = (np.isnan(centre) & np.isnan(got)) | ((got >= lo) & (got <= hi))
    assert np.all(ok), (name, x[~ok][:5], got[~ok][:5], centre[~ok][:5])
    print(f"{name}: {np.mean((got == centre) | (np.isnan(got) & np.isnan(centre))):.4f} equal to the correctly rounded evaluation")


def test_gamma_against_its_mathematics(gpu):
    """The Lanczos fit (g = 7, n = 9) is good to ~1e-15 relative; the reference's powf overflows
    t^(x + 0.5) past x ~ 141, so the comparison stops there."""
    import scipy.special as sps
    r = np.random.default_rng(100)
    x = r.uniform(0.5, 140, 20000)
    np.testing.assert_allclose(K.gamma(x), sps.gamma(x), rtol=5e-13)
    xn = r.uniform(-10, 0.5, 20000)
    xn = xn[np.abs(xn - np.round(xn)) > 1e-3]
    np.testing.assert_allclose(K.gamma(xn), sps.gamma(xn), rtol=1e-10)
    lg = K.lgamma(x)
    assert np.all(np.abs(lg - sps.gammaln(x)) <= 5e-13 * (1 + np.abs(sps.gammaln(x))))
    assert K.gamma(np.array([5.0]))[0] == pytest.approx(24.0, rel=1e-14)
    assert K.gamma(np.array([150.0]))[0] == np.inf  # the reference's overflow, kept


def test_cbrt_within_one_ulp(gpu):
    """ufuncs.rs:367 x.cbrt(): every binade, subnormals, negatives; special values exact."""
    r = np.random.default_rng(101)
    x = np.concatenate([r.choice([-1.0, 1.0], 60000) * np.exp2(r.uniform(-1074, 1023, 60000)),
                        r.uniform(-10, 10, 20000), np.arange(1.0, 2000.0) ** 3])
    got = K.cbrt(x)
    err = ulp_err(got, np.cbrt(x.astype(LD)))
    print(f"cbrt: max {err.max():.3f} ulp")
    assert err.max() <= 1.0
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 8.0, -27.0, 5e-324])
    assert_same(K.cbrt(sp)[:7], np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 2.0, -3.0]))
    x32 = r.uniform(-1e6, 1e6, 20000).astype(np.float32)
    assert ulp_err(K.cbrt(x32), np.cbrt(x32.astype(LD)), np.float32).max() <= 0.5 + 1e-9


def test_reference_named_arithmetic(gpu):
    """ufuncs.rs:182-303 by name, one IEEE rounding each, and the flat reductions."""
    r = np.random.default_rng(102)
    a, b = r.normal(size=10001), r.normal(size=10001)
    for name, ref in (("add", np.add), ("subtract", np.subtract), ("multiply", np.multiply), ("divide", np.divide)):
        np.testing.assert_array_equal(getattr(K, name)(a, b), ref(a, b))
        np.testing.assert_array_equal(getattr(K, name + "_scalar")(a, 0.37), ref(a, 0.37))
    np.testing.assert_array_equal(K.maximum(a, b), np.where(a > b, a, b))
    np.testing.assert_array_equal(K.abs(a), np.abs(a))
    assert K.sum(a) == pytest.approx(math.fsum(a), abs=1e-12)
    assert K.mean(a) == pytest.approx(math.fsum(a) / a.size, abs=1e-15)
    assert K.var(a) == pytest.approx(np.var(a), rel=1e-13)
