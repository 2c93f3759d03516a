"""Generic-T elementwise on the GPU: the integer zip / map (kernels/elementwise.rs:217-258 for
T = i64, i32, u64, u32 with arithmetic.rs's and bitwise_ops.rs's operators), the broadcasting
zip (operations.rs:370-440 add/subtract/multiply/divide_broadcast) and the strided zip that reads
non-contiguous views in place and takes shift amounts of another integer type (binary_serial's
heterogeneous T, U: bitwise_ops.rs:190-243).

Bars: bit-exact. Integers against the oracle's restatement of Rust's release arithmetic (wrapping
+ - *, truncating / and %, masked shifts), pinned in test_oracle.py; floats against numpy (IEEE
ops, one rounding each, the same bits)."""
import numpy as np
import pytest

from numrs_amd import kernels as K
from numrs_amd._lib import NrkError

pytestmark = pytest.mark.gpu
INTS = [np.int64, np.int32, np.uint64, np.uint32]
ZIP_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "rem": 4, "maximum": 5, "minimum": 6, "bitand": 13, "bitor": 14,
           "bitxor": 15, "shl": 16, "shr": 17}


def _full_range(r, dt, n):
    info = np.iinfo(dt)
    return r.integers(info.min, info.max, n, dtype=dt, endpoint=True)


@pytest.mark.parametrize("dt", INTS)
def test_integer_zip_bit_exact(gpu, oracle, dt):
    r = np.random.default_rng(np.dtype(dt).itemsize + (0 if np.issubdtype(dt, np.signedinteger) else 100))
    n = 100_003
    a, b = _full_range(r, dt, n), _full_range(r, dt, n)
    b[b == 0] = 7
    if np.issubdtype(dt, np.signedinteger):
        b[(b == -1) & (a == np.iinfo(dt).min)] = 3
    a[:4] = [np.iinfo(dt).max, np.iinfo(dt).min, 0, 1]
    for name, op in ZIP_OPS.items():
        bb = (b % 200).astype(dt) if name in ("shl", "shr") else b
        np.testing.assert_array_equal(K.binary_dispatch(a, bb, name), oracle.ref_zip_int(op, a, bb), err_msg=name)


def test_integer_zip_reference_examples_and_errors(gpu):
    """bitwise_ops.rs doc examples; a zero divisor or MIN / -1 raises (Rust panics) and leaves
    nothing written."""
    a, b = np.array([13, 17, 21], np.int64), np.array([9, 7, 15], np.int64)
    assert list(K.binary_dispatch(a, b, "bitand")) == [9, 1, 5]
    assert list(K.binary_dispatch(np.array([5, 10, 15], np.int32), np.array([1, 2, 3], np.int32), "shl")) == [10, 40, 120]
    assert list(K.binary_dispatch(np.array([40, 80, 120], np.int32), np.array([1, 2, 3], np.int32), "shr")) == [20, 20, 15]
    with pytest.raises(NrkError):
        K.binary_dispatch(np.array([1, 2], np.int64), np.array([1, 0], np.int64), "div")
    with pytest.raises(NrkError):
        K.binary_dispatch(np.array([np.iinfo(np.int32).min], np.int32), np.array([-1], np.int32), "div")
    with pytest.raises(NrkError):
        K.binary_dispatch(np.array([1], np.uint32), np.array([0], np.uint32), "rem")


@pytest.mark.parametrize("dt", INTS)
def test_integer_map_bit_exact(gpu, oracle, dt):
    r = np.random.default_rng(7 + np.dtype(dt).itemsize)
    x = _full_range(r, dt, 50_001)
    signed = np.issubdtype(dt, np.signedinteger)
    chains = [[("mul", 3), ("add", 12345)], [("sub", 7), ("square",)], [("clamp_lo", 100), ("clamp_hi", 1 << 20)],
              [("div", 7)], [("rsub", 5)]]
    if signed:
        chains += [[("neg",), ("abs",)], [("sign",)], [("mul", -3), ("div", -5)]]
    from numrs_amd.kernels import SOP
    for steps in chains:
        ops = [SOP[s[0]] for s in steps]
        sc = [int(s[1]) if len(s) > 1 else 0 for s in steps]
        np.testing.assert_array_equal(K.unary_dispatch_int(x, steps), oracle.ref_map_int(ops, sc, x), err_msg=str(steps))
    if not signed:
        with pytest.raises(NrkError):
            K.unary_dispatch_int(x, [("neg",)])


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int64, np.int32])
def test_broadcast_ops_match_numpy(gpu, oracle, dt):
    """add/subtract/multiply/divide_broadcast over numpy's broadcast rules (the reference's
    broadcast_shape), with no broadcast_to copies: floats bit-identical to numpy, integers to the
    restatement (numpy's // floors, Rust's / truncates)."""
    r = np.random.default_rng(3)
    shapes = [((3, 1, 5), (4, 1)), ((7,), ()), ((2, 3, 4, 5), (3, 1, 5)), ((1, 1000), (257, 1)), ((6,), (6,)),
              ((5, 0, 3), (1, 3))]
    for sa, sb in shapes:
        if np.issubdtype(dt, np.floating):
            a, b = r.standard_normal(sa).astype(dt), r.standard_normal(sb).astype(dt)
        else:
            a, b = r.integers(-1000, 1000, sa).astype(dt), r.integers(1, 1000, sb).astype(dt)
        for name, f in (("add", np.add), ("sub", np.subtract), ("mul", np.multiply), ("div", np.true_divide)):
            got = getattr(K, {"add": "add_broadcast", "sub": "subtract_broadcast", "mul": "multiply_broadcast",
                              "div": "divide_broadcast"}[name])(a, b)
            av, bv = np.broadcast_arrays(a, b)
            if np.issubdtype(dt, np.floating):
                want = f(av, bv).astype(dt)
            else:
                want = oracle.ref_zip_int(ZIP_OPS[name], np.ascontiguousarray(av), np.ascontiguousarray(bv)).reshape(av.shape)
            assert got.shape == want.shape, (sa, sb)
            np.testing.assert_array_equal(got, want, err_msg=f"{name} {sa} {sb}")
    with pytest.raises(ValueError):
        K.add_broadcast(np.ones((3, 2), dt), np.ones((4,), dt))


def test_strided_views_and_mixed_shift_types(gpu, oracle):
    """Views read in place (transposed, reversed, sliced): the same bits as on contiguous copies.
    Shift amounts of another integer type (i64 << u32, i32 >> i64): the amount's value, masked to
    the width of the shifted type, as Rust's release `<<` / `>>`."""
    r = np.random.default_rng(11)
    m = r.standard_normal((300, 200))
    t = m.T
    rev = m[::-1, ::3]
    got = K.zip_strided("add", t, t[:, ::-1])
    np.testing.assert_array_equal(got, np.ascontiguousarray(t) + np.ascontiguousarray(t[:, ::-1]))
    np.testing.assert_array_equal(K.zip_strided("mul", rev, rev[::-1]), rev * rev[::-1])
    col = m[:, :1]
    np.testing.assert_array_equal(K.zip_strided("sub", m, col), m - col)
    a = r.integers(-(1 << 40), 1 << 40, 5000, dtype=np.int64)
    s32 = r.integers(0, 1 << 32, 5000, dtype=np.uint64).astype(np.uint32)
    want = oracle.ref_zip_int(16, a, (s32.astype(np.uint64) & 63).astype(np.int64))
    np.testing.assert_array_equal(K.left_shift(a, s32), want)
    b32 = r.integers(-(1 << 31), 1 << 31, 5000, dtype=np.int64).astype(np.int32)
    s64 = r.integers(-300, 300, 5000, dtype=np.int64)
    want = oracle.ref_zip_int(17, b32, (s64.astype(np.uint64) & np.uint64(31)).astype(np.int32))
    np.testing.assert_array_equal(K.right_shift(b32, s64), want)
    bt = r.integers(-1000, 1000, (40, 30), dtype=np.int32).T  # a transposed integer view
    np.testing.assert_array_equal(K.zip_strided("bitxor", bt, bt[::-1]), np.ascontiguousarray(bt) ^ np.ascontiguousarray(bt[::-1]))
