"""Fused elementwise -> reduction (one HBM pass) through nrk_dev_map_reduce_{f32,f64}.

Checked against the restatement of the reference's unfused sequence: y = (x * 1.5) + (-0.25)
as two rounded maps, then the reductions over y. Index and extremum results are bit-exact; sums
and variances are within the stated tolerance of the double-double statistics of y, and the row
results are bit-identical to the unfused GPU path (map, then reduce)."""
import ctypes as C

import numpy as np
import pytest

from numrs_amd import kernels as K
from numrs_amd import rng
from numrs_amd._lib import ReduceOutputs, call
from numrs_amd.device import DeviceArray

pytestmark = pytest.mark.gpu
EPS = {np.float64: 2.0 ** -53, np.float32: 2.0 ** -24}
FIELDS = ("sum", "mean", "var", "min", "max", "argmin", "argmax")


def run_fused(x: np.ndarray, steps, ddof=0, want_rows=True, want_full=True, offset_elems=0, plain=False, fields=FIELDS):
    rows, cols = x.shape
    dt = x.dtype
    dx = DeviceArray((x.size + offset_elems,), dt)
    buf = np.zeros(x.size + offset_elems, dt)
    buf[offset_elems:] = x.reshape(-1)
    dx.fill_from(buf)
    xptr = dx.ptr + offset_elems * dt.itemsize
    outs = {}
    ro, fo = ReduceOutputs(), ReduceOutputs()
    for f in fields:
        odt = np.uint64 if f.startswith("arg") else dt
        if want_rows:
            outs["row_" + f] = DeviceArray((rows,), odt)
            setattr(ro, f, outs["row_" + f].ptr)
        if want_full:
            outs["full_" + f] = DeviceArray((1,), odt)
            setattr(fo, f, outs["full_" + f].ptr)
    ops, n_ops = K.chain(*steps)
    fn = ("nrk_dev_map_reduce_plain_" if plain else "nrk_dev_map_reduce_") + ("f64" if dt == np.float64 else "f32")
    call(fn, C.cast(ops, C.c_void_p), C.c_int32(n_ops), C.c_void_p(xptr), C.c_size_t(rows), C.c_size_t(cols),
         C.c_size_t(ddof), C.byref(ro) if want_rows else None, C.byref(fo) if want_full else None, None)
    call("nrk_device_synchronize")
    return {k: v.download() for k, v in outs.items()}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 1), (4, 7), (16, 8192), (9, 8193), (3, 40_000), (257, 100),
                                   (1024, 1024), (2048, 3072), (1100, 2048), (4096, 1024)])  # the last four: the lean kernel
def test_fused_vs_oracle(gpu, oracle, dtype, shape):
    rows, cols = shape
    x = rng.normal(rng.BASE_SEED + 1100 + rows * 7 + cols, rows * cols, dtype).reshape(rows, cols)
    steps = [("mul", 1.5), ("add", -0.25)]
    got = run_fused(x, steps, ddof=0)
    y = oracle.ref_map([2, 0], [1.5, -0.25], x).reshape(rows, cols)
    for op in ("max", "min", "argmax", "argmin"):
        np.testing.assert_array_equal(got["row_" + op], oracle.ref_axis(K.AXIS[op], y, rows, cols, 1), err_msg=op)
    assert got["full_argmax"][0] == oracle.ref("argmax", y)
    assert got["full_argmin"][0] == oracle.ref("argmin", y)
    assert got["full_max"][0] == oracle.ref("max", y) and got["full_min"][0] == oracle.ref("min", y)
    ex_rows = oracle.dd_row_stats(y, 0)
    s_abs = np.abs(y.astype(np.float64)).sum(axis=1)
    e = EPS[dtype]
    assert np.all(np.abs(got["row_sum"] - (ex_rows[:, 0] + ex_rows[:, 1])) <= 8 * e * s_abs + e * np.abs(ex_rows[:, 0]))
    if cols > 1:
        assert np.all(np.abs(got["row_var"] - ex_rows[:, 3]) <= 64 * e * ex_rows[:, 3])
    ex = oracle.dd_stats(y, 0)
    assert abs(got["full_sum"][0] - ex["sum"]) <= 8 * e * float(s_abs.sum()) + e * abs(ex["sum"])
    if rows * cols > 1:
        assert abs(got["full_var"][0] - ex["var"]) <= 64 * e * ex["var"]
    assert abs(got["full_mean"][0] - ex["mean"]) <= 8 * e * float(s_abs.sum()) / (rows * cols) + 2 * e * abs(ex["mean"])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fused_rows_bit_identical_to_unfused(gpu, dtype):
    rows, cols = 64, 8192
    x = rng.normal(rng.BASE_SEED + 1200, rows * cols, dtype).reshape(rows, cols)
    got = run_fused(x, [("mul", 1.5), ("add", -0.25)], want_full=False)
    y = K.unary_dispatch(x, [("mul", 1.5), ("add", -0.25)])
    for op in ("sum", "mean", "var", "max", "argmax"):
        np.testing.assert_array_equal(got["row_" + op], K.reduce_axis(op, y, rows, cols, 1), err_msg=op)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fused_misaligned_pointer_same_bits(gpu, dtype):
    """A device pointer off 16-byte alignment takes the scalar-load path; the summation tree is the
    same, so every output is bit-identical to the aligned run."""
    x = rng.normal(rng.BASE_SEED + 1300, 12 * 5000, dtype).reshape(12, 5000)
    a = run_fused(x, [("mul", 1.5), ("add", -0.25)])
    b = run_fused(x, [("mul", 1.5), ("add", -0.25)], offset_elems=1)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_fused_nan_rows(gpu, oracle):
    x = rng.normal(rng.BASE_SEED + 1400, 6 * 9000, np.float32).reshape(6, 9000)
    x[1, 0] = np.nan      # argmax pinned to 0 for row 1
    x[2, 8500] = np.nan   # interior NaN: skipped by argmax, poisons max
    got = run_fused(x, [])
    assert got["row_argmax"][1] == 0 and np.isnan(got["row_max"][1]) and np.isnan(got["row_max"][2])
    np.testing.assert_array_equal(got["row_argmax"], oracle.ref_axis(6, x, 6, 9000, 1))
    assert got["full_argmax"][0] == oracle.ref("argmax", x) and np.isnan(got["full_max"][0])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1024, 2048), (1, 1 << 21)])
def test_lean_path_alignment_nan_and_ties(gpu, oracle, dtype, shape):
    """The lean kernel (whole segments, >= 1024 chunks): a pointer off 16-byte alignment reads the
    same elements one by one, so every output is bit-identical; NaN at a row's first element pins
    argmax to 0, an interior NaN is skipped by argmax and poisons max; equal maxima go to the first
    index, the sign of a zero maximum is the first zero's (row and whole array, against the
    oracle's sequential folds)."""
    rows, cols = shape
    x = rng.normal(rng.BASE_SEED + 1500 + rows, rows * cols, dtype).reshape(rows, cols)
    want_rows = rows > 1
    a = run_fused(x, [("mul", 1.5), ("add", -0.25)], want_rows=want_rows)
    b = run_fused(x, [("mul", 1.5), ("add", -0.25)], want_rows=want_rows, offset_elems=1)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    z = x.copy()
    flat = z.reshape(-1)
    if want_rows:
        z[3, 0] = np.nan
        z[5, 777] = np.nan
        z[7, :] = -0.0
        z[7, 1000] = 0.0
        z[9, 100] = z[9, 1500] = 50.0
    else:
        flat[123_457] = np.nan
        flat[:] = np.where(np.arange(flat.size) % 4096 == 17, 80.0, flat)
    got = run_fused(z, [], want_rows=want_rows)
    assert got["full_argmax"][0] == oracle.ref("argmax", z)
    assert np.isnan(got["full_max"][0]) == np.isnan(oracle.ref("max", z))
    if want_rows:
        np.testing.assert_array_equal(got["row_argmax"], oracle.ref_axis(6, z, rows, cols, 1))
        np.testing.assert_array_equal(got["row_argmin"], oracle.ref_axis(5, z, rows, cols, 1))
        rm, om = got["row_max"], oracle.ref_axis(4, z, rows, cols, 1)
        np.testing.assert_array_equal(np.isnan(rm), np.isnan(om))
        ok = ~np.isnan(om)
        np.testing.assert_array_equal(rm[ok], om[ok])
        np.testing.assert_array_equal(np.signbit(rm[ok]), np.signbit(om[ok]))
        assert got["row_argmax"][3] == 0 and got["row_argmax"][9] == 100


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, (1 << 21) + 777), (1, 1100 * 1024 + 5), (1, (1 << 24) + 3 * 1024 + 17),
                                   (4097, 4099)])
def test_lean_whole_array_with_tail(gpu, oracle, dtype, shape):
    """Whole-array reductions of any length: whole chunks on the lean kernel plus one workgroup for
    the elements after the last whole chunk. Against the oracle (extrema and indices bit-exact,
    sums and variances within the reduction bars); a pointer off 16-byte alignment and a repeated
    call give the same bits; a NaN or a tied maximum placed in the tail follows the sequential
    rules (NaN poisons max and is skipped by argmax; the first of equal maxima wins)."""
    rows, cols = shape
    n = rows * cols
    x = rng.normal(rng.BASE_SEED + 1600 + cols, n, dtype).reshape(rows, cols)
    steps = [("mul", 1.5), ("add", -0.25)]
    a = run_fused(x, steps, want_rows=False)
    b = run_fused(x, steps, want_rows=False, offset_elems=1)
    c = run_fused(x, steps, want_rows=False)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        np.testing.assert_array_equal(a[k], c[k], err_msg=k)
    y = oracle.ref_map([2, 0], [1.5, -0.25], x).reshape(rows, cols)
    for op in ("max", "min", "argmax", "argmin"):
        assert a["full_" + op][0] == oracle.ref(op, y), op
    ex = oracle.dd_stats(y, 0)
    e = EPS[dtype]
    s_abs = float(np.abs(y.astype(np.float64)).sum())
    assert abs(a["full_sum"][0] - ex["sum"]) <= 8 * e * s_abs + e * abs(ex["sum"])
    assert abs(a["full_var"][0] - ex["var"]) <= 64 * e * ex["var"]
    assert abs(a["full_mean"][0] - ex["mean"]) <= 8 * e * s_abs / n + 2 * e * abs(ex["mean"])
    z = x.copy()
    flat = z.reshape(-1)
    big = dtype(1000.0)
    flat[n // 3] = big       # ties: the first index wins, the tied copy sits in the tail
    flat[n - 2] = big
    flat[n - 5] = -big       # the minimum in the tail
    got = run_fused(z, [], want_rows=False)
    assert got["full_argmax"][0] == oracle.ref("argmax", z) == n // 3
    assert got["full_argmin"][0] == oracle.ref("argmin", z) == n - 5
    flat[n - 1] = np.nan     # a NaN in the tail: max poisoned, argmax unchanged
    got = run_fused(z, [], want_rows=False)
    assert np.isnan(got["full_max"][0]) and got["full_argmax"][0] == oracle.ref("argmax", z) == n // 3


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1024, 2048), (16, 8192), (9, 8193), (3000, 100), (1, (1 << 21) + 777)])
def test_infinite_elements_give_binary64_sums(gpu, dtype, shape):
    """An infinite element makes its row's and the whole array's sum and mean infinite (not the NaN
    of a TwoSum error term); +inf and -inf together make NaN; the variance is NaN; on every path
    (row chunks, split rows, misaligned rows, short rows, a whole array with a tail)."""
    rows, cols = shape
    x = rng.normal(rng.BASE_SEED + 2100 + cols, rows * cols, dtype).reshape(rows, cols)
    x[0, 5] = -np.inf
    if rows > 2:
        x[2, 1] = np.inf
        x[2, cols - 1] = -np.inf
    want_rows = rows > 1
    got = run_fused(x, [], want_rows=want_rows)
    y = x.astype(np.float64)
    with np.errstate(invalid="ignore"):
        if want_rows:
            rs = y.sum(axis=1)
            np.testing.assert_array_equal(np.isnan(got["row_sum"]), np.isnan(rs))
            np.testing.assert_array_equal(np.isinf(got["row_sum"]), np.isinf(rs))
            assert got["row_sum"][0] == -np.inf and got["row_mean"][0] == -np.inf and np.isnan(got["row_var"][0])
            if rows > 2:
                assert np.isnan(got["row_sum"][2])
        fs = y.sum()
    assert np.isnan(got["full_sum"][0]) == np.isnan(fs)
    if not np.isnan(fs):
        assert got["full_sum"][0] == -np.inf and got["full_mean"][0] == -np.inf
    assert np.isnan(got["full_var"][0])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 5, 1000, 1023, 1024 + 5, 3000, 1_000_003])
def test_whole_array_nan_first_and_interior(gpu, oracle, dtype, n):
    """Whole arrays of every size class (one workgroup, a few chunks, many): a NaN at element 0
    pins argmax and argmin to 0; an interior NaN is skipped by them; max and min are NaN either
    way (against the oracle's sequential folds)."""
    x = rng.normal(rng.BASE_SEED + 2200 + n, n, dtype).reshape(1, n)
    for pos in sorted({0, n // 2, n - 1}):
        z = x.copy()
        z[0, pos] = np.nan
        got = run_fused(z, [], want_rows=False)
        assert np.isnan(got["full_max"][0]) and np.isnan(got["full_min"][0]), pos
        assert got["full_argmax"][0] == oracle.ref("argmax", z), pos
        assert got["full_argmin"][0] == oracle.ref("argmin", z), pos
        if pos == 0:
            assert got["full_argmax"][0] == 0 and got["full_argmin"][0] == 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(4096, 3001), (4100, 2500), (5000, 1030)])
def test_general_rows_with_whole_array(gpu, oracle, dtype, shape):
    """Long rows (not a multiple of the segment) from 2048 rows: a chunk is a row (row results
    straight from the stream) and the whole-array results come from the record tree over the row
    records. Against the oracle; a pointer off 16-byte alignment and a repeated call give the same
    bits, and the row results equal the rows-only call's."""
    rows, cols = shape
    x = rng.normal(rng.BASE_SEED + 2300 + cols, rows * cols, dtype).reshape(rows, cols)
    steps = [("mul", 1.5), ("add", -0.25)]
    a = run_fused(x, steps)
    b = run_fused(x, steps, offset_elems=1)
    c = run_fused(x, steps)
    d = run_fused(x, steps, want_full=False)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        np.testing.assert_array_equal(a[k], c[k], err_msg=k)
        if k.startswith("row_"):
            np.testing.assert_array_equal(a[k], d[k], err_msg=k)
    y = oracle.ref_map([2, 0], [1.5, -0.25], x).reshape(rows, cols)
    for op in ("max", "min", "argmax", "argmin"):
        assert a["full_" + op][0] == oracle.ref(op, y), op
        np.testing.assert_array_equal(a["row_" + op], oracle.ref_axis(K.AXIS[op], y, rows, cols, 1), err_msg=op)
    ex = oracle.dd_stats(y, 0)
    e = EPS[dtype]
    s_abs = float(np.abs(y.astype(np.float64)).sum())
    assert abs(a["full_sum"][0] - ex["sum"]) <= 8 * e * s_abs + e * abs(ex["sum"])
    assert abs(a["full_var"][0] - ex["var"]) <= 64 * e * ex["var"]
    z = x.copy()
    z[0, 0] = np.nan  # the first element: whole-array argmax / argmin pinned to 0
    got = run_fused(z, [])
    assert got["full_argmax"][0] == 0 and got["full_argmin"][0] == 0 and np.isnan(got["full_max"][0])
