"""Writes the golden fixtures under tests/golden/ (run once, in the build container).

numpy_reference.json  -- the `statistical_operations` and `arithmetic_operations` groups of the
                         reference's NumPy-generated data file
                         tests/py/array_operations_reference_data.json (read as JSON data only;
                         generated upstream by tests/py/array_operations_reference_tests.py,
                         seed 42).
known_answers.json    -- known answers transcribed from the reference's own tests and docs, each
                         with the file:line it comes from.
"""
from __future__ import annotations

import json
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF_JSON = Path("/root/reference/tests/py/array_operations_reference_data.json")


def numpy_subset() -> dict:
    data = json.loads(REF_JSON.read_text())
    return {
        "source": "reference tests/py/array_operations_reference_data.json (NumPy, seed 42)",
        "statistical_operations": data["statistical_operations"],
        "arithmetic_operations": data["arithmetic_operations"],
    }


def known_answers() -> dict:
    return {
        "determinism_pins": {
            "source": "tests/test_reduction_determinism.rs:7-14,126-197",
            "data": "x_i = i * 0.5 - 12345 for i in 0..50000 (f64 and f32)",
            "n": 50000,
            "sum_f64_bits": 4710066088337997824,
            "mean_f64_bits": 4639648798144987136,
            "var_ddof0_f64_bits": 4722259316520779776,
            "std_ddof0_f64_bits": 4664657056377925821,
            "sum_f32_bits": 1256988984,
        },
        "var_small": {
            "source": "src/kernels/reduce.rs:577-583 (np.var([1,2,3,4,5]) = 2.0, ddof=1 -> 2.5)",
            "x": [1.0, 2.0, 3.0, 4.0, 5.0], "ddof0": 2.0, "ddof1": 2.5,
        },
        "min_max_known": {
            "source": "src/kernels/reduce.rs:822-826,995-999",
            "x": [5.0, -3.0, 8.0, 0.5, -10.0, 2.0], "min": -10.0, "max": 8.0,
        },
        "min_max_nan": {
            "source": "src/kernels/reduce.rs:843-860 (any NaN -> NaN)",
            "cases": [[1.0, "nan", 3.0, -2.0, "nan"], ["nan", 1.0, 2.0], [1.0, 2.0, "nan"], ["nan"], ["nan", "nan"]],
        },
        "min_max_inf": {
            "source": "src/kernels/reduce.rs:862-867",
            "cases": [{"x": ["inf", 1.0], "min": 1.0, "max": "inf"}, {"x": ["-inf", 1.0], "min": "-inf", "max": 1.0}],
        },
        "lane_aligned_nan": {
            "source": "src/kernels/reduce.rs:879-911 (64 x 1.0, x[0]=5, x[10]=NaN; and NaN at 0,8,32,63)",
            "len": 64, "fill": 1.0, "first": 5.0, "nan_positions": [10, 0, 8, 32, 63],
        },
        "argmax_doc": {
            "source": "src/math/statistics/mod.rs:65-77",
            "x": [1.0, 3.0, 2.0, 4.0, 5.0, 1.0], "shape": [2, 3], "axis1": [1, 1], "flat": 4,
        },
        "argmin_doc": {
            "source": "src/math/statistics/mod.rs:204-216",
            "x": [5.0, 3.0, 2.0, 4.0, 1.0, 6.0], "shape": [2, 3], "axis1": [2, 1], "flat": 4,
        },
        "matmul_3x3": {
            "source": "tests/test_linalg_reference.rs:163-196",
            "a": [1, 2, 3, 4, 5, 6, 7, 8, 9], "c": [30, 36, 42, 66, 81, 96, 102, 126, 150], "v": [1, 2, 3],
            "av": [14, 32, 50],
        },
        "rem_known": {
            "source": "src/array/arithmetic.rs:509-512 (10 % 3, 21 % 4, 33 % 5 on i64; same values as floats)",
            "a": [10.0, 21.0, 33.0], "b": [3.0, 4.0, 5.0], "out": [1.0, 1.0, 3.0],
        },
    }


def main() -> None:
    (HERE / "numpy_reference.json").write_text(json.dumps(numpy_subset(), indent=1))
    (HERE / "known_answers.json").write_text(json.dumps(known_answers(), indent=1))
    print("wrote", HERE / "numpy_reference.json", HERE / "known_answers.json")


if __name__ == "__main__":
    main()
