"""Build recipe for the oracle shared library (test infrastructure only).

    python -m oracle.build              # oracle/liboracle.so
    python -m oracle.build --sanitize   # oracle/_san/sanitize_check (ASan + UBSan, host only)

Compiles oracle/oracle.c with gcc into oracle/liboracle.so. -ffp-contract=off keeps every
`a * b + c` of the restatement as two roundings (Rust never contracts); -O3 -march=x86-64-v3 (AVX2; the library also runs on the GPU host) for the
CPU-baseline timings. The reference itself (Rust) cannot be built here (no rustc/cargo, no
vendored scirs2-core), so there is no oracle/_ref build.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "oracle.c"
LIB = HERE / "liboracle.so"
CFLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-std=gnu11",
          "-Wall", "-Wno-unused-variable", "-pthread"]


def build(verbose: bool = True) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", *CFLAGS, str(SRC), "-o", str(tmp), "-lm"]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{proc.stdout}")
    if proc.stdout.strip():
        print(proc.stdout, file=sys.stderr)
    tmp.replace(LIB)
    if verbose:
        print(f"built {LIB.name}", file=sys.stderr)
    return LIB


SAN_DIR = HERE / "_san"
SAN_BIN = SAN_DIR / "sanitize_check"
# oracle.c with a driver that calls every function on exactly-sized buffers (test infrastructure)
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-ffp-contract=off", "-std=gnu11", "-Wall", "-Wno-unused-variable", "-pthread"]


def build_sanitized(verbose: bool = True) -> Path:
    """AddressSanitizer + UndefinedBehaviorSanitizer build of the oracle and its driver
    (oracle/sanitize_check.c) into oracle/_san/sanitize_check: host code only, gcc."""
    SAN_DIR.mkdir(exist_ok=True)
    cmd = ["gcc", *SAN_FLAGS, str(SRC), str(HERE / "sanitize_check.c"), "-o", str(SAN_BIN), "-lm"]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"sanitized oracle build failed: {' '.join(cmd)}\n{proc.stdout}")
    if verbose:
        print(f"built {SAN_BIN.relative_to(HERE.parent)}", file=sys.stderr)
    return SAN_BIN


if __name__ == "__main__":
    if "--sanitize" in sys.argv[1:]:
        build_sanitized()
    else:
        build()
