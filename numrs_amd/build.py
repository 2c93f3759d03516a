"""Build recipe for libnumrs_amd.so (gfx950).

    python -m numrs_amd.build [--changed]

The library is compiled with hipcc, one object per source, in-tree: the shared object lands at
numrs_amd/libnumrs_amd.so so that it travels with the repository snapshot. Every device pass
disables packed-FP32 instruction selection and FP contraction (elementwise parity relies on two
roundings for `x * a + b`).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libnumrs_amd.so"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; that line is noise.
    noise = "is not a recognized feature for this target"
    out = "\n".join(l for l in proc.stdout.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _includes(path: Path, seen: set) -> None:
    """The project headers `path` includes, recursively (quoted includes found in csrc/ or include/)."""
    for line in path.read_text().splitlines():
        line = line.strip()
        if not line.startswith('#include "'):
            continue
        name = line.split('"')[1]
        for d in (CSRC, INCLUDE):
            h = d / name
            if h.exists() and h not in seen:
                seen.add(h)
                _includes(h, seen)


def _stale(src: Path, obj: Path) -> bool:
    """True if obj is missing or older than src, this recipe, or any project header src includes."""
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    deps: set = set()
    _includes(src, deps)
    return any(d.stat().st_mtime > t for d in [src, Path(__file__), *deps])


def build_lib(verbose: bool = True, changed_only: bool = False) -> Path:
    """Compile every csrc/*.hip and link libnumrs_amd.so. Always recompiles everything unless
    changed_only, which recompiles only objects older than their source or any header."""
    if not Path(HIPCC).exists() and shutil.which(HIPCC) is None:
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    all_sources = sorted(CSRC.glob("*.hip"))
    objs_all = [BUILD / (s.stem + ".o") for s in all_sources]
    pairs = [(s, o) for s, o in zip(all_sources, objs_all) if not changed_only or _stale(s, o)]
    sources = [s for s, _ in pairs]
    objs = objs_all

    def compile_one(src_obj):
        src, obj = src_obj
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        return obj

    workers = min(max(1, len(sources)), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(compile_one, pairs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB.relative_to(ROOT)} ({len(sources)} of {len(all_sources)} sources compiled)", file=sys.stderr)
    return LIB


def main(argv: list[str]) -> int:
    build_lib(changed_only="--changed" in argv)
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv[1:]))
