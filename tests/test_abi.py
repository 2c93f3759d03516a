"""The C-ABI library loads and exports every symbol include/numrs_amd.h declares (CPU only:
no compute call is made, so this runs without a GPU)."""
import ctypes
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    from numrs_amd import _lib

    lib = _lib.load()
    declared = _lib.declared_symbols()
    assert len(declared) >= 60
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"
    # and the binding table covers everything that returns a status
    unbound = [s for s in declared if s not in _lib._SIGS and s not in _lib._SIZE_FNS and s not in ("nrk_status_string", "nrk_last_error", "nrk_version")]
    assert not unbound, f"declared but not bound in numrs_amd/_lib.py: {unbound}"


def test_exports_are_c_symbols_not_mangled():
    so = ROOT / "numrs_amd" / "libnumrs_amd.so"
    out = subprocess.run(["nm", "-D", "--defined-only", str(so)], capture_output=True, text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    from numrs_amd import _lib

    for s in _lib.declared_symbols():
        assert s in names


def test_status_strings_and_version():
    from numrs_amd import _lib

    lib = _lib.load()
    assert lib.nrk_status_string(0) == b"ok"
    assert lib.nrk_status_string(1) == b"invalid argument"
    assert lib.nrk_version() == b"0.1.0"


def test_header_compiles_as_c_and_cxx(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "numrs_amd.h"\nint main(void){ nrk_scalar_op op = {NRK_SOP_MUL, 0, 1.5}; (void)op; return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{ROOT / 'include'}", "-c", str(src), "-o", str(tmp_path / "t.o")], check=True)
    subprocess.run(["g++", "-x", "c++", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'include'}", "-c", str(src), "-o", str(tmp_path / "t2.o")], check=True)


def test_product_package_has_no_oracle_or_cpu_fallback():
    """The product path never imports the oracle and never computes on the CPU."""
    pkg = ROOT / "numrs_amd"
    for f in pkg.rglob("*.py"):
        text = f.read_text()
        assert "import oracle" not in text and "from oracle" not in text, f
    for f in (pkg / "csrc").iterdir():
        assert "oracle" not in f.read_text().lower(), f


def test_fused_chain_refuses_elementary_functions():
    """The fused map -> reduce kernels are built without the elementary functions (exp .. tanh):
    such a chain is refused with NRK_ERR_UNSUPPORTED before any device work (no GPU needed)."""
    import ctypes as C

    from numrs_amd import kernels as K
    from numrs_amd._lib import NrkError, ReduceOutputs, call

    ops, n = K.chain(("mul", 2.0), ("exp",))
    ro, fo = ReduceOutputs(), ReduceOutputs()
    with pytest.raises(NrkError) as e:
        call("nrk_dev_map_reduce_f64", C.cast(ops, C.c_void_p), n, C.c_void_p(16), 4, 4, 0, C.byref(ro), C.byref(fo), None)
    assert "elementwise map only" in str(e.value)
