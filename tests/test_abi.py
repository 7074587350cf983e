"""CPU checks of the C ABI: the library loads and exports every function that
include/mnpbem_hip.h declares, the ctypes table covers them, and the host-only
entry points behave (no GPU calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mnpbem_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mnp_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from mnpbem_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from mnpbem_amd import build
        build.build()
    return _lib.lib()


def test_header_declares_entry_points():
    names = declared()
    assert "mnp_greenret_create" in names and "mnp_bemret_solve" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol(lib):
    from mnpbem_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (mnp_\w+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    for name in declared():
        assert hasattr(lib, name)


def test_ctypes_table_matches_header():
    from mnpbem_amd import _lib
    assert sorted(_lib.PROTOTYPES) == declared()


def test_host_only_entry_points(lib):
    from mnpbem_amd import _lib
    op = _lib.GreenOptions()
    lib.mnp_green_options_default(C.byref(op))
    assert (op.order, op.rel_cutoff, op.abs_cutoff, op.npol_rho, op.npol_phi, op.closed) == \
        (5, 3.0, 0.0, 7, 5, 1)
    assert lib.mnp_version() >= 100
    assert isinstance(lib.mnp_last_error(), bytes)
    # argument validation happens before any device work
    st = lib.mnp_greenret_create(None, 0, None, 0, None, None)
    assert st == 1 and b"NULL" in lib.mnp_last_error()
    assert lib.mnp_bemret_create(None, 0, None) == 1
    assert lib.mnp_gemm(7, 1, 1, 1, _lib.MnpComplex(1, 0), None, 1, None, 1,
                        _lib.MnpComplex(0, 0), None, 1) == 1


def test_product_path_has_no_oracle_imports():
    """The shipped package never imports the oracle (it is the checker)."""
    pkg = os.path.join(ROOT, "mnpbem_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.sub(r"#.*|//.*", "", src).replace("oracles", ""), f
