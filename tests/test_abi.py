"""CPU-side checks of the C ABI library: it loads and exports exactly the
entry points include/solaris_amd.h declares (no compute call needs a GPU)."""
import ctypes
import os
import re

from solaris_amd import runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "solaris_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sol_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = runtime.library()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_covers_header():
    assert sorted(runtime.exported_symbols()) == declared_symbols()


def test_version_and_error_without_device():
    lib = runtime.library()
    assert b"gfx950" in lib.sol_version()
    # sol_create with total == 0 fails before touching a device
    from solaris_amd import _abi
    nb = _abi.NBodies()
    h = ctypes.c_void_p()
    r = lib.sol_create(ctypes.byref(h), ctypes.byref(nb), 0, 0)
    assert r == 6  # SOL_ERR_ALLOCATION, BodyData::Allocate with total == 0
    assert b"allocation" in lib.sol_last_error(None)
