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


def test_device_code_has_no_packed_fp32_instructions():
    """Every HIP source is compiled with packed FP32 disabled; the gfx950
    assembly must contain no v_pk_{add,mul,fma}_f32."""
    import re
    from solaris_amd import build
    for src in build.HIP_SOURCES:
        asm = build.device_assembly(src)
        assert "s_endpgm" in asm or src.endswith(".cpp")   # kernels were built (host-only sources have none)
        assert not re.search(r"v_pk_(add|mul|fma)_f32", asm), src


def test_pair_kernel_inner_loop_instruction_count():
    """The FP64 pair loop is 16 FP64 VALU operations plus one v_rsq_f64 per
    pair, the mixed one 14 FP32 operations (the close-pair test and select
    included) plus one v_rsq_f32 (unrolled by four; SGPR-fed source
    records)."""
    import collections
    import re
    from solaris_amd import build
    asm = build.device_assembly("pairs.hip")
    names = re.findall(r"^(_Z\S+):", asm, re.M)

    def loop_valu(tag):
        sym = [n for n in names if tag in n][0]
        start = asm.index(sym + ":")
        body = asm[start:asm.index("s_endpgm", start)]
        blocks = re.split(r"\n\.LBB\S+:[^\n]*", body)
        best = max(blocks, key=lambda b: len(re.findall(r"^\s+v_", b, re.M)))
        c = collections.Counter(l.split()[0] for l in best.split("\n")
                                if l.strip().startswith("v_"))
        return c
    f = loop_valu("k_pairs_fp64ILb0ELb0ELb0E")
    assert sum(f.values()) == 4 * 17 + 1 and f["v_rsq_f64_e32"] == 4   # + one loop-carried move
    m = loop_valu("k_pairs_mixedILb0ELb0ELb0E")
    assert sum(m.values()) == 4 * 15 and m["v_rsq_f32_e32"] == 4
