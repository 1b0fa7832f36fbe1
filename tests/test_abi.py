"""The C-ABI library loads and exports every symbol include/pbamd.h declares;
the oracle exports what pb_oracle.h declares.  No GPU calls."""
import os
import re
import ctypes

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pbo?_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    import particlebase_amd as pb
    lib = pb.load()
    names = declared(os.path.join(ROOT, "include", "pbamd.h"))
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(pb.EXPORTS) <= set(names)


def test_oracle_symbols_exported(orc):
    names = declared(os.path.join(ROOT, "oracle", "pb_oracle.h"))
    lib = orc.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_product_does_not_reference_oracle():
    # the product path never loads, links or imports the oracle
    for dirpath, _, files in os.walk(os.path.join(ROOT, "particlebase_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".hh", ".cpp", ".cc")) or f == "Makefile":
                txt = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "pb_oracle" not in txt and "liboracle" not in txt, f


def test_context_fails_loudly_without_device():
    import particlebase_amd as pb
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    with pytest.raises(pb.PBError):
        pb.Context(0)


def test_null_handles_rejected():
    # a null context or system is an argument error, returned before any
    # device call (so this runs without a GPU)
    import particlebase_amd as pb
    lib = pb.load()
    u64, u32 = ctypes.c_uint64(), ctypes.c_uint32()
    assert lib.pb_ctx_limits(None, ctypes.byref(u64), ctypes.byref(u32), None, None, None) == pb.PB_ERR_ARG
    assert lib.pb_ctx_info(None, None, 0, None) == pb.PB_ERR_ARG
    assert lib.pb_ctx_sync(None) == pb.PB_ERR_ARG
    buf = (ctypes.c_double * 3)()
    assert lib.pb_sys_get(None, 0, buf) == pb.PB_ERR_ARG
    assert lib.pb_sys_set(None, 0, buf) == pb.PB_ERR_ARG
    assert lib.pb_ctx_destroy(None) == pb.PB_ERR_ARG
    assert lib.pb_dem_rewind(None) == pb.PB_ERR_ARG


def test_gr_histogram_flush_bound():
    # k_gr counts pairs in a uint32 LDS histogram: a tile pair adds at most
    # GR_T^2 pairs to a bin and the histogram is flushed at least every
    # GR_FLUSH_TILES tile pairs, which must keep every bin below 2^32
    src = open(os.path.join(ROOT, "particlebase_amd", "csrc", "pb_bin.hip")).read()
    gt = int(re.search(r"#define GR_T (\d+)", src).group(1))
    ft = int(re.search(r"#define GR_FLUSH_TILES (\d+)u", src).group(1))
    assert gt * gt * ft < 2 ** 32
    assert "++since >= GR_FLUSH_TILES" in src
