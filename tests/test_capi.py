"""The C-ABI library loads and exports every entry point that include/xadv.h declares (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="xadv.h", prefix="xadv_"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_api():
    syms = declared_symbols()
    for s in ("xadv_batch_create", "xadv_step", "xadv_readout", "xadv_stats", "xadv_fields", "xadv_metrics"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from xadvcool_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libxadv.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"
    assert set(_lib.EXPORTS) == set(declared_symbols())


def test_surrogate_library_exports_every_declared_symbol():
    """libxadv_nn.so (the warm-start surrogate's kernels) against include/xadv_nn.h."""
    from xadvcool_amd import surrogate as S
    if not os.path.exists(S.NN_LIB):
        pytest.fail("libxadv_nn.so not built (run __graft_entry__.build())")
    syms = declared_symbols("xadv_nn.h", "xnn_")
    assert "xnn_conv_fwd" in syms and "xnn_conv_wgrad" in syms and "xnn_adam" in syms
    lib = ctypes.CDLL(S.NN_LIB)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_version_and_error_calls_need_no_device():
    from xadvcool_amd import _lib
    lib = _lib.load()
    assert b"gfx950" in lib.xadv_version()
    assert lib.xadv_last_error() == b""


def test_bad_arguments_are_errors_not_crashes():
    """Argument checks run before any device call."""
    from xadvcool_amd import _lib
    lib = _lib.load()
    d = _lib.BatchDesc(16, 1, 1, 0, 1, 4, 4, 4, 256, 0, 0, 0)  # precision 16 is not supported
    h = ctypes.c_void_p()
    assert lib.xadv_batch_create(ctypes.byref(d), ctypes.byref(h)) == _lib.XADV_E_ARG
    assert b"precision" in lib.xadv_last_error()
    d = _lib.BatchDesc(32, 1, 1, 0, 1, 4, 4, 4, 256, 0, 0, 0, 0, 2)  # alternate_sweep is a 0/1 flag
    assert lib.xadv_batch_create(ctypes.byref(d), ctypes.byref(h)) == _lib.XADV_E_ARG
    assert b"alternate_sweep" in lib.xadv_last_error()
    d = _lib.BatchDesc(32, 1, 1, 0, 1, 4, 4, 4, 256, 0, 0, 0, 0, 1, 9)  # thermal lattice D3Q9 does not exist
    assert lib.xadv_batch_create(ctypes.byref(d), ctypes.byref(h)) == _lib.XADV_E_ARG
    assert b"thermal_q" in lib.xadv_last_error()
    for fluid, general in ((0, 0), (1, 1)):  # the D3Q7 lattice runs coupled, channel boundaries only
        d = _lib.BatchDesc(32, fluid, 1, general, 1, 4, 4, 4, 256, 0, 0, 0, 0, 1, 7)
        assert lib.xadv_batch_create(ctypes.byref(d), ctypes.byref(h)) == _lib.XADV_E_ARG
        assert b"D3Q7" in lib.xadv_last_error()
    # the thermal prefetch is a 0/1 flag of the fused D3Q19 pass on compact storage: not with dense storage,
    # the thermal lattice off, or the D3Q7 lattice
    for storage, thermal, q, flag in ((0, 1, 19, 2), (_lib.STORAGE_DENSE, 1, 19, 1), (0, 0, 19, 1), (0, 1, 7, 1)):
        d = _lib.BatchDesc(32, 1, thermal, 0, 1, 4, 4, 4, 256, 0, 0, 0, storage, 1, q, 0, flag)
        assert lib.xadv_batch_create(ctypes.byref(d), ctypes.byref(h)) == _lib.XADV_E_ARG
        assert b"thermal_prefetch" in lib.xadv_last_error()
    assert lib.xadv_step(None, 1, 0) == _lib.XADV_E_ARG
