"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares; the ABI structs have the reference's sizes; context
creation fails cleanly (NULL + message) when there is no device."""
import ctypes as C
import os
import re

import mav1hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mav1hip_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(mav1hip.LIB_PATH)
    names = declared_functions()
    assert len(names) > 15
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_struct_sizes():
    assert C.sizeof(mav1hip.FrameParams) == 392
    assert C.sizeof(mav1hip.BlockInfo) == 108
    assert C.sizeof(mav1hip.SbInfo) == 356


def test_create_fails_cleanly_without_device(monkeypatch):
    import conftest
    if conftest._gpu_available():
        return
    cfg = mav1hip.default_config(352, 288)
    try:
        mav1hip.Context(cfg)
    except mav1hip.HipError as e:
        assert "device" in str(e)
    else:
        raise AssertionError("context created without a device")


def test_create_rejects_bad_size():
    cfg = mav1hip.default_config(8, 8)
    try:
        mav1hip.Context(cfg)
    except mav1hip.HipError as e:
        assert "size" in str(e) or "device" in str(e)
    else:
        raise AssertionError("context created for an 8x8 frame")


def test_pool_import_rejects_missing_contexts():
    lib = C.CDLL(mav1hip.LIB_PATH)
    lib.mav1hip_pool_import.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
    assert lib.mav1hip_pool_import(None, 0, None, 0) == -1  # MAV1HIP_EINVAL
