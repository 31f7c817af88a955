"""The C-ABI libraries load without a GPU and export exactly the entry points
include/magic_hip.h and include/magic_h5.h declare (no compute calls here)."""
import ctypes
import os
import re

from magic_master_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "magic_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(magic_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_the_binding_exports():
    assert _declared() == sorted(hip.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = hip.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.magic_strerror(0) == b"ok"
    assert b"rho" in lib.magic_strerror(1)


def test_params_struct_matches_header():
    text = open(os.path.join(ROOT, "include", "magic_hip.h")).read()
    body = re.search(r"typedef struct \{(.*?)\} magic_params;", text, re.S).group(1)
    names = re.findall(r"([A-Za-z0-9_]+)\s*[,;]", re.sub(r"\b(int|double)\b", "", body))
    assert names == [n for n, _ in hip.CParams._fields_]


def test_create_rejects_bad_arguments_without_touching_the_gpu():
    lib = hip.load_library()
    h = ctypes.c_void_p()
    # mbc < 2 and meqn out of range are rejected before any HIP call
    assert lib.magic_create(8, 8, 8, 6, 1, 32, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, ctypes.byref(h)) == 5
    assert lib.magic_create(8, 8, 8, 11, 2, 32, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, ctypes.byref(h)) == 5
    # mz far above mx: conduct.f:86-91's out-of-row z-ghost writes would land
    # in rows written after them, an order the device does not replay
    assert lib.magic_create(8, 8, 600, 6, 2, 32, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, ctypes.byref(h)) == 4
    # unknown state form; FP32 form with mbc beyond its per-layer table
    assert lib.magic_create_form(8, 8, 8, 6, 2, 32, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 7, ctypes.byref(h)) == 5
    assert lib.magic_create_form(8, 8, 8, 6, 9, 32, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, hip.FORM_FP32P,
                                 ctypes.byref(h)) == 4


def test_h5_library_exports_every_declared_symbol():
    from magic_master_amd import h5
    text = open(os.path.join(ROOT, "include", "magic_h5.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(magic_[a-z0-9_]+)\s*\(", text)))
    assert names == ["magic_h5_frame_name", "magic_in3_h5", "magic_out3_h5"]
    lib = h5.load_library()
    for name in names:
        assert hasattr(lib, name), name
