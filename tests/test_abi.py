"""The C-ABI library loads and exports every entry point include/lsh.h declares
(no compute: this runs without a GPU)."""
import os
import re

import pytest

from lapstack_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "lsh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lsh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_seam():
    names = _declared()
    for must in ["lsh_abi", "lsh_available", "lsh_free_str", "lsh_frames_new", "lsh_frames_process",
                 "lsh_fuser_new", "lsh_fuser_push", "lsh_fuser_push_frames", "lsh_fuser_finish",
                 "lsh_depth_from_slices"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(_declared()) <= bound, set(_declared()) - bound


def test_abi_version_and_probe_do_not_crash():
    lib = _lib.load()
    assert lib.lsh_abi() == _lib.ABI_VERSION
    msg = _lib.take_err(lib.lsh_available())
    assert msg is None or isinstance(msg, str)


def test_errors_come_back_as_strings():
    import ctypes as C
    lib = _lib.load()
    out = C.c_void_p()
    msg = _lib.take_err(lib.lsh_fuser_new(0, 0, None, C.byref(out)))
    assert msg and "null" in msg or "empty" in (msg or "")
    msg = _lib.take_err(lib.lsh_frames_new(0, 10, 3, C.byref(out)))
    assert msg and "empty" in msg
    msg = _lib.take_err(lib.lsh_fuser_push(None, None, 0))
    assert msg and "null" in msg
    msg = _lib.take_err(lib.lsh_depth_from_source(_lib.SLICE_FN(), None, 2, None, 0, 8, 8, None, None, None, 0, None,
                                                  None, _lib.LOG_FN(), None))
    assert msg and "null" in msg
