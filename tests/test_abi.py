"""The C-ABI library loads and exports every entry point include/chemps2_amd.h declares."""
import ctypes
import os
import re

import chemps2_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "chemps2_amd.h")).read()
    return sorted(set(re.findall(r"\b(chm_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared() == sorted(chemps2_amd.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(chemps2_amd.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_no_device_create_fails_cleanly():
    # Without a visible GPU chm_ctx_create must fail with an error code, not crash.
    import torch

    if torch.cuda.is_available():
        return
    lib = chemps2_amd.load()
    h = ctypes.c_void_p()
    assert lib.chm_ctx_create(0, ctypes.byref(h)) != 0
