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


def test_checkpoint_library_exports():
    # include/chemps2_amd_io.h: the HDF5 checkpoint entry points (libchemps2_amd_io.so)
    text = open(os.path.join(ROOT, "include", "chemps2_amd_io.h")).read()
    names = sorted(set(re.findall(r"\b(chm_[a-z0-9_]+)\s*\(", text)))
    assert names == sorted(chemps2_amd.IO_EXPORTS)
    if not os.path.exists(chemps2_amd.IO_LIB_PATH):
        import pytest

        pytest.skip("checkpoint library not built (no HDF5)")
    lib = chemps2_amd.load_io()
    for name in names:
        assert hasattr(lib, name), name
