"""The C ABI without a GPU: the library loads, exports every symbol include/rdhip.h
declares, and rejects invalid arguments before touching the device."""

import ctypes as C
import os
import re

import pytest

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "rdhip.h")).read()
    return sorted(set(re.findall(r"\b(rdhip_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from ridgedetection_amd import _lib
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(_lib.EXPORTS) == set(names)


def test_invalid_arguments_are_rejected_without_a_device():
    from ridgedetection_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.rdhip_create(C.byref(h), 0, 0, None) == -1          # dim 0
    assert b"dimension" in lib.rdhip_last_error()
    assert lib.rdhip_create(C.byref(h), 0, 13, None) == -1         # dim 13
    assert lib.rdhip_create(C.byref(h), 7, 3, None) == -1          # dtype
    assert lib.rdhip_create(None, 0, 3, None) == -1
    assert lib.rdhip_ridge_detection(None, 0.1, 0.2, 0, None) == -1
    assert lib.rdhip_ridge_detection(None, -1.0, 0.2, 0, None) == -1
    assert lib.rdhip_tiled_ridge_detection(None, 99, 0.1, 0.2, 100, None, 0, 0, None) == -1
    assert lib.rdhip_tiled_ridge_detection(None, 5, 0.1, 0.2, 2, None, 0, 0, None) == -1
    assert lib.rdhip_set_points(None, None, -5, 0, 0, 0) == -1
    assert lib.rdhip_destroy(None) == 0
    assert lib.rdhip_version() >= 100


def test_noalloc_choose_validates_before_touching_the_device():
    from ridgedetection_amd import _lib
    lib = _lib.load()
    fake = C.c_void_p(256)   # never dereferenced: each call fails validation first
    assert lib.rdhip_choose_noalloc(0, 3, fake, (1 << 24) + 1, 0, 0.1, fake, 0, fake, fake, None) == -1
    assert b"RDHIP_NOALLOC_MAX_POINTS" in lib.rdhip_last_error()
    assert lib.rdhip_choose_noalloc(0, 3, fake, 0, 0, 0.1, fake, 0, fake, fake, None) == -1
    assert lib.rdhip_choose_noalloc(5, 3, fake, 10, 0, 0.1, fake, 0, fake, fake, None) == -1
    assert lib.rdhip_choose_noalloc(0, 3, fake, 10, 0, 0.1, fake, 9, fake, fake, None) == -1
