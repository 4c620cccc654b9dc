"""The C-ABI library loads and exports every entry point include/b2g.h declares."""
import ctypes
import os
import re

from conftest import REPO
from pypybox2d_amd import _native as N


def declared():
    text = open(os.path.join(REPO, "include", "b2g.h")).read()
    return set(re.findall(r"B2G_API\s+int\s+(b2g_\w+)\s*\(", text))


def test_header_and_binding_agree():
    assert declared() == set(N.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = N.load()
    for name in declared():
        assert hasattr(lib, name), name
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr)
    assert lib.b2g_version() == 1


def test_create_without_gpu_fails_cleanly_or_succeeds():
    lib = N.load()
    h = ctypes.c_void_p()
    desc = N.CDesc(N.B2G_FP32, 0, 8.0, 64, 0)
    rc = lib.b2g_create(ctypes.byref(desc), ctypes.byref(h))
    assert rc in (N.B2G_OK, N.B2G_E_NODEVICE)
    if rc == N.B2G_OK:
        assert lib.b2g_destroy(h) == N.B2G_OK
    bad = N.CDesc(7, 0, 8.0, 64, 0)
    assert lib.b2g_create(ctypes.byref(bad), ctypes.byref(h)) in (N.B2G_E_INVALID, N.B2G_E_NODEVICE)


def test_parameter_keys_agree():
    text = open(os.path.join(REPO, "include", "b2g.h")).read()
    keys = {k.lower(): int(v) for k, v in re.findall(r"#define B2G_PARAM_(\w+) (\d+)", text)}
    assert keys == N.PARAMS
