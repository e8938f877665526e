"""The C-ABI library loads on a machine without a GPU and exports every
entry point include/hypertrace_hip.h declares (no compute calls)."""
import os
import re
import subprocess

from hypertrace_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "hypertrace_hip.h")).read()
    return set(re.findall(r"\b(ht_[a-z0-9_]+)\s*\(", text))


def test_header_symbols_exported():
    lib = native.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    nm = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ht_[a-z0-9_]+)", nm))
    assert names <= exported, names - exported


def test_binding_covers_header():
    assert declared() == set(native.SIGNATURES)


def test_no_device_reports_backend_error():
    import ctypes
    h = ctypes.c_void_p()
    rc = native.lib().ht_ctx_create(10_000, ctypes.byref(h))
    assert rc < 0
    assert native.lib().ht_last_error()
