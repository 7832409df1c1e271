"""The C ABI library loads and exports every function include/*.h declares
(CPU: no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(gal_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gal_create" in names and "gal_run_churn" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    import gallatin_amd
    lib = gallatin_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(gallatin_amd.CAPI_SYMBOLS) == declared_functions()


def test_struct_layouts_match_header():
    import gallatin_amd
    assert ctypes.sizeof(gallatin_amd.Config) == 8 * 5 + 4 * 10
    assert ctypes.sizeof(gallatin_amd.Info) == 8 * 5 + 4 * 4 + 8 * 3 + 8 * 24 + 8 * 3 + 3 * 8 * 24
    assert gallatin_amd.STEP_DTYPE.itemsize == 8 + 8 + 64 * 8


def test_errors_are_reported_not_fatal():
    """Invalid geometry is rejected on the host before any device call."""
    import gallatin_amd
    lib = gallatin_amd.lib()
    cfg = gallatin_amd.Config()
    lib.gal_config_default(ctypes.byref(cfg), 1 << 30, 1)
    cfg.bytes_per_segment = 4 << 20  # < 4096 * 4096
    h = ctypes.c_void_p()
    rc = lib.gal_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b"4096" in lib.gal_last_error()


def test_flag_constants_match_header():
    """Every GAL_FLAG_* of gallatin_capi.h has the same value in the Python
    package."""
    import gallatin_amd
    text = open(os.path.join(ROOT, "include", "gallatin_capi.h")).read()
    flags = dict(re.findall(r"#define (GAL_FLAG_\w+) (\d+)u", text))
    assert "GAL_FLAG_LAZY_BOOT" in flags and len(flags) >= 5
    for name, value in flags.items():
        assert getattr(gallatin_amd, name) == int(value), name


def test_lazy_boot_with_deterministic_mode_is_refused():
    """Rejected on the host, before any device call."""
    import gallatin_amd
    lib = gallatin_amd.lib()
    cfg = gallatin_amd.Config()
    lib.gal_config_default(ctypes.byref(cfg), 1 << 30, 1)
    cfg.flags = gallatin_amd.GAL_FLAG_LAZY_BOOT | gallatin_amd.GAL_FLAG_DETERMINISTIC
    h = ctypes.c_void_p()
    assert lib.gal_create(ctypes.byref(cfg), ctypes.byref(h)) == -1 and not h.value
    assert b"eagerly" in lib.gal_last_error()


def test_product_does_not_import_oracle():
    """The shipped package never reaches into oracle/."""
    pkg = os.path.join(ROOT, "gallatin_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in text.replace("oracle/_ref", ""), fn
    for dirpath, _, files in os.walk(os.path.join(ROOT, "include")):
        for fn in files:
            assert "oracle" not in open(os.path.join(dirpath, fn)).read(), fn
