"""The C-ABI library loads and exports every entry point include/mrg.h declares (no GPU needed)."""
import ctypes
import os
import re

from minirender_amd import _native as N

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "mrg.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mrg_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("mrg_device_open", "mrg_scene_upload", "mrg_intersect", "mrg_job_submit", "mrg_job_pause",
                 "mrg_job_resume", "mrg_job_cancel", "mrg_job_read_film", "mrg_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) == set(N.SIGNATURES)


def test_version_and_error_without_device():
    L = N.lib()
    assert b"gfx950" in L.mrg_version()
    h = ctypes.c_void_p()
    rc = L.mrg_device_open(-1, None, ctypes.byref(h))
    assert rc != 0
    assert L.mrg_last_error()
