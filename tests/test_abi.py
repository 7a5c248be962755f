"""libcsx.so loads and exports every entry point declared in include/csx.h."""
import ctypes
import os
import re
import subprocess

from crystalscatter_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "csx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(csx_[a-z0-9_]+)\s*\(", text))


def test_header_declares_the_boundary():
    d = _declared()
    for name in ("csx_create", "csx_destroy", "csx_upload", "csx_render", "csx_fit_fqs",
                 "csx_fit_batch_run", "csx_engine_calc", "csx_engine_fit"):
        assert name in d


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in sorted(_declared()):
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (csx_[a-z0-9_]+)", out))
    assert _declared() <= exported


def test_python_signatures_cover_the_header():
    assert {n for n, _, _ in _lib.SIGNATURES} == _declared()


def test_scalars_struct_layout_matches_header():
    # the C struct has 11 int32, 22 + 18 + 1 + 15 doubles, then 6 int32 (with natural alignment)
    assert ctypes.sizeof(_lib.Scalars) == 11 * 4 + 4 + (22 + 18 + 1 + 15) * 8 + 6 * 4
