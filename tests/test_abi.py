"""The C-ABI library loads without a GPU and exports every entry point that
include/mambart.h declares; the Python binding declares each of them."""
import os
import re
import subprocess

from mambatracer_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mambart.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int32_t|const char\*)\s+(mt_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "mt_trace_closest" in names and "mt_compact" in names and len(names) > 40


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mt_\w+)", out))
    assert set(declared()) <= exported, set(declared()) - exported


def test_binding_covers_header():
    assert set(declared()) == set(_lib.SIGNATURES), set(declared()) ^ set(_lib.SIGNATURES)


def test_error_path_without_gpu_call():
    """Host-only entry points report argument errors through mt_last_error."""
    lib = _lib.load()
    rc = lib.mt_bvh_build(0, None, None, None, None, None, 1)
    assert rc == -1
    assert "mt_bvh_build" in _lib.last_error()


def test_library_has_gfx950_code():
    out = subprocess.run(["readelf", "-S", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert ".hip_fatbin" in out
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_binding_argument_counts_match_header():
    """Each ctypes signature has as many parameters as the header's declaration
    (a drifted binding would pass arguments into the wrong slots)."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    decls = re.findall(r"^\s*(?:int32_t|const char\*)\s+(mt_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S)
    assert len(decls) > 40
    for name, params in decls:
        params = params.strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        assert name in _lib.SIGNATURES, name
        assert len(_lib.SIGNATURES[name]) == n, (name, n, len(_lib.SIGNATURES[name]))
