"""The C-ABI library loads and exports exactly what include/gdpd.h declares.

No compute call is made (this runs without a GPU).
"""
import ctypes as C
import os
import re

from gpumdpd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gdpd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gdpd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gdpd_create" in names and "gdpd_pair_compute" in names and len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_prototypes_cover_header():
    assert sorted(_lib.PROTOTYPES) == declared_functions()


def test_load_binds_prototypes():
    lib = _lib.load()
    assert lib.gdpd_create.restype == C.c_int


def test_system_bytes_without_gpu():
    from gpumdpd_amd import make_desc, system_bytes
    d = make_desc(24000, 1, (0, 0, 0), (20, 20, 20), 0.04)
    b = system_bytes(d)
    assert 24000 * 200 < b < 24000 * 400 + (1 << 24)
