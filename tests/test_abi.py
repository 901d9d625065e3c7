"""The C-ABI library loads and exports every function include/emd_gfx950.h declares, and
the Python prototypes cover exactly that set.  No device calls: runs on the CPU."""
import os
import re

from cuda_emd import _lib

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "emd_gfx950.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return set(re.findall(r"\b(emd_[a-z_]+)\s*\(", text))


def test_library_exports_every_declared_function():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_prototypes_match_header():
    assert set(_lib.PROTOTYPES) == declared_functions()


def test_host_only_queries():
    lib = _lib.load()
    assert lib.emd_version() >= 1
    assert lib.emd_lds_slot_bytes() == 28
    for wg in (64, 128, 256, 512, 1024):
        assert 0 < lib.emd_static_lds_bytes(wg) < 4096
