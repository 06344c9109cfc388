"""The C-ABI library loads and exports every entry point its header declares
(no device calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    inc = os.path.join(ROOT, 'include')
    for f in os.listdir(inc):
        if f.endswith('.h'):
            txt = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r'\b(mdf_\w+)\s*\(', txt))
    return names


def test_library_exports_header_symbols():
    from electrolites_amd import _lib
    lib = _lib.load()
    names = declared()
    assert names, 'no declarations found'
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.mdf_abi_version() == _lib.ABI_VERSION


def test_binding_covers_header():
    from electrolites_amd import _lib
    assert declared() <= set(_lib.exported())


def test_workspace_queries_are_host_only():
    from electrolites_amd import _lib
    assert _lib.query('mdf_get_j_workspace', 7000, 1180416, 1) > 0
    assert _lib.query('mdf_get_k_workspace', 7000, 1536, 320, 0, 0, 256) > 256 * 1536 * 320 * 8
    assert _lib.query('mdf_get_k_workspace', 7000, 1536, 320, 0, 1, 256) > 256 * 1536 * 320 * 4
