'''
The C ABI (CPU): libowmi.so loads and exports every function declared in
include/owmi.h, and the Python binding declares exactly those. No compute
call is made without a GPU.
'''
import ctypes
import os
import re

from optweight_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, 'include', 'owmi.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(owmi_[a-z0-9_]+)\s*\(', txt)))


def test_library_exports_header():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
        assert isinstance(getattr(lib, n), ctypes._CFuncPtr)


def test_binding_matches_header():
    assert _lib.symbols() == header_functions()


def test_abi_version_and_error_string():
    lib = _lib.load()
    assert lib.owmi_abi_version() == 1
    assert isinstance(lib.owmi_last_error(), bytes)


def test_invalid_arguments_fail_cleanly():
    '''Argument validation runs on the host before any device work.'''
    lib = _lib.load()
    assert lib.owmi_lmul(7, None, 0, None, None, 10, 1, 1, None) != 0
    assert b'dtype' in lib.owmi_last_error()
    assert lib.owmi_trunc_alm(_lib.F64, None, None, 5, 9, 1, None) != 0
    assert b'lmax_out' in lib.owmi_last_error()
    assert lib.owmi_sht_synthesis(None, _lib.F64, 0, None, 0, None, 0, 1, 0, None) != 0
