"""The C ABI library loads on a CPU-only host and exports every declared symbol.

No compute call is made here (there may be no GPU); only symbol presence and
the ctypes signature table are checked.
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER_DIR = os.path.join(ROOT, 'include')


def declared_functions():
    names = set()
    for fn in os.listdir(HEADER_DIR):
        if fn.endswith('.h'):
            text = open(os.path.join(HEADER_DIR, fn)).read()
            text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
            names |= set(re.findall(r'\b(qg_[a-z0-9_]+)\s*\(', text))
    return sorted(names)


def _lib():
    from qradient_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip('libqradient_amd.so not built (run __graft_entry__.build())')
    return _native, ctypes.CDLL(_native.LIB_PATH)


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ('qg_ctx_create', 'qg_mcclean_plan', 'qg_mcclean_grad_run', 'qg_qaoa_grad_run',
                     'qg_state_rot', 'qg_last_error'):
        assert required in names


def test_library_exports_every_declared_symbol():
    _native, lib = _lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    _native, _ = _lib()
    assert set(_native.SIGNATURES) == set(declared_functions())


def test_version_and_error_string_without_gpu():
    _native, lib = _lib()
    lib.qg_version.restype = ctypes.c_int
    assert lib.qg_version() >= 1
    lib.qg_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.qg_last_error(), bytes)


def test_invalid_arguments_are_rejected_before_any_device_call():
    _native, lib = _lib()
    _native.load()
    with pytest.raises(ValueError):
        _native.call('qg_mcclean_plan', None, 4, 2, 0, 0, 0, None)
    with pytest.raises(ValueError):
        _native.call('qg_plan_workspace_size', None, None)
