"""The C ABI: the built library loads and exports every symbol the header
declares; argument checks fire before any device work.  CPU only (no kernel
is launched)."""
import ctypes
import os
import re

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, 'include', 'mbpt_hip.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(mbpt_[a-z0-9_]+)\s*\(', text)))


@pytest.fixture(scope='module')
def lib():
    from mbptcode_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f'{_lib.LIB_PATH} is not built; run __graft_entry__.build()')
    return _lib.lib()


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
    from mbptcode_amd import _lib
    assert sorted(_lib.EXPORTS) == names


def test_abi_version(lib):
    assert lib.mbpt_abi_version() == 1


def test_argument_errors_are_value_errors(lib):
    from mbptcode_amd import _lib
    h = ctypes.c_void_p()
    eps = (ctypes.c_double * 3)(-1.0, 0.5, 1.0)
    rc = lib.mbpt_sigma_create(ctypes.byref(h), 0, 1, 2, 1, 2.0, None, None, None,
                               eps, None)
    assert rc == _lib.ERR_VALUE
    assert b'positive' in lib.mbpt_last_error()
    rc = lib.mbpt_sigma_create(ctypes.byref(h), 10, 200, 2, 1, 2.0, None, None, None,
                               eps, None)
    assert rc == _lib.ERR_VALUE and b'128' in lib.mbpt_last_error()
    with pytest.raises(ValueError):
        _lib.check(rc, 'create')
    rc = lib.mbpt_sigma_apply(None, 1, None, None, None)
    assert rc == _lib.ERR_VALUE
    rc = lib.mbpt_chi0_build(None, 1, None, None, None, 0, None)
    assert rc == _lib.ERR_VALUE
    mu_bad = ctypes.c_double(0.7)   # above a virtual level
    rc = lib.mbpt_chi0_create(ctypes.byref(h), 4, 1, 2, 1, ctypes.c_void_p(8), eps,
                              mu_bad, None)
    assert rc == _lib.ERR_VALUE and b'mu' in lib.mbpt_last_error()


def test_unknown_precision_is_unsupported(lib):
    from mbptcode_amd import _lib
    h = ctypes.c_void_p()
    eps = (ctypes.c_double * 3)(-1.0, 0.5, 1.0)
    rc = lib.mbpt_sigma_create(ctypes.byref(h), 4, 1, 2, 7, 2.0, ctypes.c_void_p(8),
                               ctypes.c_void_p(8), ctypes.c_void_p(8), eps, None)
    assert rc == _lib.ERR_UNSUPPORTED
    with pytest.raises(NotImplementedError):
        _lib.check(rc, 'create')


def test_destroy_null_is_harmless(lib):
    assert lib.mbpt_sigma_destroy(None) == 0
    assert lib.mbpt_chi0_destroy(None) == 0
    assert lib.mbpt_sigma_held_bytes(None) == 0
