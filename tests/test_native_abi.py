"""The C-ABI library loads and exports every symbol include/qsx.h declares
(no device calls: runs without a GPU)."""
import os
import re

from tests.conftest import ROOT


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, 'include'))):
        if fn.endswith('.h'):
            text = open(os.path.join(ROOT, 'include', fn)).read()
            text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
            names += re.findall(r'^\s*int(?:32|64)_t\s+(qsx_\w+)\s*\(', text, flags=re.M)
    return names


def test_library_exports_every_declared_symbol():
    from qspectra_amd import _native
    lib = _native.load_library()
    names = declared_functions()
    assert len(names) >= 4
    for n in names:
        assert hasattr(lib, n), n
        assert n in _native.SIGNATURES, n
    assert lib.qsx_version() >= 100


def test_argument_validation_without_device():
    """Invalid sizes are rejected on the host before any launch."""
    from qspectra_amd import _native
    lib = _native.load_library()
    st = lib.qsx_apply_dense_c128(None, 0, 0, None, 1, 1, None, None)
    assert st == -1
    assert 'null' in _native.last_error() or 'bad' in _native.last_error()
    st = lib.qsx_rk4_dense(7, None, 0, None, 0, 1, None, 1, 1, None, None, 1, None, 0, 1,
                           None, None, None, None)
    assert st == -1 and 'precision' in _native.last_error()
