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


def test_fourier_workspace_covers_the_chirp_tables():
    """Short grids need the N-entry twiddle table only; from 2048 points on
    the workspace also holds the 2N-entry half-angle table and one 8192-point
    kernel spectrum per piece/chunk pair (two for 4096 samples onto 8191
    points, the absorption grid)."""
    from qspectra_amd import _native
    lib = _native.load_library()
    assert lib.qsx_fourier_workspace_bytes(0) == -1
    assert lib.qsx_fourier_workspace_bytes(255) == 255 * 16
    assert lib.qsx_fourier_workspace_bytes(2047) == 2047 * 16
    assert lib.qsx_fourier_workspace_bytes(8191) >= (3 * 8191 + 2 * 8192) * 16


def _c_kind(param):
    """ctypes class a C parameter declaration passes as."""
    import ctypes
    p = param.strip()
    if re.match(r'(const\s+)?char\s*\*', p):
        return ctypes.c_char_p
    if '*' in p or p.startswith('hipStream_t'):
        return ctypes.c_void_p
    base = p.rsplit(None, 1)[0].replace('const ', '').strip()
    return {'int32_t': ctypes.c_int32, 'int64_t': ctypes.c_int64, 'double': ctypes.c_double,
            'size_t': ctypes.c_size_t}[base]


def declared_prototypes():
    protos = {}
    for fn in sorted(os.listdir(os.path.join(ROOT, 'include'))):
        if fn.endswith('.h'):
            text = open(os.path.join(ROOT, 'include', fn)).read()
            text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
            text = re.sub(r'//[^\n]*', '', text)
            for ret, name, args in re.findall(r'^\s*(int(?:32|64)_t)\s+(qsx_\w+)\s*\(([^)]*)\)',
                                               text, flags=re.M):
                params = [a for a in args.split(',') if a.strip() not in ('', 'void')]
                protos[name] = (ret, [_c_kind(a) for a in params])
    return protos


def test_ctypes_signatures_match_header_prototypes():
    """Every argument list handed to ctypes has the count and the C type of
    the header's prototype, position by position, and the return type (a
    mismatch passes garbage or truncates a pointer on the host)."""
    import ctypes
    from qspectra_amd import _native
    protos = declared_prototypes()
    assert set(protos) == set(_native.SIGNATURES)
    for name, (ret, kinds) in protos.items():
        got = _native.SIGNATURES[name]
        assert len(got) == len(kinds), (name, len(got), len(kinds))
        for i, (a, b) in enumerate(zip(got, kinds)):
            assert a is b, (name, i, a, b)
        want_ret = ctypes.c_int64 if ret == 'int64_t' else ctypes.c_int32
        assert _native.RESTYPES.get(name, ctypes.c_int32) is want_ret, name
