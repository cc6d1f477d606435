"""The C-ABI library loads and exports every entry point include/pgj.h declares (CPU, no compute)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'pgj.h')
LIB = os.path.join(ROOT, 'partial-green-jacobi_amd', 'pgj', 'libpgj.so')


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r'\b(pgj_[a-z0-9_]+)\s*\(', text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason='libpgj.so not built')
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from pgj import _lib
    names = declared()
    assert sorted(_lib.SIGNATURES) == names


@pytest.mark.skipif(not os.path.exists(LIB), reason='libpgj.so not built')
def test_host_side_queries_without_gpu():
    from pgj import _lib
    lib = _lib.lib()
    assert lib.pgj_version() == 1
    g = _lib.make_grid((64, 64, 64), (1 / 64,) * 3, 15 / 26, 5 / 13)
    assert lib.pgj_plan_bytes(g) >= 6 * 64 * 64 * 33 * 8
    assert lib.pgj_resident_ws_bytes(g, 6) > 6 * 3 * 64 * 64 * 33 * 16
    assert lib.pgj_resident_state_bytes(6) >= 6 * (5 * 8 + 2)
    # argument validation happens on the host and raises through the binding
    bad = _lib.make_grid((1, 64, 64), (1 / 64,) * 3, 15 / 26, 5 / 13)
    with pytest.raises(ValueError):
        _lib.check(lib.pgj_hex8_apply(bad, 1, None, None, None, None, None))
    with pytest.raises(ValueError):
        _lib.check(lib.pgj_resident_iterate(g, 6, 6, 0, 6, None, None, None, None, None, None, None, None, None,
                                            None, None, 3, None))
