"""CPU checks of the C ABI: the libraries load and export every declared symbol."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(f2\w+)\s*\(", text)))


def test_f2jk_exports_every_declared_symbol():
    from forte2_amd import _native

    lib = _native.jk_lib()
    names = _declared("f2jk.h")
    assert "f2jk_build_jk" in names and len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_native.F2JK_SYMBOLS)


def test_f2jk_errors_without_gpu_are_reported():
    from forte2_amd import _native

    lib = _native.jk_lib()
    # a NULL context is rejected with a message, no device touched
    rc = lib.f2jk_set_B(None, 0, None, 1, 1)
    assert rc == -1
    assert b"null context" in lib.f2jk_last_error()
    rc = lib.f2jk_ctx_destroy(None)
    assert rc == 0


def test_f2probe_exports_every_declared_symbol():
    from forte2_amd import _native

    lib = _native.probe_lib()
    names = _declared("f2probe.h")
    assert set(names) == set(_native.F2PROBE_SYMBOLS)
    for n in names:
        assert hasattr(lib, n), n
    # argument checks run before any device call
    ms = ctypes.c_double()
    assert lib.f2probe_read(None, 16, 1, None, 0, None, ctypes.byref(ms)) == -1
    assert b"null argument" in lib.f2probe_last_error()


def test_f2ints_exports_every_declared_symbol():
    from forte2_amd import _ints_proto, _native

    lib = _native.ints_lib()
    names = _declared("f2ints.h")
    assert "f2ints_coulomb_3c_pairs_aux" in names
    assert set(names) == set(_ints_proto.F2INTS_SYMBOLS)
    for n in names:
        assert hasattr(lib, n), n

