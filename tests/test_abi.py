"""The C-ABI library loads and exports every symbol include/*.h declares (no GPU calls)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbhip_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    from nbody_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    headers = glob.glob(os.path.join(ROOT, "include", "*.h"))
    assert headers
    names = []
    for h in headers:
        names += declared_symbols(h)
    assert "nbhip_step" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers exactly the declared surface
    assert set(_lib.SIGNATURES) == set(n for n in names if n in _lib.SIGNATURES)
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_default_params_are_the_reference_defaults():
    from nbody_amd import _lib
    p = _lib.params()
    # include/nbody/state.h:23-32
    assert p.size == 10000.0 and abs(p.theta - 0.5) < 1e-7 and p.G == 1.0 and p.wrap == 1
    assert p.eps == 0.0 and p.mode == _lib.MODE_N2 and p.leaf_size == 1


def test_no_device_is_reported_not_raised():
    # On a host without a GPU the probe reports a reason; with a gfx950 device it succeeds.
    from nbody_amd import _lib
    ok, reason = _lib.probe()
    assert ok or reason


def test_body_limits_refused_without_a_device():
    # the count checks come before the context check, so they answer on any machine; the
    # text of a failure with no context is the thread's NULL-context message
    import ctypes as C
    from nbody_amd import _lib
    L = _lib.load()
    assert L.nbhip_reserve(None, _lib.MAX_BODIES + 1) == _lib.ERR_INVALID
    assert b"exceeds the limit" in L.nbhip_last_error(None)
    assert L.nbhip_reserve(None, 2 ** 31) == _lib.ERR_INVALID
    assert L.nbhip_reserve(None, 2 ** 40) == _lib.ERR_INVALID
    assert b"exceeds the limit" in L.nbhip_last_error(None)
    # within the limit the missing context is what is reported
    assert L.nbhip_reserve(None, _lib.MAX_BODIES) == _lib.ERR_INVALID
    assert b"null context" in L.nbhip_last_error(None)
    spec = _lib.ics(_lib.IC_PLUMMER)
    assert L.nbhip_generate(None, C.byref(spec), 2 ** 31) == _lib.ERR_INVALID
    assert b"exceeds the limit" in L.nbhip_last_error(None)
    assert L.nbhip_generate(None, C.byref(spec), 1000) == _lib.ERR_INVALID
    assert b"null context" in L.nbhip_last_error(None)
    # the largest count keeps a tile's padding lanes inside int
    assert _lib.MAX_BODIES + 4096 < 2 ** 31 - 1
    assert _lib.MAX_NLOGN_BODIES < 2 ** 30
