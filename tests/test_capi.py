"""The C-ABI library loads and exports every symbol include/clblm_hip.h declares."""
import ctypes
import os
import re

from clblm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "clblm_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char \*|double)\s*\*?\s*(clblm_\w+)\s*\(", hdr, re.M))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_lib.EXPORTS), declared ^ set(_lib.EXPORTS)


def test_version_without_gpu():
    lib = _lib.load()
    assert lib.clblm_version() >= 1


def test_level2layer_matches_the_oracle():
    """clblm_level2layer is host code: bit-for-bit the oracle's level2layer
    (clblm_noScattJac.f90:1208-1279), and it rejects a layer of uniform density."""
    import numpy as np
    from oracle.py import level2layer
    rng = np.random.default_rng(7)
    nlay, nmol = 9, 3
    p = np.exp(np.linspace(np.log(1013.0), np.log(2.0), nlay + 1))
    t = 250.0 + rng.normal(0, 20, nlay + 1)
    tl = 0.5 * (t[1:] + t[:-1]) + rng.normal(0, 1, nlay)
    q = np.abs(rng.normal(1e17, 3e16, (nmol, nlay + 1)))
    q[2, 4] = 0.0
    lib = _lib.load()
    outs = [np.zeros(nlay), np.zeros(nlay), np.zeros((nmol, nlay)), np.zeros((nmol, nlay))]
    _lib.check(lib.clblm_level2layer(nlay, nmol, p.ctypes.data, t.ctypes.data, tl.ctypes.data, q.ctypes.data,
                                     *[o.ctypes.data for o in outs]))
    for a, b in zip(outs, level2layer(p, t, tl, q)):
        assert np.array_equal(a, b)
    t2, p2 = t.copy(), p.copy()
    t2[3] = t2[2]
    p2[3] = p2[2]          # layer 2 with the same density at both levels
    assert lib.clblm_level2layer(nlay, nmol, p2.ctypes.data, t2.ctypes.data, tl.ctypes.data, q.ctypes.data,
                                 *[o.ctypes.data for o in outs]) != 0
