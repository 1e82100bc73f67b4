"""CPU checks of the C-ABI library: it loads without a GPU and exports every
function that include/*.h declares; parameter helpers agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["include/gpufhe_kernels.h", "include/gpufhe_capi.h"]


def declared():
    names = set()
    for h in HEADERS:
        src = open(os.path.join(ROOT, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b((?:gf|Launch)\w+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    import gpufhe
    names = declared()
    assert len(names) > 40
    missing = sorted(n for n in names if not hasattr(gpufhe.lib, n))
    assert not missing, missing


def test_native_params_match_oracle():
    import gpufhe
    for count, bits, n in ((4, 35, 1024), (24, 50, 1 << 16), (6, 60, 1 << 16)):
        out = np.zeros(count, np.uint64)
        assert gpufhe.lib.gf_native_primes(out.ctypes.data_as(gpufhe.u64p), count, bits, n, None, 0) == 0
        assert np.array_equal(out, O.native_primes(count, bits, n))
        assert gpufhe.lib.gf_native_root(n, int(out[0])) == O.lib.or_native_root(n, int(out[0]))
    mods = O.native_primes(6, 50, 1 << 12)
    s1, s2 = np.zeros(5, np.uint64), np.zeros(5, np.uint64)
    gpufhe.lib.gf_native_rescale_consts(s1.ctypes.data_as(gpufhe.u64p), s2.ctypes.data_as(gpufhe.u64p),
                                        mods.ctypes.data_as(gpufhe.u64p), 5)
    r1, r2 = np.zeros(5, np.uint64), np.zeros(5, np.uint64)
    O.lib.or_rescale_consts(O.P(r1), O.P(r2), O.P(mods), 5)
    assert np.array_equal(s1, r1) and np.array_equal(s2, r2)


def test_errors_are_reported_not_crashing():
    import gpufhe
    out = np.zeros(1, np.uint64)
    # 70-bit primes are out of range: error code + message, no abort
    assert gpufhe.lib.gf_native_primes(out.ctypes.data_as(gpufhe.u64p), 1, 70, 1024, None, 0) == -1
    assert b"bits" in gpufhe.lib.gf_last_error()
    _ = C
