"""Cross-checks the oracle's SHAKE against the reference's own host C code
(src/fips202/fips202.c), built from the reference tree into oracle/_ref/ by
oracle/Makefile.ref.  Skipped where that library has not been built (the
reference tree is not available everywhere)."""
import ctypes
import os

import numpy as np
import pytest

import oracle

REF_LIB = os.path.join(os.path.dirname(oracle.__file__), "_ref", "libref_fips202.so")


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (no reference tree here)")
@pytest.mark.parametrize("fn", ["shake128", "shake256"])
def test_oracle_shake_matches_reference_fips202(fn):
    ref = ctypes.CDLL(REF_LIB)
    f = getattr(ref, fn)
    f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
    rng = np.random.default_rng(11)
    for inlen in (0, 1, 8, 33, 135, 136, 137, 168, 169, 1000, 2592, 4128):
        data = rng.integers(0, 256, inlen + 1, dtype=np.uint8)
        for outlen in (1, 32, 136, 168, 841):
            out = np.zeros(outlen, np.uint8)
            f(out.ctypes.data, outlen, data.ctypes.data, inlen)
            mine = getattr(oracle, fn)(data[:inlen].tobytes(), outlen)
            assert out.tobytes() == mine, (inlen, outlen)
