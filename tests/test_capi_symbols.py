"""The C-ABI library loads and exports every symbol the public headers declare
(no GPU calls)."""
import ctypes

from cudilithium_amd import _lib


def test_library_exports_all_declared_symbols():
    L = _lib.load()
    names = _lib.declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_sizes_match_reference_params():
    L = _lib.load()
    # include/params.h:71-73
    assert [L.dil_public_key_bytes(m) for m in (2, 3, 5)] == [1312, 1952, 2592]
    assert [L.dil_secret_key_bytes(m) for m in (2, 3, 5)] == [2528, 4000, 4864]
    assert [L.dil_signature_bytes(m) for m in (2, 3, 5)] == [2420, 3293, 4595]
    assert L.dil_signature_bytes(4) == 0
