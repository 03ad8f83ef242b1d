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


def test_pool_rows_fit_reference_formulas():
    """Pools sized with the reference callers' formulas
    (tests/test_vectors.cu:64-70, 81-100, 117-128, 149-166) are large enough."""
    from cudilithium_amd import api
    L = api.L()
    for mode in (2, 3, 5):
        assert L.dil_keypair_pool_row_bytes(mode) <= api.reference_row_bytes(mode, api.POOL_KEYPAIR)
        assert L.dil_sign_temp_row_bytes(mode) <= api.reference_row_bytes(mode, api.POOL_SIGN_TEMP)
        for mlen in (0, 1, 32, 33, 64, 1000, 4096, 65536):
            assert L.dil_sign_pool_row_bytes(mode, mlen) <= api.reference_row_bytes(mode, api.POOL_SIGN, mlen)
            assert L.dil_verify_pool_row_bytes(mode, mlen) <= api.reference_row_bytes(mode, api.POOL_VERIFY, mlen)
    # the reference formula itself, Dilithium3 with 32-byte messages
    assert api.reference_row_bytes(3, api.POOL_SIGN, 32) == 80896


def test_sizes_match_reference_params():
    L = _lib.load()
    # include/params.h:71-73
    assert [L.dil_public_key_bytes(m) for m in (2, 3, 5)] == [1312, 1952, 2592]
    assert [L.dil_secret_key_bytes(m) for m in (2, 3, 5)] == [2528, 4000, 4864]
    assert [L.dil_signature_bytes(m) for m in (2, 3, 5)] == [2420, 3293, 4595]
    assert L.dil_signature_bytes(4) == 0


def test_dropin_sign_rejects_pools_not_16_byte_aligned():
    """The sign pools' int32 arrays are accessed 16 bytes per lane, so the
    drop-in rejects pitches (or pool pointers) that are not multiples of 16
    before any device work -- hipMallocPitch pitches always qualify."""
    from cudilithium_amd import api
    L = api.L()
    mode, batch, mlen = 3, 4, 32
    row = L.dil_sign_pool_row_bytes(mode, mlen)
    temp = L.dil_sign_temp_row_bytes(mode)
    elems = (api.SignLutElement * batch)()
    fake = ctypes.c_void_p(0x100000)  # never dereferenced: the call returns before device work
    lut = api.TaskLutC(elems, batch, fake, fake, fake, fake, fake, fake)
    host = (ctypes.c_uint8 * 64)()
    siglen = ctypes.c_size_t()
    for sign_pitch, temp_pitch in ((row + 4, (temp + 15) // 16 * 16), ((row + 15) // 16 * 16, temp + 8)):
        rc = L.dil_crypto_sign_signature(mode, host, 3293, ctypes.byref(siglen), host, mlen, mlen, host,
                                         fake, sign_pitch, fake, temp_pitch, ctypes.byref(lut), 2048, batch, None)
        assert rc == -1, rc  # DIL_EINVAL


def test_keccak_microbench_rejects_bad_arguments():
    """Argument checks run before any device call (DIL_EINVAL = -1)."""
    L = _lib.load()
    fn = L.dil_keccak_microbench
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    us, rate = ctypes.c_double(), ctypes.c_double()
    for form, blocks, threads, nperm in ((3, 1, 64, 1), (0, 1, 128, 1), (0, 0, 64, 1), (0, 1, 64, 0)):
        assert fn(form, blocks, threads, nperm, ctypes.byref(us), ctypes.byref(rate)) == -1
    assert fn(0, 1, 64, 1, None, ctypes.byref(rate)) == -1
