"""dil_verify_device_noalloc (work rows in the library's fixed device array)
against dil_verify_device (caller workspace) through the C-ABI: two calls in a
row on the same GPU, the second with a smaller batch, give the same verdicts
as the workspace entry point (and the oracle's on a sample); batches beyond
the array's capacity are refused with DIL_ERANGE."""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _lib():
    from cudilithium_amd import api
    L = api.L()
    L.dil_verify_noalloc_max_items.restype = ctypes.c_size_t
    L.dil_verify_noalloc_max_items.argtypes = [ctypes.c_int]
    L.dil_verify_device_noalloc.restype = ctypes.c_int
    L.dil_verify_device_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                            ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    return L


def _inputs(mode, n, seed):
    from cudilithium_amd import api
    from cudilithium_amd.device import DeviceBuffer
    rng = np.random.default_rng(seed)
    seeds = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    pk, sk = api.keypair_device(mode, seeds=seeds)
    msgs = [rng.integers(0, 256, int(rng.integers(0, 1500)), dtype=np.uint8).tobytes() for _ in range(n)]
    sig = api.SignContext(mode, n, n).sign(msgs, sk)
    flip = rng.choice(n, n // 7, replace=False)
    for i in flip:
        b = int(rng.integers(0, sig.shape[1] * 8))
        sig[i, b // 8] ^= 1 << (b % 8)
    return pk, msgs, sig, flip, DeviceBuffer.from_array(pk), DeviceBuffer.from_array(sig), api.DeviceMessages(msgs)


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_verify_noalloc_matches_workspace_entry_point(mode):
    from cudilithium_amd import api
    from cudilithium_amd.device import DeviceBuffer, synchronize
    L = _lib()
    cap = L.dil_verify_noalloc_max_items(mode)
    assert cap == (16 << 20) // {2: 1792, 3: 1792, 5: 2048}[mode]
    pkb, _, sigb = api.sizes(mode)
    wrow = L.dil_verify_work_row_bytes(mode)
    for call, n in enumerate((cap, 777)):  # the full array, then a smaller batch over its leftovers
        pk, msgs, sig, flip, d_pk, d_sig, dm = _inputs(mode, n, 100 * mode + call)
        d_ret, d_ret2 = DeviceBuffer(4 * n), DeviceBuffer(4 * n)
        d_ret.memset(0x5A)
        assert L.dil_verify_device_noalloc(mode, d_ret.ptr, d_sig.ptr, sigb, d_pk.ptr, pkb, dm.data.ptr, dm.offs.ptr,
                                           dm.lens.ptr, n, None) == 0
        d_work = DeviceBuffer(n * wrow)
        assert L.dil_verify_device(mode, d_ret2.ptr, d_sig.ptr, sigb, d_pk.ptr, pkb, dm.data.ptr, dm.offs.ptr,
                                   dm.lens.ptr, n, d_work.ptr, wrow, None) == 0
        synchronize()
        got, ref = d_ret.download((n,), np.int32), d_ret2.download((n,), np.int32)
        assert np.array_equal(got, ref)
        want = np.zeros(n, np.int32)
        want[flip] = -1
        assert np.array_equal(got, want)
        sample = np.random.default_rng(call).choice(n, 64, replace=False)
        assert np.array_equal(got[sample], oracle.verify_batch(mode, sig[sample], [msgs[i] for i in sample],
                                                               pk[sample]))
    # beyond the array: refused before any GPU work
    assert L.dil_verify_device_noalloc(mode, d_ret.ptr, d_sig.ptr, sigb, d_pk.ptr, pkb, dm.data.ptr, dm.offs.ptr,
                                       dm.lens.ptr, cap + 1, None) == -3
