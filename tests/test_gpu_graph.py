"""HIP-graph captured pipelines (dil_*_graph_create / dil_graph_launch): every
replay equals the eager call's bytes and the oracle's, with eager work and
synchronisations between replays and the outputs cleared before each one."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

REPLAYS = 3


def _msgs(seed, n, lo, hi):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, int(rng.integers(lo, hi + 1)), dtype=np.uint8).tobytes() for _ in range(n)]


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_sign_graph_replays_match_eager_and_oracle(mode):
    from cudilithium_amd import api
    from cudilithium_amd.device import DeviceBuffer, synchronize
    n = 96
    _, sk = oracle.keypair_counter_batch(mode, n, rand_index=1000 + mode)
    msgs = _msgs(mode, n, 0, 700)
    exp, exp_att = oracle.sign_batch(mode, msgs, sk)
    _, _, sigb = api.sizes(mode)
    ctx = api.SignContext(mode, n, n)
    d_sk, dm = DeviceBuffer.from_array(sk), api.DeviceMessages(msgs)
    d_sig, d_eager = DeviceBuffer(n * sigb), DeviceBuffer(n * sigb)
    g = ctx.graph(d_sig, d_sk, dm)
    assert g.n == n
    for r in range(REPLAYS):
        d_sig.memset(0)
        g.launch()
        synchronize()
        got = d_sig.download((n, sigb), np.uint8)
        assert np.array_equal(got, exp), f"replay {r}"
        assert np.array_equal(ctx.attempt_counts(n), exp_att), f"replay {r}"
        assert ctx.status() == 0
        # eager call on the same context between replays
        ctx.sign_device(d_eager, d_sk, dm)
        synchronize()
        assert np.array_equal(d_eager.download((n, sigb), np.uint8), exp)


def test_verify_graph_replays_with_bit_flips():
    from cudilithium_amd import api
    from cudilithium_amd.device import DeviceBuffer, synchronize
    mode, n = 5, 80
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=31)
    msgs = _msgs(7, n, 0, 1500)
    sig, _ = oracle.sign_batch(mode, msgs, sk)
    flip = np.arange(0, n, 7)
    rng = np.random.default_rng(8)
    for i in flip:
        b = int(rng.integers(0, sig.shape[1] * 8))
        sig[i, b // 8] ^= 1 << (b % 8)
    exp = oracle.verify_batch(mode, sig, msgs, pk)
    assert (exp[flip] == -1).all() and (np.delete(exp, flip) == 0).all()
    d_sig, d_pk, dm = DeviceBuffer.from_array(sig), DeviceBuffer.from_array(pk), api.DeviceMessages(msgs)
    wrow = api.L().dil_verify_work_row_bytes(mode)
    d_work, d_ret = DeviceBuffer(n * wrow), DeviceBuffer(4 * n)
    g = api.verify_graph(mode, d_ret, d_sig, d_pk, dm, d_work, wrow)
    for r in range(REPLAYS):
        d_ret.memset(0x55)
        g.launch()
        synchronize()
        assert np.array_equal(d_ret.download((n,), np.int32), exp), f"replay {r}"
        # eager verify of a different batch through the same work rows between replays
        assert np.array_equal(api.verify_device(mode, sig[:9], msgs[:9], pk[:9]), exp[:9])


@pytest.mark.parametrize("mode", [2, 5])
def test_keypair_graph_replays_match_oracle(mode):
    from cudilithium_amd import api
    from cudilithium_amd.device import DeviceBuffer, synchronize
    n = 50
    seeds = np.random.default_rng(40 + mode).integers(0, 256, (n, 32), dtype=np.uint8)
    epk, esk = oracle.keypair_batch(mode, seeds)
    pkb, skb, _ = api.sizes(mode)
    d_pk, d_sk = DeviceBuffer(n * pkb), DeviceBuffer(n * skb)
    wrow = (api.L().dil_keypair_work_row_bytes(mode) + 255) // 256 * 256
    d_work, d_seeds = DeviceBuffer(n * wrow), DeviceBuffer.from_array(seeds)
    g = api.keypair_graph(mode, d_pk, d_sk, n, d_work, wrow, d_seeds=d_seeds)
    for r in range(REPLAYS):
        d_pk.memset(0)
        d_sk.memset(0)
        g.launch()
        synchronize()
        assert np.array_equal(d_pk.download((n, pkb), np.uint8), epk), f"replay {r}"
        assert np.array_equal(d_sk.download((n, skb), np.uint8), esk), f"replay {r}"
        pk2, _ = api.keypair_device(mode, seeds=seeds[:3])
        assert np.array_equal(pk2, epk[:3])
