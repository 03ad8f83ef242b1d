"""Full-size parity of the two batched configurations the engine is measured
on, at their stated sizes, through HIP-graph replay:

- Dilithium5 verify of 65,536 engine-signed (pk, 1 KiB message, sig)
  triples, exactly round(0.1 * 65536) = 6554 of them (Fisher-Yates choice)
  with one flipped bit: every verdict equals the oracle's, flipped items -1,
  the others 0 (the reference's gpu_verify, src/verify.cu:28-325, and
  crypto_sign_verify, src/api.cu:454-515).
- Dilithium3 signing of 16,384 items with log-uniform 32 B - 4 KiB messages,
  keys from seeds, with the autotuned tiling: every signature and every
  per-item attempt count equals the oracle's (crypto_sign_signature,
  src/api.cu:63-452).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_config4_dil5_verify_65536_graph_all_verdicts():
    from cudilithium_amd import api, workloads
    from cudilithium_amd.device import DeviceBuffer, synchronize
    mode, n = 5, 65536
    seeds, msgs, flip_idx, flip_bit = workloads.config4(n)
    assert len(flip_idx) == 6554 and len(np.unique(flip_idx)) == 6554
    pkb, skb, sigb = api.sizes(mode)
    mlist = [m.tobytes() for m in msgs]
    pk, sk = api.keypair_device(mode, seeds=seeds)
    sig = api.SignContext(mode, n, n).sign(mlist, sk)
    for i, b in zip(flip_idx, flip_bit):
        sig[i, b // 8] ^= 1 << (b % 8)
    d_sig, d_pk = DeviceBuffer.from_array(sig), DeviceBuffer.from_array(pk)
    dm = api.DeviceMessages(mlist)
    wrow = api.L().dil_verify_work_row_bytes(mode)
    d_work, d_ret = DeviceBuffer(n * wrow), DeviceBuffer(4 * n)
    graph = api.verify_graph(mode, d_ret, d_sig, d_pk, dm, d_work, wrow)
    rets = []
    for _ in range(2):
        d_ret.memset(0x55)
        graph.launch()
        synchronize()
        rets.append(d_ret.download((n,), np.int32))
    assert np.array_equal(rets[0], rets[1])
    ret = rets[0]
    oret = oracle.verify_batch(mode, sig, mlist, pk)
    assert np.array_equal(ret, oret)
    want = np.zeros(n, np.int32)
    want[flip_idx] = -1
    assert np.array_equal(ret, want)


def test_config3_dil3_sign_16384_tuned_graph_all_signatures():
    from cudilithium_amd import api, autotune, workloads
    from cudilithium_amd.device import DeviceBuffer, synchronize
    mode, n = 3, 16384
    seeds, msgs = workloads.config3(n)
    pkb, skb, sigb = api.sizes(mode)
    pk, sk = api.keypair_device(mode, seeds=seeds)
    epk, esk = oracle.keypair_batch(mode, seeds)
    assert np.array_equal(pk, epk) and np.array_equal(sk, esk)
    assert autotune.lookup(mode, n) is not None  # the shipped tiling for this bucket
    ctx = api.SignContext(mode, n)  # autotuned slots, tiling and claim
    d_sk = DeviceBuffer.from_array(sk)
    dm = api.DeviceMessages(msgs)
    d_sig = DeviceBuffer(n * sigb)
    graph = ctx.graph(d_sig, d_sk, dm)
    sigs = []
    for _ in range(2):
        d_sig.memset(0)
        graph.launch()
        synchronize()
        sigs.append(d_sig.download((n, sigb), np.uint8))
    assert ctx.status() == 0
    attempts = ctx.attempt_counts(n)
    assert np.array_equal(sigs[0], sigs[1])
    exp, exp_att = oracle.sign_batch(mode, msgs, sk)
    assert np.array_equal(sigs[0], exp)
    assert np.array_equal(attempts, exp_att)
