"""GPU parity of keygen / sign / verify against the reference's KAT pins and
the oracle, through the C-ABI (drop-in and device-resident entry points)."""
import hashlib

import numpy as np
import pytest

import oracle
from tests import kat

pytestmark = pytest.mark.gpu


def _engines():
    from cudilithium_amd.api import ENGINE_PERSISTENT, ENGINE_PERSTAGE
    return [ENGINE_PERSISTENT, ENGINE_PERSTAGE]


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_gpu_reproduces_reference_tvecs(mode):
    """tests/test_vectors.cu on the GPU: 10,000 counter-seeded keys, 32-byte
    messages, exec_threshold 2048, verify; SHA-256 of the tvecs file must be
    the reference's pin (tests/CMakeLists.txt:21-23)."""
    from cudilithium_amd.api import Dilithium
    d = Dilithium(mode)
    m = kat.messages()
    pk, sk = d.keypair(kat.NVECTORS, rand_index=0)
    sig = d.sign(m, sk, exec_threshold=2048)
    ret = d.verify(sig, m, pk)
    assert (ret == 0).all()
    assert kat.tvecs_sha256(m, pk, sk, sig, ret) == kat.PINS[mode]


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_perstage_engine_matches_oracle(mode):
    from cudilithium_amd.api import ENGINE_PERSTAGE, Dilithium
    d = Dilithium(mode)
    n = 300
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=77)
    m = np.random.default_rng(mode).integers(0, 256, (n, 48), dtype=np.uint8)
    exp, _ = oracle.sign_batch(mode, m, sk)
    # exec_threshold 64 < batch exercises full rounds then tail rounds
    got = d.sign(m, sk, exec_threshold=64, engine=ENGINE_PERSTAGE)
    assert np.array_equal(got, exp)


def test_persistent_dropin_fewer_slots_than_items():
    """Drop-in sign with exec_threshold 8 on 300 Dilithium3 items: the
    persistent engine gets exec_threshold * temp_pitch / align256(sig) (79)
    slots, keeping slot state in sign-pool rows; bytes equal the oracle's."""
    from cudilithium_amd.api import Dilithium, ENGINE_PERSISTENT
    mode, n = 3, 300
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=4242)
    m = np.random.default_rng(4242).integers(0, 256, (n, 33), dtype=np.uint8)
    exp, _ = oracle.sign_batch(mode, m, sk)
    got = Dilithium(mode).sign(m, sk, exec_threshold=8, engine=ENGINE_PERSISTENT)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("engine", _engines())
def test_config1_dil2_batch128_explicit_seeds(engine):
    """Dilithium2, batch 128, 32-byte seeds and 64-byte messages from the
    seeded xoshiro256** stream: GPU bytes equal the oracle's, verdicts 0, for
    the persistent engine and the single-launch-per-stage form."""
    from cudilithium_amd.api import Dilithium, keypair_device
    from cudilithium_amd.workloads import config1
    seeds, m = config1()
    pk, sk = keypair_device(2, seeds=seeds)
    epk, esk = oracle.keypair_batch(2, seeds)
    assert np.array_equal(pk, epk) and np.array_equal(sk, esk)
    d = Dilithium(2)
    sig = d.sign(m, sk, engine=engine)
    esig, _ = oracle.sign_batch(2, m, sk)
    assert np.array_equal(sig, esig)
    assert (d.verify(sig, m, pk) == 0).all()


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_keypair_counter_and_seed_forms_agree(mode):
    from cudilithium_amd.api import Dilithium, keypair_device
    n = 200
    pk, sk = Dilithium(mode).keypair(n, rand_index=1234)
    pk2, sk2 = keypair_device(mode, n=n, rand_index=1234)
    epk, esk = oracle.keypair_counter_batch(mode, n, rand_index=1234)
    assert np.array_equal(pk, epk) and np.array_equal(sk, esk)
    assert np.array_equal(pk2, epk) and np.array_equal(sk2, esk)


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_keypair_device_unaligned_rows(mode):
    """Keygen into rows at odd addresses and odd pitches, with work rows at an
    8-byte (not 16-byte) pitch: the byte and 8-byte store paths of ExpandS
    and the packers give the same keys as packed, aligned rows."""
    from cudilithium_amd import _lib, api
    from cudilithium_amd.device import DeviceBuffer
    L = api.L()
    n = 45
    pkb, skb, _ = api.sizes(mode)
    seeds = np.frombuffer(hashlib.shake_128(b"unaligned rows %d" % mode).digest(32 * n), np.uint8).reshape(n, 32)
    epk, esk = oracle.keypair_batch(mode, seeds)
    pkp, skp, sp = pkb + 5, skb + 3, 33
    wrow = (L.dil_keypair_work_row_bytes(mode) + 7) // 8 * 8
    wrow += 8 if wrow % 16 == 0 else 0  # 8-aligned, not 16-aligned
    d_pk, d_sk = DeviceBuffer(n * pkp + 16), DeviceBuffer(n * skp + 16)
    d_seeds = DeviceBuffer(n * sp + 16)
    host_seeds = np.zeros(n * sp + 16, np.uint8)
    for i in range(n):
        host_seeds[1 + i * sp:1 + i * sp + 32] = seeds[i]
    d_seeds.upload(host_seeds)
    d_work = DeviceBuffer(n * wrow + 256)
    _lib.check(L.dil_keypair_device(mode, d_pk.ptr + 1, pkp, d_sk.ptr + 3, skp, d_seeds.ptr + 1, sp, 0, n,
                                    d_work.ptr + 8, wrow, None), "keypair_device")
    pk_all = d_pk.download((n * pkp + 16,), np.uint8)
    sk_all = d_sk.download((n * skp + 16,), np.uint8)
    for i in range(n):
        assert np.array_equal(pk_all[1 + i * pkp:1 + i * pkp + pkb], epk[i]), (mode, i)
        assert np.array_equal(sk_all[3 + i * skp:3 + i * skp + skb], esk[i]), (mode, i)


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_verify_bit_flips_match_oracle(mode):
    from cudilithium_amd.api import Dilithium
    n = 256
    rng = np.random.default_rng(100 + mode)
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=5)
    m = rng.integers(0, 256, (n, 40), dtype=np.uint8)
    sig, _ = oracle.sign_batch(mode, m, sk)
    flip = rng.random(n) < 0.5
    for i in np.nonzero(flip)[0]:
        bit = rng.integers(0, sig.shape[1] * 8)
        sig[i, bit // 8] ^= 1 << (bit % 8)
    ret = Dilithium(mode).verify(sig, m, pk)
    exp = oracle.verify_batch(mode, sig, m, pk)
    assert np.array_equal(ret, exp)
    assert (ret[~flip] == 0).all()
    assert (ret[flip] == -1).mean() > 0.95


def test_verify_wrong_siglen_rejects_all():
    from cudilithium_amd.api import Dilithium
    d = Dilithium(3)
    pk, sk = oracle.keypair_counter_batch(3, 8)
    m = np.zeros((8, 32), np.uint8)
    sig, _ = oracle.sign_batch(3, m, sk)
    assert (d.verify(sig, m, pk, siglen=d.sig_bytes - 1) == -1).all()


@pytest.mark.parametrize("mode", [2, 5])
def test_verify_malformed_hints(mode):
    """Non-increasing hint indices, a count above omega, a decreasing count
    and non-zero padding are all rejected, as the oracle does."""
    from cudilithium_amd.api import Dilithium
    p = oracle.params(mode)
    n = 16
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=9)
    m = np.arange(n * 16, dtype=np.uint8).reshape(n, 16)
    sig, _ = oracle.sign_batch(mode, m, sk)
    hoff = 32 + p.l * p.polyz_bytes
    bad = sig.copy()
    for i in range(n):
        h = bad[i, hoff:]
        kind = i % 4
        cnt = int(h[p.omega + p.k - 1])
        if kind == 0 and cnt >= 2:
            h[1] = h[0]  # not strictly increasing (first polynomial)
        elif kind == 1:
            h[p.omega] = p.omega + 1  # count beyond omega
        elif kind == 2 and cnt < p.omega:
            h[p.omega - 1] = 7  # padding not zero
        else:
            h[p.omega + 1] = 0 if h[p.omega] > 0 else 1  # decreasing / inconsistent counts
    ret = Dilithium(mode).verify(bad, m, pk)
    exp = oracle.verify_batch(mode, bad, m, pk)
    assert np.array_equal(ret, exp)


def test_verify_device_variable_message_lengths():
    from cudilithium_amd.api import verify_device
    mode = 5
    n = 64
    rng = np.random.default_rng(3)
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=11)
    msgs = [rng.integers(0, 256, int(rng.integers(0, 3000)), dtype=np.uint8).tobytes() for _ in range(n)]
    msgs[0] = b""
    sig, _ = oracle.sign_batch(mode, msgs, sk)
    assert (verify_device(mode, sig, msgs, pk) == 0).all()
    msgs2 = list(msgs)
    for i in range(0, n, 3):
        msgs2[i] = msgs2[i] + b"x"
    ret = verify_device(mode, sig, msgs2, pk)
    exp = oracle.verify_batch(mode, sig, msgs2, pk)
    assert np.array_equal(ret, exp)


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_persistent_engine_matches_oracle_and_perstage(mode):
    """Device-resident batch with per-item message lengths (0 B .. 5 KiB):
    the persistent engine, the per-stage engine and the oracle agree byte for
    byte, for several slot tilings (including ones that force speculative
    nonces and many queue refills)."""
    from cudilithium_amd.api import ENGINE_PERSISTENT, ENGINE_PERSTAGE, SignContext
    rng = np.random.default_rng(40 + mode)
    n = 333
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=321)
    msgs = [rng.integers(0, 256, int(rng.integers(0, 5000)), dtype=np.uint8).tobytes() for _ in range(n)]
    msgs[1] = b""
    exp, att = oracle.sign_batch(mode, msgs, sk)
    for nslots, spb, occ in [(64, 8, 2), (256, 16, 3), (1024, 32, 4), (7, 1, 2), (2048, 32, 2)]:
        ctx = SignContext(mode, n, nslots)
        ctx.set_tiling(spb, occ)
        got = ctx.sign(msgs, sk, ENGINE_PERSISTENT)
        assert np.array_equal(got, exp), (nslots, spb, occ)
    ctx = SignContext(mode, n, 100)
    assert np.array_equal(ctx.sign(msgs, sk, ENGINE_PERSTAGE), exp)
    assert (oracle.verify_batch(mode, exp, msgs, pk) == 0).all()


def test_persistent_engine_repeated_calls_and_small_batches():
    """The context is reusable: the queue head and slot state are reset per
    call; batches of 1, 2 and 65 items."""
    from cudilithium_amd.api import SignContext
    mode = 3
    pk, sk = oracle.keypair_counter_batch(mode, 65, rand_index=2)
    ctx = SignContext(mode, 65, 512)
    for n in (1, 2, 65, 2):
        msgs = [bytes([n, i]) * 20 for i in range(n)]
        exp, _ = oracle.sign_batch(mode, msgs, sk[:n])
        assert np.array_equal(ctx.sign(msgs, sk[:n]), exp)


def test_dropin_persistent_matches_perstage_full_rounds():
    from cudilithium_amd.api import ENGINE_PERSISTENT, ENGINE_PERSTAGE, Dilithium
    d = Dilithium(2)
    pk, sk = oracle.keypair_counter_batch(2, 500, rand_index=8)
    m = np.random.default_rng(8).integers(0, 256, (500, 64), dtype=np.uint8)
    a = d.sign(m, sk, exec_threshold=128, engine=ENGINE_PERSISTENT)
    b = d.sign(m, sk, exec_threshold=128, engine=ENGINE_PERSTAGE)
    assert np.array_equal(a, b)
    assert np.array_equal(a, oracle.sign_batch(2, m, sk)[0])


@pytest.mark.parametrize("claim", [1, 3, 8, 8 | 0x10000, 0x10000])
def test_persistent_engine_claim_cap(claim):
    """Fewer items per workgroup than slots (several nonces per item per
    round), and guided claims (0x10000, DIL_CLAIM_GUIDED: at most remaining
    items / resident workgroups near the end of the queue), change no
    signature or attempt count."""
    from cudilithium_amd import api
    mode, n = 3, 150
    _, sk = oracle.keypair_counter_batch(mode, n, rand_index=4000 + (claim & 0xFF))
    m = [bytes([claim & 0xFF]) * (i % 97) for i in range(n)]
    exp, exp_att = oracle.sign_batch(mode, m, sk)
    ctx = api.SignContext(mode, n, 256)
    ctx.set_tiling(8, 2)
    ctx.set_claim(claim)
    ctx.profile_phases(True)
    got = ctx.sign(m, sk)
    assert np.array_equal(got, exp)
    assert np.array_equal(ctx.attempt_counts(n), exp_att)
    assert sum(ctx.phase_cycles().values()) > 0


def test_verify_rejects_hint_padding_pair_like_specification():
    """Two non-zero hint padding bytes 32 positions apart: the reference's
    XOR-of-ballots decode (src/verify.cu:81-85) would accept; the engine
    follows the specification and rejects, as the oracle does."""
    from cudilithium_amd.api import Dilithium
    mode, n = 5, 128
    p = oracle.params(mode)
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=31)
    msgs = np.array([[i] * 40 for i in range(n)], dtype=np.uint8)
    sig, _ = oracle.sign_batch(mode, msgs, sk)
    hoff = p.sig_bytes - (p.omega + p.k)
    cand = [i for i in range(n) if int(sig[i, hoff + p.omega + p.k - 1]) + 32 < p.omega]
    assert cand
    bad = sig[cand].copy()
    for j, i in enumerate(cand):
        pos = int(sig[i, hoff + p.omega + p.k - 1])
        bad[j, hoff + pos] = 1
        bad[j, hoff + pos + 32] = 1
    d = Dilithium(mode)
    assert (d.verify(sig[cand], msgs[cand], pk[cand]) == 0).all()
    got = d.verify(bad, msgs[cand], pk[cand])
    assert (got == -1).all()
    assert np.array_equal(got, oracle.verify_batch(mode, bad, msgs[cand], pk[cand]))


def test_guided_claims_across_modes_and_sizes():
    """Guided claims at several batch sizes (one workgroup up to several
    resident generations) and both register budgets: bytes and attempt
    counts equal the oracle's; unknown claim flags are rejected."""
    from cudilithium_amd import _lib, api
    for mode, n, spb, occ in ((2, 1, 1, 2), (5, 37, 8, 2), (3, 600, 16, 2), (2, 600, 8, 3)):
        _, sk = oracle.keypair_counter_batch(mode, n, rand_index=4100 + n)
        m = [bytes([n & 0xFF]) * (i % 61) for i in range(n)]
        exp, exp_att = oracle.sign_batch(mode, m, sk)
        ctx = api.SignContext(mode, n, n)
        ctx.set_tiling(spb, occ)
        ctx.set_claim(2 | api.CLAIM_GUIDED)
        assert np.array_equal(ctx.sign(m, sk), exp), (mode, n)
        assert np.array_equal(ctx.attempt_counts(n), exp_att), (mode, n)
    with pytest.raises(_lib.DilError):
        ctx.set_claim(0x20000)
