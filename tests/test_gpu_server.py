"""Serving loop (dil_server_*): a mixed stream of keygen / sign / verify
requests over resident keys, batched continuously, gives every request the
oracle's bytes and verdict, the same result as when it runs alone, and
sensible admission errors and timestamps."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

KEYS = 8


def _server(budget=384 << 20, lanes=1, max_batch=128, window=0, long_bytes=4096, long_lanes=0):
    from cudilithium_amd import api
    srv = api.Server(budget, lanes=lanes, max_batch=max_batch, max_keys=(KEYS, KEYS, KEYS), batch_window_us=window,
                     long_message_bytes=long_bytes, long_lanes=long_lanes)
    keys = {}
    for m in (2, 3, 5):
        seeds = np.random.default_rng(70 + m).integers(0, 256, (KEYS, 32), dtype=np.uint8)
        srv.stores[m].keygen(0, seeds)
        keys[m] = oracle.keypair_batch(m, seeds)
    return srv, keys


def _mix(n, keys, seed, max_mlen=3000):
    from cudilithium_amd import api
    rng = np.random.default_rng(seed)
    ops = rng.choice([api.OP_KEYGEN, api.OP_SIGN, api.OP_VERIFY], size=n, p=[0.2, 0.5, 0.3])
    modes = rng.choice([2, 3, 5], size=n)
    kidx = rng.integers(0, KEYS, n)
    msgs = [rng.integers(0, 256, int(rng.integers(0, max_mlen + 1)), dtype=np.uint8).tobytes() for _ in range(n)]
    seeds = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    sigs = [None] * n
    flipped = np.zeros(n, bool)
    for m in (2, 3, 5):
        sel = np.nonzero((ops == api.OP_VERIFY) & (modes == m))[0]
        if len(sel):
            s, _ = oracle.sign_batch(m, [msgs[i] for i in sel], keys[m][1][kidx[sel]])
            for j, i in enumerate(sel):
                sg = s[j].copy()
                if rng.random() < 0.3:
                    b = int(rng.integers(0, sg.size * 8))
                    sg[b // 8] ^= 1 << (b % 8)
                    flipped[i] = True
                sigs[i] = sg
    return ops, modes, kidx, msgs, seeds, sigs, flipped


def _expected(ops, modes, kidx, msgs, seeds, sigs, keys):
    from cudilithium_amd import api
    exp = [None] * len(ops)
    for m in (2, 3, 5):
        for op in (api.OP_KEYGEN, api.OP_SIGN, api.OP_VERIFY):
            sel = np.nonzero((ops == op) & (modes == m))[0]
            if not len(sel):
                continue
            if op == api.OP_KEYGEN:
                pk, sk = oracle.keypair_batch(m, seeds[sel])
                for j, i in enumerate(sel):
                    exp[i] = np.concatenate([pk[j], sk[j]])
            elif op == api.OP_SIGN:
                s, _ = oracle.sign_batch(m, [msgs[i] for i in sel], keys[m][1][kidx[sel]])
                for j, i in enumerate(sel):
                    exp[i] = s[j]
            else:
                v = oracle.verify_batch(m, np.stack([sigs[i] for i in sel]), [msgs[i] for i in sel],
                                        keys[m][0][kidx[sel]])
                for j, i in enumerate(sel):
                    exp[i] = int(v[j])
    return exp


def _check(batch, ops, exp):
    from cudilithium_amd import api
    for i, r in enumerate(batch.reqs):
        assert r.done == 1 and r.status == 0, (i, r.status)
        if ops[i] == api.OP_VERIFY:
            assert r.verdict == exp[i], i
        else:
            assert np.array_equal(batch.out[i], exp[i]), i
        assert r.t_submit_ns <= r.t_launch_ns <= r.t_done_ns


def test_server_mixed_requests_match_oracle_and_alone_runs():
    from cudilithium_amd import api
    srv, keys = _server()
    n = 400
    ops, modes, kidx, msgs, seeds, sigs, flipped = _mix(n, keys, 1)
    exp = _expected(ops, modes, kidx, msgs, seeds, sigs, keys)
    assert all(exp[i] == -1 for i in np.nonzero(flipped)[0])
    batch = api.RequestBatch(ops, modes, kidx, msgs, sigs, seeds)
    srv.submit(batch.reqs)
    srv.drain()
    _check(batch, ops, exp)
    st = srv.stats()
    assert st.completed == n and st.rejected == 0 and st.batches >= 9  # one batch per (op, mode) group at least
    assert list(st.items_by_op) == [int((ops == o).sum()) for o in range(3)]
    # isolation: a few of the same requests, each submitted alone
    for i in list(range(0, n, 37)):
        one = api.RequestBatch(ops[i:i + 1], modes[i:i + 1], kidx[i:i + 1], msgs[i:i + 1], sigs[i:i + 1],
                               seeds[i:i + 1])
        srv.submit(one.reqs)
        assert srv.wait(one.reqs[0], 30_000_000) == 0
        if ops[i] == api.OP_VERIFY:
            assert one.reqs[0].verdict == batch.reqs[i].verdict
        else:
            assert np.array_equal(one.out[0], batch.out[i])
    srv.close()


def test_server_replay_with_arrival_times_and_window():
    from cudilithium_amd import api
    srv, keys = _server(lanes=2, max_batch=64, window=200)
    n = 240
    ops, modes, kidx, msgs, seeds, sigs, _ = _mix(n, keys, 2, max_mlen=600)
    exp = _expected(ops, modes, kidx, msgs, seeds, sigs, keys)
    batch = api.RequestBatch(ops, modes, kidx, msgs, sigs, seeds)
    gaps = np.random.default_rng(3).exponential(20_000, n)  # 20 us mean gap
    srv.replay(batch.reqs, np.cumsum(gaps).astype(np.uint64))
    srv.drain()
    _check(batch, ops, exp)
    srv.close()


def test_server_admission_errors():
    from cudilithium_amd import api
    srv, keys = _server(budget=160 << 20, lanes=1, max_batch=64)
    cap = srv.stats().max_msg_bytes
    assert 0 < cap < (160 << 20)
    big = bytes(cap + 64)
    ops = np.array([api.OP_SIGN, api.OP_SIGN, api.OP_VERIFY, api.OP_SIGN])
    modes = np.array([3, 3, 5, 2])
    kidx = np.array([0, KEYS + 1, 0, 1])  # second names a key past the store
    msgs = [big, b"abc", b"abc", b"ok"]
    sig5 = np.zeros(4595, np.uint8)
    batch = api.RequestBatch(ops, modes, kidx, msgs, [None, None, sig5, None], np.zeros((4, 32), np.uint8))
    srv.submit(batch.reqs)
    srv.drain()
    st = [r.status for r in batch.reqs]
    assert st[0] == -3 and st[1] == -1 and st[2] == 0 and st[3] == 0
    assert batch.reqs[2].verdict == -1
    exp, _ = oracle.sign_batch(2, [b"ok"], keys[2][1][1:2])
    assert np.array_equal(batch.out[3], exp[0])

    # a sign request under a verify-only key is refused at admission; verify
    # under that key is served
    vpk, vsk = oracle.keypair_counter_batch(5, 1, rand_index=77)
    srv.stores[5].load(KEYS - 1, vpk)
    vsig, _ = oracle.sign_batch(5, [b"hello"], vsk)
    batch = api.RequestBatch(np.array([api.OP_SIGN, api.OP_VERIFY, api.OP_SIGN]), np.array([5, 5, 5]),
                             np.array([KEYS - 1, KEYS - 1, 0]), [b"hello", b"hello", b"hello"],
                             [None, vsig[0], None], np.zeros((3, 32), np.uint8))
    srv.submit(batch.reqs)
    srv.drain()
    assert [r.status for r in batch.reqs] == [-1, 0, 0]
    assert batch.reqs[1].verdict == 0
    exp, _ = oracle.sign_batch(5, [b"hello"], keys[5][1][0:1])
    assert np.array_equal(batch.out[2], exp[0])
    srv.close()


@pytest.mark.parametrize("long_bytes,long_lanes", [(512, 0), (512, 2), (0, 0)])
def test_server_message_classes(long_bytes, long_lanes):
    """Messages above long_message_bytes go to lanes of their own (512: most
    of this mix; one or two lanes per long-message group), or every message
    shares one lane per group (0); results match the oracle either way."""
    from cudilithium_amd import api
    srv, keys = _server(budget=256 << 20, lanes=1, max_batch=64, long_bytes=long_bytes, long_lanes=long_lanes)
    n = 160
    ops, modes, kidx, msgs, seeds, sigs, _ = _mix(n, keys, 11 + long_bytes, max_mlen=3000)
    exp = _expected(ops, modes, kidx, msgs, seeds, sigs, keys)
    batch = api.RequestBatch(ops, modes, kidx, msgs, sigs, seeds)
    srv.submit(batch.reqs)
    srv.drain()
    _check(batch, ops, exp)
    st = srv.stats()
    assert st.completed == n and st.rejected == 0
    srv.close()
