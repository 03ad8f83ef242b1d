"""The autotuner on small grids at a small bucket: it records its choice in a
cache file, SignContext / the library then use that choice, and the bytes
and verdicts under the chosen tilings equal the oracle's."""
import json

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture
def cache(tmp_path, monkeypatch):
    from cudilithium_amd import autotune
    path = str(tmp_path / "tuning.json")
    monkeypatch.setattr(autotune, "CACHE", path)
    yield path
    # the kernel tilings are process-wide: back to the compiled defaults
    from cudilithium_amd import api
    for kind in autotune.TILE_KINDS.values():
        for mode in (2, 3, 5):
            api.L().dil_set_kernel_tiling(kind, mode, -1, 0)


def test_sign_tuner_choice_is_looked_up_and_exact(cache):
    from cudilithium_amd import api, autotune, workloads
    mode, n = 2, 700
    best, results = autotune.tune_sign(mode, n, reps=1, factors=(1.0, 2.0), spbs=(8,), occs=(2,), claims=(1, 2),
                                       guided=(0, 1), log=lambda s: None)
    assert len(results) == 8
    autotune.write_entry(f"sign{mode}:{autotune.bucket(n)}",
                         {"slot_factor": best["slot_factor"], "slots_per_block": 8, "occupancy": 2, "claim": 4,
                          "guided": 1, "batch_ms": best["batch_ms"]})
    assert json.load(open(cache))[f"sign{mode}:1024"]["slot_factor"] == best["slot_factor"]
    assert autotune.lookup(mode, n) == (best["slot_factor"], 8, 2, 4, 1)
    ctx = api.SignContext(mode, n)
    assert ctx.nslots == int(n * best["slot_factor"]) and ctx.slots_per_block == 8
    assert ctx.claim == 4 | api.CLAIM_GUIDED
    seeds = workloads.xoshiro_bytes(0x77, n * 32).reshape(n, 32)
    msgs = workloads.messages_loguniform(0x78, n, 32, 4096)
    pk, sk = oracle.keypair_batch(mode, seeds)
    exp, att = oracle.sign_batch(mode, msgs, sk)
    assert np.array_equal(ctx.sign(msgs, sk), exp)
    assert np.array_equal(ctx.attempt_counts(n), att)


@pytest.mark.parametrize("op,mode,values", [("keygen", 3, (2, 4)), ("verify", 5, (2, 8)), ("ntt", 0, (2, 8))])
def test_kernel_tuner_choice_is_applied_and_exact(cache, op, mode, values):
    from cudilithium_amd import api, autotune, ntt, workloads
    L = api.L()
    n = 600
    best, results = autotune.tune_kernel(op, mode, n, reps=1, values=values, log=lambda s: None)
    assert [r["value"] for r in results] == list(values)
    key = f"{op}{mode if op != 'ntt' else ''}:{autotune.bucket(n)}"
    autotune.write_entry(key, {"value": best["value"], "batch_ms": best["batch_ms"]})
    kind = autotune.TILE_KINDS[op]
    L.dil_set_kernel_tiling(kind, mode, -1, 0)
    assert autotune.apply_cache(L) == 1
    assert L.dil_get_kernel_tiling(kind, mode, n) == best["value"]
    assert L.dil_get_kernel_tiling(kind, mode, 20000) == 0  # other buckets keep the default
    for v in values:  # every tiling of the grid gives the oracle's bytes
        L.dil_set_kernel_tiling(kind, mode, -1, v)
        if op == "keygen":
            seeds = np.random.default_rng(v).integers(0, 256, (n, 32), dtype=np.uint8)
            pk, sk = api.keypair_device(mode, seeds=seeds)
            epk, esk = oracle.keypair_batch(mode, seeds)
            assert np.array_equal(pk, epk) and np.array_equal(sk, esk), v
        elif op == "verify":
            pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=v)
            m = np.random.default_rng(v).integers(0, 256, (n, 40), dtype=np.uint8)
            sig, _ = oracle.sign_batch(mode, m, sk)
            sig[::7, 9] ^= 4
            ret = api.Dilithium(mode).verify(sig, m, pk)
            assert np.array_equal(ret, oracle.verify_batch(mode, sig, m, pk)), v
        else:
            a, b = workloads.config2(n)
            assert np.array_equal(ntt.pipeline(a, b), oracle.ntt_pipeline(a, b)), v
    # an unsupported value is refused
    assert L.dil_set_kernel_tiling(kind, mode, -1, 3) != 0
