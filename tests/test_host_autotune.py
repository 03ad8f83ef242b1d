"""Host logic of the autotuner (no GPU): batch buckets, the tuning file's
entries and how SignContext reads them, and that every recorded entry is one
the library accepts."""
import json

from cudilithium_amd import api, autotune


def test_buckets():
    assert [autotune.bucket(n) for n in (1, 1024, 1025, 4096, 10000, 16384, 16385, 10**6)] == \
        [1024, 1024, 4096, 4096, 16384, 16384, 65536, 65536]
    assert [autotune.bucket_index(n) for n in (1, 2000, 16384, 10**6)] == [0, 1, 2, 3]


def test_cache_roundtrip_and_lookup(tmp_path):
    path = str(tmp_path / "tuning.json")
    assert autotune.load_cache(path) == {}
    assert autotune.lookup(3, 16384, path) is None
    autotune.write_entry("sign3:16384", {"slot_factor": 1.0, "slots_per_block": 32, "occupancy": 2, "claim": 16,
                                         "guided": 1, "batch_ms": 3.3}, path)
    autotune.write_entry("verify5:65536", {"value": 2, "batch_ms": 4.8}, path)
    assert autotune.lookup(3, 10000, path) == (1.0, 32, 2, 16, 1)
    assert autotune.lookup(3, 1000, path) is None
    # entries written before the claim / guided fields existed read as 0
    autotune.write_entry("sign2:1024", {"slot_factor": 2.0, "slots_per_block": 16}, path)
    assert autotune.lookup(2, 1, path) == (2.0, 16, 0, 0, 0)
    with open(path) as f:
        assert set(json.load(f)) == {"sign3:16384", "verify5:65536", "sign2:1024"}


def test_shipped_tuning_file_is_complete_and_valid():
    """Every (op, mode, bucket) has an entry, and each holds values the
    library's setters accept (slot counts, register budgets, claims, kernel
    tilings)."""
    cache = autotune.load_cache()
    for m in (2, 3, 5):
        for b in autotune.BUCKETS:
            s = cache[f"sign{m}:{b}"]
            assert s["slot_factor"] in autotune.SLOT_FACTORS
            assert s["slots_per_block"] in autotune.SLOTS_PER_BLOCK
            assert s["occupancy"] in autotune.OCCUPANCY
            assert 0 <= s["claim"] <= s["slots_per_block"] and s["guided"] in (0, 1)
            for op in ("keygen", "verify"):
                assert cache[f"{op}{m}:{b}"]["value"] in autotune.TILE_VALUES
    for b in autotune.BUCKETS:
        assert cache[f"ntt:{b}"]["value"] in (2, 4, 8)


def test_claim_flag_matches_the_header():
    """The Python flag is the header's DIL_CLAIM_GUIDED (the C-ABI's check of
    claim arguments is exercised by the GPU tests)."""
    import pathlib
    import re
    hdr = (pathlib.Path(__file__).resolve().parents[1] / "include" / "cudilithium_amd.h").read_text()
    assert int(re.search(r"#define DIL_CLAIM_GUIDED (0x[0-9a-fA-F]+)", hdr).group(1), 16) == api.CLAIM_GUIDED
