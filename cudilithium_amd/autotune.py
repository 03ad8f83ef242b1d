"""Autotuner of the persistent signing engine.

For a (mode, batch-size bucket) it times the engine over candidate tilings --
attempt slots (a multiple of the batch), slots per workgroup, register budget
(waves per SIMD) and items claimed per workgroup (all slots, or fewer so that
each item gets several nonces per round) -- on synthetic inputs of that shape,
and records the fastest in
cudilithium_amd/tuning.json.  SignContext looks the bucket up when the caller
does not choose a tiling.

    python -m cudilithium_amd.autotune --mode 3 --batch 16384 [--write]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuning.json")
SLOT_FACTORS = (0.5, 1.0, 1.5, 2.0)
SLOTS_PER_BLOCK = (8, 16, 32)
OCCUPANCY = (2, 3, 4)
CLAIM_DIVISORS = (1, 2, 4)  # claim = slots per workgroup / divisor


def bucket(n: int) -> int:
    b = 1
    while b < n:
        b <<= 1
    return b


def load_cache() -> dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(mode: int, batch: int):
    """(slot_factor, slots_per_block, occupancy, claim) recorded for this mode
    and batch bucket, or None."""
    e = load_cache().get(f"sign{mode}:{bucket(batch)}")
    return (e["slot_factor"], e["slots_per_block"], e.get("occupancy", 0), e.get("claim", 0)) if e else None


def tune_sign(mode: int, batch: int, reps: int = 3, factors=SLOT_FACTORS, spbs=SLOTS_PER_BLOCK, occs=OCCUPANCY,
              claims=CLAIM_DIVISORS, log=print):
    from . import api, workloads
    from .device import DeviceBuffer, synchronize

    seeds = workloads.xoshiro_bytes(0xA070 + mode, batch * 32).reshape(batch, 32)
    msgs = workloads.messages_loguniform(0xA071 + mode, batch, 32, 4096)
    _, sk = api.keypair_device(mode, seeds=seeds)
    _, _, sigb = api.sizes(mode)
    d_sk, dm, d_sig = DeviceBuffer.from_array(sk), api.DeviceMessages(msgs), DeviceBuffer(batch * sigb)
    results = []
    for f in factors:
        nslots = max(1, int(batch * f))
        ctx = api.SignContext(mode, batch, nslots)
        for spb in spbs:
            if nslots // spb < 1:
                continue
            for occ, div in ((o, d) for o in occs for d in claims):
                claim = max(1, spb // div) if div > 1 else 0
                ctx.set_tiling(spb, occ)
                ctx.set_claim(claim)
                ctx.sign_device(d_sig, d_sk, dm)
                synchronize()
                ctx.enable_timing(reps)
                t0 = time.perf_counter()
                for _ in range(reps):
                    ctx.sign_device(d_sig, d_sk, dm)
                synchronize()
                wall = (time.perf_counter() - t0) / reps
                loop = float(np.mean([t[1] for t in ctx.timings()]))
                ctx.enable_timing(0)
                executed = ctx.executed_attempts()
                needed = int(ctx.attempt_counts(batch).sum())
                r = {"slot_factor": f, "nslots": nslots, "slots_per_block": spb, "occupancy": occ, "claim": claim,
                     "blocks": nslots // spb, "batch_ms": wall * 1e3, "loop_ms": loop,
                     "executed_attempts": executed, "needed_attempts": needed,
                     "speculative_overhead": executed / max(1, needed) - 1}
                results.append(r)
                log(json.dumps(r))
        del ctx
    best = min(results, key=lambda r: r["batch_ms"])
    return best, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", type=int, default=3)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--occupancy", type=int, nargs="*", default=list(OCCUPANCY))
    ap.add_argument("--write", action="store_true", help="record the best tiling in tuning.json")
    args = ap.parse_args()
    best, _ = tune_sign(args.mode, args.batch, args.reps, occs=tuple(args.occupancy),
                        log=lambda s: print(s, file=sys.stderr, flush=True))
    print(json.dumps({"best": best}))
    if args.write:
        cache = load_cache()
        cache[f"sign{args.mode}:{bucket(args.batch)}"] = {
            "slot_factor": best["slot_factor"], "slots_per_block": best["slots_per_block"],
            "occupancy": best["occupancy"], "claim": best["claim"], "batch_ms": best["batch_ms"]}
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
