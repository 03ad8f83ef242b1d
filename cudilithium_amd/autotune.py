"""Autotuner of the batched kernels.

For a (mode, operation, batch-size bucket) it times the operation over its
candidate tilings on synthetic inputs of that shape and records the fastest
in cudilithium_amd/tuning.json:

  sign{m}:{b}     persistent signing engine: attempt slots (a multiple of the
                  batch), slots per workgroup, register budget (waves per
                  SIMD), items claimed per workgroup (all slots, or fewer so
                  that each item gets several nonces per round) and guided
                  claims near the end of the queue (on / off)
  keygen{m}:{b}   keys per workgroup of keygen's main kernel
  verify{m}:{b}   items per workgroup of verify's main kernel
  ntt:{b}         wavefronts per workgroup of the standalone NTT

Buckets are 1024, 4096, 16384 and 65536 items (a batch takes the smallest
bucket at least its size, the last one above).  SignContext looks its entry
up; the kernel tilings are handed to the library when it is loaded
(apply_cache, dil_set_kernel_tiling).  Results never depend on the tiling.

    python -m cudilithium_amd.autotune --op sign --mode 3 --batch 16384 [--write]
    python -m cudilithium_amd.autotune --op verify --mode 5 --batch 65536 --write
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
BUCKETS = (1024, 4096, 16384, 65536)
SLOT_FACTORS = (0.5, 1.0, 1.5, 2.0)
SLOTS_PER_BLOCK = (8, 16, 32)
OCCUPANCY = (2, 3, 4)
CLAIM_DIVISORS = (1, 2, 4)  # claim = slots per workgroup / divisor
TILE_KINDS = {"keygen": 0, "verify": 1, "ntt": 2}  # DIL_TILE_* of include/cudilithium_amd.h
TILE_VALUES = (2, 4, 8, 16)


def bucket(n: int) -> int:
    for b in BUCKETS:
        if n <= b:
            return b
    return BUCKETS[-1]


def bucket_index(b: int) -> int:
    return BUCKETS.index(bucket(b))


def load_cache(path: str | None = None) -> dict:
    path = path or CACHE
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def write_entry(key: str, value: dict, path: str | None = None) -> None:
    cache = load_cache(path)
    cache[key] = value
    with open(path or CACHE, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)


def lookup(mode: int, batch: int, path: str | None = None):
    """(slot_factor, slots_per_block, occupancy, claim, guided) recorded for
    signing in this mode and batch bucket, or None."""
    e = load_cache(path).get(f"sign{mode}:{bucket(batch)}")
    return (e["slot_factor"], e["slots_per_block"], e.get("occupancy", 0), e.get("claim", 0),
            e.get("guided", 0)) if e else None


def apply_cache(L, path: str | None = None) -> int:
    """Hands the recorded kernel tilings to the library; returns how many."""
    count = 0
    for key, e in load_cache(path).items():
        op, _, b = key.partition(":")
        name = op.rstrip("0123456789")
        if name not in TILE_KINDS or not b.isdigit():
            continue
        mode = int(op[len(name):] or 0)
        if L.dil_set_kernel_tiling(TILE_KINDS[name], mode, bucket_index(int(b)), int(e["value"])) == 0:
            count += 1
    return count


def _time(fn, reps: int, min_seconds: float = 0.05) -> float:
    """Mean time per call over at least `reps` calls and `min_seconds` (a
    ~50 us kernel timed over 3 calls is mostly launch and clock noise)."""
    from .device import synchronize
    fn()
    synchronize()
    t0 = time.perf_counter()
    fn()
    synchronize()
    one = max(time.perf_counter() - t0, 1e-6)
    reps = max(reps, min(1000, int(min_seconds / one) + 1))
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    synchronize()
    return (time.perf_counter() - t0) / reps


def tune_sign(mode: int, batch: int, reps: int = 3, factors=SLOT_FACTORS, spbs=SLOTS_PER_BLOCK, occs=OCCUPANCY,
              claims=CLAIM_DIVISORS, guided=(0, 1), log=print):
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
            for occ, div, g in ((o, d, g) for o in occs for d in claims for g in guided):
                claim = max(1, spb // div) if div > 1 else 0
                ctx.set_tiling(spb, occ)
                ctx.set_claim(claim | (api.CLAIM_GUIDED if g else 0))
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
                     "guided": g,
                     "blocks": nslots // spb, "batch_ms": wall * 1e3, "loop_ms": loop,
                     "executed_attempts": executed, "needed_attempts": needed,
                     "speculative_overhead": executed / max(1, needed) - 1}
                results.append(r)
                log(json.dumps(r))
        del ctx
    best = min(results, key=lambda r: r["batch_ms"])
    return best, results


def tune_kernel(op: str, mode: int, batch: int, reps: int = 3, values=TILE_VALUES, log=print):
    """Times keygen / verify / ntt over the tilings the kernels are built for
    (dil_set_kernel_tiling accepts them) at this batch's bucket.  Returns the
    best {"value", "batch_ms"} and all results; the library's setting for the
    bucket is left at the best value."""
    from . import api, workloads
    from .device import DeviceBuffer
    L = api.L()
    kind, bi = TILE_KINDS[op], bucket_index(batch)
    if op == "keygen":
        pkb, skb, _ = api.sizes(mode)
        seeds = DeviceBuffer.from_array(workloads.xoshiro_bytes(0xA072 + mode, batch * 32).reshape(batch, 32))
        wrow = (L.dil_keypair_work_row_bytes(mode) + 255) // 256 * 256
        d_pk, d_sk, d_work = DeviceBuffer(batch * pkb), DeviceBuffer(batch * skb), DeviceBuffer(batch * wrow)

        def run():
            api._lib.check(L.dil_keypair_device(mode, d_pk.ptr, pkb, d_sk.ptr, skb, seeds.ptr, 32, 0, batch,
                                                d_work.ptr, wrow, None), "keypair_device")
    elif op == "verify":
        pkb, _, sigb = api.sizes(mode)
        seeds = workloads.xoshiro_bytes(0xA073 + mode, batch * 32).reshape(batch, 32)
        msgs = [m.tobytes() for m in workloads.xoshiro_bytes(0xA074 + mode, batch * 1024).reshape(batch, 1024)]
        pk, sk = api.keypair_device(mode, seeds=seeds)
        sig = api.SignContext(mode, batch, batch).sign(msgs, sk)
        d_sig, d_pk, dm = DeviceBuffer.from_array(sig), DeviceBuffer.from_array(pk), api.DeviceMessages(msgs)
        wrow = L.dil_verify_work_row_bytes(mode)
        d_work, d_ret = DeviceBuffer(batch * wrow), DeviceBuffer(4 * batch)

        def run():
            api._lib.check(L.dil_verify_device(mode, d_ret.ptr, d_sig.ptr, sigb, d_pk.ptr, pkb, dm.data.ptr,
                                               dm.offs.ptr, dm.lens.ptr, batch, d_work.ptr, wrow, None), "verify")
    else:
        from . import ntt
        a, b = workloads.config2(batch)
        da, db, dc = DeviceBuffer.from_array(a), DeviceBuffer.from_array(b), DeviceBuffer(a.nbytes)

        def run():
            ntt.pipeline_device(dc, da, db, batch)
    results = []
    for v in values:
        if L.dil_set_kernel_tiling(kind, mode, bi, v) != 0:
            continue  # not built for this mode (LDS)
        ms = _time(run, reps) * 1e3
        results.append({"value": v, "batch_ms": ms})
        log(json.dumps({"op": op, "mode": mode, "batch": batch, "value": v, "batch_ms": ms}))
    best = min(results, key=lambda r: r["batch_ms"])
    api._lib.check(L.dil_set_kernel_tiling(kind, mode, bi, best["value"]), "set_kernel_tiling")
    return best, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--op", default="sign", choices=["sign", "keygen", "verify", "ntt"])
    ap.add_argument("--mode", type=int, default=3)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--occupancy", type=int, nargs="*", default=list(OCCUPANCY))
    ap.add_argument("--write", action="store_true", help="record the best tiling in tuning.json")
    ap.add_argument("--cache", default=None, help="tuning file to record in (default: the package's tuning.json)")
    args = ap.parse_args()
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    if args.op == "sign":
        best, _ = tune_sign(args.mode, args.batch, args.reps, occs=tuple(args.occupancy), log=log)
        entry = {"slot_factor": best["slot_factor"], "slots_per_block": best["slots_per_block"],
                 "occupancy": best["occupancy"], "claim": best["claim"], "guided": best["guided"],
                 "batch_ms": best["batch_ms"]}
        key = f"sign{args.mode}:{bucket(args.batch)}"
    else:
        best, _ = tune_kernel(args.op, args.mode, args.batch, args.reps, log=log)
        entry = {"value": best["value"], "batch_ms": best["batch_ms"]}
        key = f"{args.op}{args.mode if args.op != 'ntt' else ''}:{bucket(args.batch)}"
    print(json.dumps({"key": key, "best": entry}))
    if args.write:
        write_entry(key, entry, args.cache)


if __name__ == "__main__":
    main()
