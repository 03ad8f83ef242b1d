"""Benchmark of the batched Dilithium engine on one MI355X.

Headline workload (default): Dilithium3 deterministic signing of 16384
(key, message) pairs, message lengths log-uniform in [32 B, 4 KiB], keys from
seeds, inputs resident in HBM, one fixed-stage pass plus one persistent
rejection-loop launch per step.  A step signs the whole batch.

Printed on one JSON line: throughput (signatures/s, the persistent engine
replayed as a HIP graph), the roofline of the dominant kernel (the persistent
rejection loop, integer-VALU bound), the oracle's CPU rate on a bounded
sample, the per-stage (reference architecture) GPU form on the same inputs,
a serving summary (resident keys, continuous batching: p50 / p99 latency
at two offered loads) and the drop-in call's rate with host buffers.

    python bench.py [--steps K] [--warmup W]
                    [--workload sign3|sign2|sign5|verify5|ntt|copy|serve|dropin|refcond|keccak]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


METRIC = "signs/s & verifies/s per MI355X (Dil2/3/5), bit-exact; serving p50/p99 latency"
PEAK_INT32_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # CUs x SIMDs x lanes x GHz = 39.3 T int32 ops/s
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by a 1 GiB copy's readings (read = 2 x FETCH_SIZE,
# written = WRITE_SIZE); profiles/sign3_16384_hbm_valu_counters.md.
TRAFFIC_BYTES = {
    # loop kernel per launch: 2 x FETCH_SIZE + WRITE_SIZE KiB (profiles/sign3_16384_hbm_valu_counters.md)
    "dilithium3_sign_16384_msg_loguniform_32B-4KiB": (2 * 1641558 + 1548709) * 1024,
    # vf_mu_pair + vf_prep + vf_main + vf_final per call (profiles/verify5_65536_kernel_stats.md)
    "dilithium5_verify_65536_msg_1KiB_10pct_bitflips":
        (2 * (119517 + 51717 + 247489 + 46140) + (4096 + 75070 + 65462 + 256)) * 1024,
    "ntt_pipeline_65536x256": (2 * 65576 + 65536) * 1024,
    # vf_main alone
    "dilithium5_verify_65536_msg_1KiB_10pct_bitflips_vf_main": (2 * 247489 + 65462) * 1024,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def bench_sign(args, mode: int):
    from cudilithium_amd import api, workmodel, workloads
    from cudilithium_amd.device import DeviceBuffer, synchronize

    n = args.batch or 16384
    seeds, msgs = workloads.config3(n) if mode == 3 else (
        workloads.xoshiro_bytes(workloads.CONFIG_SEED[3] + mode, n * 32).reshape(n, 32),
        workloads.messages_loguniform(workloads.CONFIG_SEED[3] + 1000 + mode, n, 32, 4096))
    pkb, skb, sigb = api.sizes(mode)
    log(f"[bench] setup: mode {mode}, {n} keys from seeds, messages {sum(map(len, msgs))} bytes")
    pk, sk = api.keypair_device(mode, seeds=seeds)
    d_sk = DeviceBuffer.from_array(sk)
    d_msgs = api.DeviceMessages(msgs)
    d_sig = DeviceBuffer(n * sigb)
    ctx = api.SignContext(mode, n, args.slots or None)  # autotuned tiling unless overridden
    nslots = ctx.nslots
    if args.spb or args.occupancy:
        ctx.set_tiling(args.spb, args.occupancy)
    if args.claim:
        ctx.set_claim(args.claim)
    if args.claim_guided:
        ctx.set_claim(ctx.claim | api.CLAIM_GUIDED)

    for _ in range(args.warmup):
        ctx.sign_device(d_sig, d_sk, d_msgs, api.ENGINE_PERSISTENT)
    synchronize()
    ctx.enable_timing(args.steps)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.sign_device(d_sig, d_sk, d_msgs, api.ENGINE_PERSISTENT)
    synchronize()
    wall = time.perf_counter() - t0
    tim = ctx.timings()
    ctx.enable_timing(0)
    sig_eager = d_sig.download((n, sigb), np.uint8)
    log(f"[bench] persistent: {args.steps} steps in {wall * 1e3:.2f} ms; loop kernel "
        f"{np.mean([t[1] for t in tim]):.3f} ms, fixed stage {np.mean([t[0] for t in tim]):.3f} ms")
    eager_value = n * args.steps / wall

    # the same call captured as a HIP graph (fixed stage's three branches +
    # the persistent loop), replayed once per step: the headline figure
    graph = ctx.graph(d_sig, d_sk, d_msgs)
    d_sig.memset(0)
    for _ in range(max(1, args.warmup)):
        graph.launch()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.launch()
    synchronize()
    wall = time.perf_counter() - t0
    sig = d_sig.download((n, sigb), np.uint8)
    attempts = ctx.attempt_counts(n)
    graph_matches_eager = bool(np.array_equal(sig, sig_eager))
    log(f"[bench] graph replay: {args.steps} steps in {wall * 1e3:.2f} ms; bytes equal eager: {graph_matches_eager}")

    # phase profile of one more call, outside the timed region
    ctx.profile_phases(True)
    ctx.sign_device(d_sig, d_sk, d_msgs, api.ENGINE_PERSISTENT)
    phases = ctx.phase_cycles()
    executed = ctx.executed_attempts()
    ctx.profile_phases(False)
    tot = sum(phases.values()) or 1
    phase_share = {k: round(v / tot, 4) for k, v in phases.items()}
    log(f"[bench] loop phase shares (workgroup cycles): {phase_share}; executed attempts "
        f"{ctx.executed_attempts()}")

    # parity on a sample (outside the timed region)
    import oracle
    rng = np.random.default_rng(1)
    idx = np.sort(rng.choice(n, size=min(n, args.check), replace=False))
    exp, exp_att = oracle.sign_batch(mode, [msgs[i] for i in idx], sk[idx])
    parity = bool(np.array_equal(sig[idx], exp)) and bool(np.array_equal(attempts[idx], exp_att)) \
        and graph_matches_eager
    log(f"[bench] parity vs oracle on {len(idx)} items: {parity}")

    # roofline of the rejection loop: algorithmic int32 ops of the attempts the
    # specification makes, over the loop kernel's event-timed duration
    loop_ms = float(np.mean([t[1] for t in tim]))
    ops = float(attempts.sum()) * workmodel.sign_attempt_ops(mode)
    achieved = ops / (loop_ms * 1e-3) / 1e12
    value = n * args.steps / wall

    perstage = None
    if not args.no_perstage:
        # the reference architecture at its default exec_threshold (2048,
        # benches/bench_dilithium.cu:278-280), and with one slot per item
        for thr in (2048, n):
            ps = api.SignContext(mode, n, thr)
            ps.sign_device(d_sig, d_sk, d_msgs, api.ENGINE_PERSTAGE)
            synchronize()
            reps = max(1, min(args.steps, 3))
            t0 = time.perf_counter()
            for _ in range(reps):
                ps.sign_device(d_sig, d_sk, d_msgs, api.ENGINE_PERSTAGE)
            synchronize()
            ps_wall = (time.perf_counter() - t0) / reps
            ps_sig = d_sig.download((n, sigb), np.uint8)
            del ps
            r = {"value": n / ps_wall, "unit": "signs/s", "exec_threshold": thr,
                 "ms_per_step": ps_wall * 1e3, "bit_exact_vs_persistent": bool(np.array_equal(ps_sig, sig))}
            log(f"[bench] per-stage form, exec_threshold {thr}: {ps_wall * 1e3:.2f} ms per batch")
            if perstage is None:
                perstage = r
            else:
                perstage["at_exec_threshold_batch"] = r

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_sign(mode, msgs, sk, args.cpu_seconds)

    cfg = {"workload": f"dilithium{mode}_sign_{n}_msg_loguniform_32B-4KiB",
           "batch": n, "mode": mode, "engine": "persistent+graph", "slots": nslots,
           "slots_per_block": ctx.slots_per_block or "auto", "occupancy": ctx.occupancy or "auto",
           "claim": (ctx.claim & 0xFFFF) or "slots", "guided_claims": bool(ctx.claim & api.CLAIM_GUIDED),
           "mean_attempts": float(attempts.mean()), "loop_phase_shares": phase_share}
    return {
        "metric": METRIC, "value": value, "unit": "signs/s", "ms_per_step": wall / args.steps * 1e3,
        "dtype": "int32", "config": cfg, "parity": "bit-exact" if parity else "MISMATCH",
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_INT32_TOPS, "unit": "TOP/s",
                     "frac": achieved / PEAK_INT32_TOPS, "traffic": TRAFFIC_BYTES.get(cfg["workload"]),
                     "kernel": "sg_persistent_kernel", "kernel_ms": loop_ms,
                     "algorithmic_ops": ops},
        "cpu_baseline": cpu, "perstage_gpu_baseline": perstage, "eager_persistent_value": eager_value,
        "speedup_vs_perstage": (value / perstage["value"]) if perstage else None,
        "speedup_vs_perstage_at_exec_threshold_batch":
            (value / perstage["at_exec_threshold_batch"]["value"]) if perstage else None,
        "step": {"fixed_ms": float(np.mean([t[0] for t in tim])), "loop_ms": loop_ms,
                 "graph_ms_per_step": wall / args.steps * 1e3,
                 "whole_step_frac": (ops + n * float(np.mean([workmodel.sign_fixed_ops(mode, len(m)) for m in msgs])))
                 / (wall / args.steps) / 1e12 / PEAK_INT32_TOPS,
                 "executed_attempts": int(executed), "needed_attempts": int(attempts.sum()),
                 "speculation_ratio": executed / float(attempts.sum())},
    }


def bench_dropin(args, mode: int = 3):
    """The reference's call surface with host buffers (api.cu:99-447,
    crypto_sign_signature): sk, messages and signatures cross PCIe inside
    every call, so this is the PCIe-inclusive rate, never the headline."""
    from cudilithium_amd import api, workloads
    from cudilithium_amd.device import HostBuffer

    n, mlen = args.batch or 16384, 64
    seeds, _ = workloads.config3(n) if mode == 3 else (
        workloads.xoshiro_bytes(workloads.CONFIG_SEED[3] + mode, n * 32).reshape(n, 32), None)
    pkb, skb, sigb = api.sizes(mode)
    pk, sk = api.keypair_device(mode, seeds=seeds)
    pk, sk = np.ascontiguousarray(pk), np.ascontiguousarray(sk)
    m = workloads.xoshiro_bytes(workloads.CONFIG_SEED[3] + 77, n * mlen).reshape(n, mlen)
    pool = api.reference_pool(mode, api.POOL_SIGN, n, mlen)
    hsk, hm, hsig = HostBuffer((n, skb), np.uint8), HostBuffer((n, mlen), np.uint8), HostBuffer((n, sigb), np.uint8)
    hsk.array[...] = sk
    hm.array[...] = m
    out, sigs = {}, {}
    reps = max(1, min(args.steps, 5))
    for host in ("pinned", "pageable"):
        for thr in (n, 2048):
            if host == "pageable" and thr != n:
                continue
            temp = api.reference_pool(mode, api.POOL_SIGN_TEMP, thr)
            lut = api.TaskLut(n, thr)
            s_, m_, sig = (hsk.array, hm.array, hsig.array) if host == "pinned" else (sk, m, np.zeros((n, sigb), np.uint8))

            def call():
                api.crypto_sign_signature(mode, sig, sigb, m_, mlen, mlen, s_, pool, pool.pitch, temp, temp.pitch,
                                          lut, thr, n)
                api.stream_sync(None)
            call()
            t0 = time.perf_counter()
            for _ in range(reps):
                call()
            t = (time.perf_counter() - t0) / reps
            key = f"{host}_exec_threshold_{thr}"
            out[key] = {"value": n / t, "ms_per_call": t * 1e3}
            sigs[key] = sig.copy()
            log(f"[bench] drop-in {key}: {t * 1e3:.2f} ms per call")
    import oracle
    idx = np.arange(0, n, max(1, n // min(n, args.check)))[:args.check]
    exp, _ = oracle.sign_batch(mode, [bytes(m[i]) for i in idx], sk[idx])
    first = next(iter(sigs.values()))
    parity = all(np.array_equal(s, first) for s in sigs.values()) and bool(np.array_equal(first[idx], exp))
    # crypto_sign_verify (api.cu:455-515) on the signatures, one flipped bit in every 10th
    bad = first.copy()
    bad[::10, 7] ^= 1
    want = np.zeros(n, np.int32)
    want[::10] = -1
    vpool = api.reference_pool(mode, api.POOL_VERIFY, n, mlen)
    hpk, hbad, hret = HostBuffer((n, pkb), np.uint8), HostBuffer((n, sigb), np.uint8), HostBuffer((n,), np.int32)
    hpk.array[...] = pk
    hbad.array[...] = bad
    for host in ("pinned", "pageable"):
        p_, s_, m_, ret = (hpk.array, hbad.array, hm.array, hret.array) if host == "pinned" else (
            pk, bad, m, np.zeros(n, np.int32))

        def vcall():
            api.crypto_sign_verify(mode, ret, s_, sigb, sigb, m_, mlen, mlen, p_, vpool, vpool.pitch, n)
            api.stream_sync(None)
        vcall()
        t0 = time.perf_counter()
        for _ in range(reps):
            vcall()
        t = (time.perf_counter() - t0) / reps
        out[f"verify_{host}"] = {"value": n / t, "unit": "verifies/s", "ms_per_call": t * 1e3}
        parity = parity and bool(np.array_equal(ret, want))
        log(f"[bench] drop-in verify {host}: {t * 1e3:.2f} ms per call")
    log(f"[bench] drop-in parity (sign paths equal, {len(idx)} items vs oracle; verdicts): {parity}")
    best = max((r for k, r in out.items() if not k.startswith("verify")), key=lambda r: r["value"])
    return {"metric": METRIC, "value": best["value"], "unit": "signs/s", "ms_per_step": best["ms_per_call"],
            "dtype": "int32", "parity": "bit-exact" if parity else "MISMATCH",
            "config": {"workload": f"dilithium{mode}_dropin_sign_{n}_msg_64B_host_buffers", "batch": n,
                       "mode": mode, "calls": out}}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_sign(mode, msgs, sk, seconds, runs: int = 5):
    """The oracle (OpenMP over items, all host threads) on the first items of
    the same workload: one warm-up run, then the median of `runs` timed runs
    (std::chrono-style wall clock around the whole batch), the sample sized
    so that all runs together take about `seconds`."""
    import oracle
    threads = oracle.max_threads()
    t0 = time.perf_counter()
    oracle.sign_batch(mode, msgs[:64], sk[:64])
    per = (time.perf_counter() - t0) / 64
    m = int(max(64, min(len(msgs), seconds / (runs + 1) / max(per, 1e-6))))
    oracle.sign_batch(mode, msgs[:m], sk[:m])  # warm-up
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle.sign_batch(mode, msgs[:m], sk[:m])
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    log(f"[bench] CPU oracle: {m} signatures, median of {runs} runs {dt * 1e3:.0f} ms "
        f"(runs {', '.join(f'{t * 1e3:.0f}' for t in times)} ms) on {threads} threads of {cpu_model()}")
    return {"value": m / dt, "unit": "signs/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "build": oracle.BUILD_FLAGS,
            "sample": f"first {m} of the batch's (key, message) pairs, OpenMP schedule(dynamic); "
                      f"1 warm-up then the median of {runs} runs"}


def bench_verify(args, mode: int = 5):
    """Dilithium5 verify of 65,536 engine-signed triples, 1 KiB messages,
    6554 single-bit flips (config 4), replayed as a HIP graph.  Parity: every
    verdict against the oracle.  Roofline: the vf_main kernel (ExpandA folded
    into A z, the transforms and UseHint), timed by HIP events on its stream."""
    from cudilithium_amd import api, workmodel, workloads
    from cudilithium_amd.device import DeviceBuffer, synchronize
    import ctypes

    n = args.batch or 65536
    seeds, msgs, flip_idx, flip_bit = workloads.config4(n)
    pkb, skb, sigb = api.sizes(mode)
    mlist = [m.tobytes() for m in msgs]
    pk, sk = api.keypair_device(mode, seeds=seeds)
    ctx = api.SignContext(mode, n, n)
    sig = ctx.sign(mlist, sk)
    del ctx
    for i, b in zip(flip_idx, flip_bit):
        sig[i, b // 8] ^= 1 << (b % 8)
    d_sig, d_pk = DeviceBuffer.from_array(sig), DeviceBuffer.from_array(pk)
    dm = api.DeviceMessages(mlist)
    wrow = api.L().dil_verify_work_row_bytes(mode)
    d_work, d_ret = DeviceBuffer(n * wrow), DeviceBuffer(4 * n)
    L = api.L()

    def run():
        api._lib.check(L.dil_verify_device(mode, d_ret.ptr, d_sig.ptr, sigb, d_pk.ptr, pkb, dm.data.ptr, dm.offs.ptr,
                                           dm.lens.ptr, n, d_work.ptr, wrow, None), "verify")
    for _ in range(args.warmup):
        run()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
    synchronize()
    eager_wall = time.perf_counter() - t0
    ret_eager = d_ret.download((n,), np.int32)
    # per-kernel durations (HIP events between the three launches)
    L.dil_verify_device_timed.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 2 + [ctypes.c_size_t, ctypes.c_void_p,
                                          ctypes.c_size_t] + [ctypes.c_void_p] * 3 + [ctypes.c_size_t,
                                          ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    ms3 = np.zeros((max(1, args.steps), 3), np.float32)
    for k in range(ms3.shape[0]):
        api._lib.check(L.dil_verify_device_timed(mode, d_ret.ptr, d_sig.ptr, sigb, d_pk.ptr, pkb, dm.data.ptr,
                                                 dm.offs.ptr, dm.lens.ptr, n, d_work.ptr, wrow,
                                                 ms3[k].ctypes.data, None), "verify_timed")
    kms = ms3.mean(axis=0)
    # the three launches captured as a HIP graph, replayed once per step
    graph = api.verify_graph(mode, d_ret, d_sig, d_pk, dm, d_work, wrow)
    d_ret.memset(0x55)
    graph.launch()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.launch()
    synchronize()
    wall = time.perf_counter() - t0
    ret = d_ret.download((n,), np.int32)
    want = np.zeros(n, np.int32)
    want[flip_idx] = -1
    import oracle
    t0 = time.perf_counter()
    oret = oracle.verify_batch(mode, sig, mlist, pk)
    log(f"[bench] verify5: oracle verdicts for all {n} items in {time.perf_counter() - t0:.1f} s")
    parity = bool(np.array_equal(ret, oret)) and bool(np.array_equal(ret, want)) and bool(np.array_equal(ret, ret_eager))
    value = n * args.steps / wall
    kops = {k: n * v for k, v in workmodel.verify_kernel_ops(mode, 1024).items()}
    main_ms = float(kms[1])
    achieved = kops["vf_main"] / (main_ms * 1e-3) / 1e12
    pipe = sum(kops.values()) / (wall / args.steps) / 1e12
    log(f"[bench] verify5: graph {wall / args.steps * 1e3:.3f} ms per batch; kernels vf_prep / vf_main / vf_final "
        f"{kms[0]:.3f} / {kms[1]:.3f} / {kms[2]:.3f} ms")
    return {"metric": METRIC, "value": value, "unit": "verifies/s", "ms_per_step": wall / args.steps * 1e3,
            "dtype": "int32", "parity": "bit-exact" if parity else "MISMATCH",
            "config": {"workload": f"dilithium{mode}_verify_{n}_msg_1KiB_10pct_bitflips", "batch": n, "mode": mode,
                       "engine": "graph", "rejected": int((ret != 0).sum()), "flipped": int(len(flip_idx))},
            "eager_value": n * args.steps / eager_wall,
            "kernel_ms": {"vf_prep": float(kms[0]), "vf_main": float(kms[1]), "vf_final": float(kms[2])},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_INT32_TOPS, "unit": "TOP/s",
                         "frac": achieved / PEAK_INT32_TOPS,
                         "traffic": TRAFFIC_BYTES.get(f"dilithium{mode}_verify_{n}_msg_1KiB_10pct_bitflips_vf_main"),
                         "kernel": "vf_main_kernel", "kernel_ms": main_ms, "algorithmic_ops": kops["vf_main"],
                         "pipeline_frac": pipe / PEAK_INT32_TOPS},
            "cpu_baseline": None}


# The reference's published Dilithium2 rates (batch 10,000, MLEN 32, 1 stream;
# results/dili2_batch.csv:101,201,301) and its best stream-mode rates
# (results/dili2_stream.csv:13,20,43), on unstated (likely A100) hardware:
# context for the like-for-like line below, not a target.
REF_PUBLISHED = {"batch": {"keypair": 907721.0, "sign": 468848.0, "verify": 927505.0},
                 "stream": {"keypair": 1460496.0, "sign": 565143.0, "verify": 1407193.0}}


def bench_reference_conditions(args, n_streams: int = 10):
    """Dilithium2 under the reference's benchmark conditions
    (benches/bench_dilithium.cu:30-175): batch 10,000, 32-byte messages,
    pinned host buffers (cudaMallocHost), exec_threshold 2048, the drop-in
    calls with host buffers, each op timed as wall time around the call plus a
    device synchronisation; 1 untimed trial, then the median of 10.  Stream
    mode as the reference's (:180-249): 10 streams on contiguous slices of
    the batch, the pools and the task tables, sign issued from one host
    thread per stream."""
    from concurrent.futures import ThreadPoolExecutor
    from cudilithium_amd import api, workloads
    from cudilithium_amd.device import HostBuffer, Stream, synchronize
    mode, n, mlen, thr = 2, 10000, 32, 2048
    pkb, skb, sigb = api.sizes(mode)
    kpool = api.reference_pool(mode, api.POOL_KEYPAIR, n)
    spool = api.reference_pool(mode, api.POOL_SIGN, n, mlen)
    tpool = api.reference_pool(mode, api.POOL_SIGN_TEMP, thr)
    vpool = api.reference_pool(mode, api.POOL_VERIFY, n, mlen)
    lut = api.TaskLut(n, thr)
    hpk, hsk = HostBuffer((n, pkb), np.uint8), HostBuffer((n, skb), np.uint8)
    hm, hsm, hret = HostBuffer((n, mlen), np.uint8), HostBuffer((n, sigb), np.uint8), HostBuffer((n,), np.int32)
    hm.array[...] = workloads.xoshiro_bytes(0xD1110000 + 77, n * mlen).reshape(n, mlen)

    def keypair():
        api.crypto_sign_keypair(mode, hpk.array, hsk.array, kpool, kpool.pitch, n)

    def sign():
        api.crypto_sign_signature(mode, hsm.array, sigb, hm.array, mlen, mlen, hsk.array, spool, spool.pitch, tpool,
                                  tpool.pitch, lut, thr, n)

    def verify():
        api.crypto_sign_verify(mode, hret.array, hsm.array, sigb, sigb, hm.array, mlen, mlen, hpk.array, vpool,
                               vpool.pitch, n)

    streams = [Stream() for _ in range(n_streams)]
    per, per_thr = n // n_streams, thr // n_streams
    slices = [(i * per, per + (n % n_streams if i == n_streams - 1 else 0)) for i in range(n_streams)]
    tslices = [(i * per_thr, per_thr + (thr % n_streams if i == n_streams - 1 else 0)) for i in range(n_streams)]
    luts = [api.TaskLutSlice(lut, f, c, slices[i][1]) for i, (f, c) in enumerate(tslices)]
    pool = ThreadPoolExecutor(max_workers=n_streams)

    def keypair_s():
        # rand_index 2 f: the slice's keys equal the whole-batch call's (the
        # reference passes 0 on every stream, repeating the first slice's keys)
        for i, (f, c) in enumerate(slices):
            api._lib.check(api.L().dil_crypto_sign_keypair(mode, hpk.ptr + f * pkb, hsk.ptr + f * skb,
                                                           kpool.ptr + f * kpool.pitch, kpool.pitch, c,
                                                           streams[i].handle, 2 * f), "keypair")

    def sign_one(i):
        f, c = slices[i]
        api.crypto_sign_signature_rows(mode, hsm.ptr + f * sigb, sigb, hm.ptr + f * mlen, mlen, mlen,
                                       hsk.ptr + f * skb, spool.ptr + f * spool.pitch, spool.pitch,
                                       tpool.ptr + tslices[i][0] * tpool.pitch, tpool.pitch, luts[i], tslices[i][1],
                                       c, streams[i].handle)

    def sign_s():
        list(pool.map(sign_one, range(n_streams)))

    def verify_s():
        for i, (f, c) in enumerate(slices):
            api._lib.check(api.L().dil_crypto_sign_verify(mode, hret.ptr + 4 * f, hsm.ptr + f * sigb, sigb, sigb,
                                                          hm.ptr + f * mlen, mlen, mlen, hpk.ptr + f * pkb,
                                                          vpool.ptr + f * vpool.pitch, vpool.pitch, c,
                                                          streams[i].handle), "verify")

    import oracle
    out = {}
    parity = True
    idx = np.arange(0, n, n // 64)
    for form, ops in (("batch", (keypair, sign, verify)), (f"stream_{n_streams}", (keypair_s, sign_s, verify_s))):
        times = {k: [] for k in ("keypair", "sign", "verify")}
        for trial in range(11):
            hsm.array[...] = 0
            hret.array[...] = 7
            for name, fn in zip(("keypair", "sign", "verify"), ops):
                t0 = time.perf_counter()
                fn()
                synchronize()
                if trial:
                    times[name].append(time.perf_counter() - t0)
        # counter-seeded keys (rand_index 0), as the reference's keypair call
        epk, esk = oracle.keypair_counter_batch(mode, n, rand_index=0)
        esig, _ = oracle.sign_batch(mode, hm.array[idx], esk[idx])
        ok = bool(np.array_equal(hpk.array, epk)) and bool(np.array_equal(hsk.array, esk)) \
            and bool(np.array_equal(hsm.array[idx], esig)) and bool((hret.array == 0).all())
        parity &= ok
        key = "batch" if form == "batch" else "stream"
        out[form] = {k: {"value": n / float(np.median(v)), "unit": "op/s", "ms_median": float(np.median(v)) * 1e3,
                         "reference_published": REF_PUBLISHED[key][k],
                         "ratio_to_reference": n / float(np.median(v)) / REF_PUBLISHED[key][k]}
                     for k, v in times.items()}
        log(f"[bench] reference conditions ({form}): " + ", ".join(
            f"{k} {r['value']:.0f} op/s ({r['ms_median']:.2f} ms; reference {r['reference_published']:.0f})"
            for k, r in out[form].items()) + f"; parity {ok}")
    pool.shutdown()
    return {"workload": f"dilithium2_dropin_batch{n}_msg32B_pinned_host_buffers_exec_threshold{thr}",
            "parity": "bit-exact" if parity else "MISMATCH", "forms": out,
            "note": "wall time per call incl. PCIe copies and device sync, median of 10 "
                    "(benches/bench_dilithium.cu:140-249); reference figures are on its own (unstated) GPU"}


def bench_ntt(args):
    from cudilithium_amd import ntt, workloads, workmodel
    from cudilithium_amd.device import DeviceBuffer, Event, synchronize
    n = args.batch or 65536
    a, b = workloads.config2(n)
    da, db, dc = DeviceBuffer.from_array(a), DeviceBuffer.from_array(b), DeviceBuffer(a.nbytes)
    for _ in range(args.warmup):
        ntt.pipeline_device(dc, da, db, n)
    synchronize()
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(args.steps):
        ntt.pipeline_device(dc, da, db, n)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / args.steps
    import oracle
    c = dc.download(a.shape, np.int32)
    idx = np.random.default_rng(3).choice(n, 256, replace=False)
    parity = bool(np.array_equal(c[idx], oracle.ntt_pipeline(a[idx], b[idx])))
    ops = n * workmodel.NTT_PIPELINE_OPS
    # config 2 sits near the ridge (16,896 ops per 3 KiB): both bounds, the
    # binding one (the larger fraction) first
    valu = {"bound": "valu", "achieved": ops / (ms * 1e-3) / 1e12, "peak": PEAK_INT32_TOPS, "unit": "TOP/s"}
    valu["frac"] = valu["achieved"] / PEAK_INT32_TOPS
    gbs = n * 3072 / (ms * 1e-3) / 1e9  # read A and B, write C
    hbm = {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS}
    first, second = (valu, hbm) if valu["frac"] >= hbm["frac"] else (hbm, valu)
    roof = dict(first, traffic=TRAFFIC_BYTES.get(f"ntt_pipeline_{n}x256"), kernel="ntt_batch_kernel",
                kernel_ms=ms, other_bound=second)
    return {"metric": METRIC, "value": n / (ms * 1e-3), "unit": "polys/s", "ms_per_step": ms, "dtype": "int32",
            "parity": "bit-exact" if parity else "MISMATCH",
            "config": {"workload": f"ntt_pipeline_{n}x256", "batch": n},
            "roofline": roof, "cpu_baseline": None}


def _pct(v, q):
    return float(np.percentile(v, q)) if len(v) else None


def bench_serve(args):
    """Serving mix (continuous batching, resident keys): saturation throughput
    with every request queued at once, then p50 / p99 latency (submit ->
    completion visible on the host) and throughput at 30 / 60 / 90 % of it
    with Poisson arrivals.  Parity: a sample of one run's results against
    the oracle, and a few requests re-run alone against their in-mix bytes."""
    from cudilithium_amd import api, workloads
    n_keys, n = 4096, args.requests
    budget = int(args.budget_gib * (1 << 30))
    keys = workloads.config5_keys(n_keys)
    caps = tuple(len(keys[m][0]) for m in workloads.SERVE_MODES)
    srv = api.Server(budget, lanes=args.lanes, max_batch=args.max_batch, max_keys=caps,
                     batch_window_us=args.window_us, long_message_bytes=args.long_bytes,
                     long_lanes=args.long_lanes)
    for m in workloads.SERVE_MODES:
        srv.stores[m].keygen(0, keys[m][1])
    st0 = srv.stats()
    log(f"[serve] {n_keys} resident keys, budget {budget >> 20} MiB over {args.lanes} lanes "
        f"(max message bytes per batch {st0.max_msg_bytes >> 20} MiB), {n} requests per run")

    def make(rate, off):
        c = workloads.config5(n, rate, n_keys, seed_offset=off)
        sigs = [None] * n
        for m in workloads.SERVE_MODES:
            sel = np.nonzero((c["ops"] == api.OP_VERIFY) & (c["modes"] == m))[0]
            if len(sel):
                ctx = api.SignContext(m, len(sel), resident_only=True)
                sg = ctx.sign_resident(srv.stores[m], c["keys"][sel].astype(np.uint32), [c["msgs"][i] for i in sel])
                for j, i in enumerate(sel):
                    sigs[i] = sg[j]
        return c, api.RequestBatch(c["ops"], c["modes"], c["keys"], c["msgs"], sigs, c["seeds"])

    prev = [srv.stats()]

    def op_breakdown():
        st, p0 = srv.stats(), prev[0]
        prev[0] = st
        out = {}
        for op, name in ((0, "keygen"), (1, "sign"), (2, "verify")):
            nb = st.batches_by_op[op] - p0.batches_by_op[op]
            if nb:
                out[name] = {"batches": nb,
                             "items_per_batch": (st.items_by_op[op] - p0.items_by_op[op]) / nb,
                             "stage_us": (st.stage_ns[op] - p0.stage_ns[op]) / nb / 1e3,
                             "device_us": (st.device_ns[op] - p0.device_ns[op]) / nb / 1e3,
                             "finish_us": (st.finish_ns[op] - p0.finish_ns[op]) / nb / 1e3}
        return out

    def summarize(b, c):
        sub, done = b.field("t_submit_ns").astype(np.int64), b.field("t_done_ns").astype(np.int64)
        lat = (done - sub) / 1e3
        status = b.field("status")
        launch = b.field("t_launch_ns").astype(np.int64)
        wait, service = (launch - sub) / 1e3, (done - launch) / 1e3
        out = {"requests": n, "errors": int((status != 0).sum()),
               "throughput_rps": n / ((done.max() - sub.min()) / 1e9),
               "p50_us": _pct(lat, 50), "p99_us": _pct(lat, 99), "mean_us": float(lat.mean()),
               "p50_wait_us": _pct(wait, 50), "p50_service_us": _pct(service, 50),
               "p99_service_us": _pct(service, 99)}
        for op, name in ((0, "keygen"), (1, "sign"), (2, "verify")):
            sel = c["ops"] == op
            out[f"p99_us_{name}"] = _pct(lat[sel], 99)
        lens = np.array([len(m) for m in c["msgs"]])
        long_sel = (c["ops"] != api.OP_KEYGEN) & (lens > args.long_bytes)
        out["p50_us_short"], out["p99_us_short"] = _pct(lat[~long_sel], 50), _pct(lat[~long_sel], 99)
        out["p50_us_long"], out["p99_us_long"] = _pct(lat[long_sel], 50), _pct(lat[long_sel], 99)
        out["long_fraction"] = float(long_sel.mean())
        out["per_op_batch_means"] = op_breakdown()
        log(f"[serve]   per-op batch means: {json.dumps(out['per_op_batch_means'])}")
        return out

    c, b = make(1e9, 0)
    srv.submit(b.reqs)
    srv.drain()
    sat = summarize(b, c)
    log(f"[serve] saturation {sat['throughput_rps']:.0f} req/s")
    runs = []
    # 30 / 60 / 90 % of the burst saturation, then fixed absolute rates
    levels = [(frac * sat["throughput_rps"], frac) for frac in args.fractions]
    levels += [(float(r), None) for r in args.rates]
    for k, (rate, frac) in enumerate(levels):
        c, b = make(rate, k + 1)
        srv.replay(b.reqs, c["arrival_ns"])
        srv.drain()
        r = summarize(b, c)
        r["offered_rps"] = rate
        r["offered_fraction"] = frac
        runs.append(r)
        log(f"[serve] offered {r['offered_rps']:.0f} req/s: p50 {r['p50_us']:.0f} us, p99 {r['p99_us']:.0f} us, "
            f"{r['throughput_rps']:.0f} req/s (p50 queue wait {r['p50_wait_us']:.0f} us, p50 / p99 service "
            f"{r['p50_service_us']:.0f} / {r['p99_service_us']:.0f} us; messages <= {args.long_bytes} B p50 / p99 "
            f"{r['p50_us_short']:.0f} / {r['p99_us_short']:.0f} us, longer {r['p50_us_long']:.0f} / "
            f"{r['p99_us_long']:.0f} us)")

    # parity of the last run on a sample, and isolation (alone vs in the mix)
    import oracle
    rng = np.random.default_rng(5)
    sample = rng.choice(n, size=min(n, args.check), replace=False)
    ok = True
    for m in workloads.SERVE_MODES:
        pk, sk = srv.stores[m].export(0, len(keys[m][0]))
        for op in (api.OP_KEYGEN, api.OP_SIGN, api.OP_VERIFY):
            sel = [i for i in sample if c["ops"][i] == op and c["modes"][i] == m]
            if not sel:
                continue
            if op == api.OP_KEYGEN:
                epk, esk = oracle.keypair_batch(m, c["seeds"][sel])
                ok &= all(np.array_equal(b.out[i], np.concatenate([epk[j], esk[j]])) for j, i in enumerate(sel))
            elif op == api.OP_SIGN:
                es, _ = oracle.sign_batch(m, [c["msgs"][i] for i in sel], sk[c["keys"][sel]])
                ok &= all(np.array_equal(b.out[i], es[j]) for j, i in enumerate(sel))
            else:
                ev = oracle.verify_batch(m, np.stack([b._sigs_by_index[i] for i in sel]), [c["msgs"][i] for i in sel],
                                         pk[c["keys"][sel]])
                ok &= all(b.reqs[i].verdict == ev[j] for j, i in enumerate(sel))
    alone_ok = True
    for i in sample[:24]:
        one = api.RequestBatch(c["ops"][i:i + 1], c["modes"][i:i + 1], c["keys"][i:i + 1], c["msgs"][i:i + 1],
                               [b._sigs_by_index.get(i)], c["seeds"][i:i + 1])
        srv.submit(one.reqs)
        srv.wait(one.reqs[0])
        alone_ok &= (one.reqs[0].verdict == b.reqs[i].verdict) if c["ops"][i] == api.OP_VERIFY else \
            bool(np.array_equal(one.out[0], b.out[i]))
    st = srv.stats()
    srv.close()
    log(f"[serve] parity vs oracle on {len(sample)} requests: {ok}; alone == in mix on 24: {alone_ok}")
    return {"metric": METRIC, "value": sat["throughput_rps"], "unit": "requests/s", "ms_per_step": None,
            "dtype": "int32", "parity": "bit-exact" if (ok and alone_ok) else "MISMATCH",
            "config": {"workload": f"serving_mix_{n_keys}_resident_keys_keygen20_sign50_verify30_msg_loguniform_32B-64KiB",
                       "requests_per_run": n, "hbm_budget_bytes": budget, "lanes": args.lanes,
                       "max_batch": args.max_batch, "batch_window_us": args.window_us,
                       "long_message_bytes": args.long_bytes, "long_lanes": args.long_lanes or args.lanes,
                       "batches": int(st.batches)},
            "saturation": sat, "loads": runs, "roofline": None, "cpu_baseline": None}


def bench_copy(args):
    """Plain device-to-device copy of `batch` MiB (default 1024): the calibration
    kernel for the FETCH_SIZE / WRITE_SIZE counters and a check of HBM bandwidth."""
    from cudilithium_amd import api
    from cudilithium_amd.device import DeviceBuffer, Event, synchronize
    nbytes = (args.batch or 1024) << 20
    src, dst = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
    L = api.L()
    api._lib.check(L.dil_memset_device(src.ptr, 0x5A, nbytes, None), "memset")
    for _ in range(args.warmup):
        api._lib.check(L.dil_memcpy_d2d(dst.ptr, src.ptr, nbytes, None), "copy")
    synchronize()
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(args.steps):
        api._lib.check(L.dil_memcpy_d2d(dst.ptr, src.ptr, nbytes, None), "copy")
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / args.steps
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    return {"metric": "device copy bandwidth (calibration)", "value": gbs, "unit": "GB/s", "ms_per_step": ms,
            "dtype": "u8", "config": {"workload": f"d2d_copy_{nbytes >> 20}MiB", "bytes_read": nbytes,
                                      "bytes_written": nbytes},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": gbs / PEAK_HBM_GBS, "traffic": None},
            "cpu_baseline": None}


def bench_keccak(args):
    """Chained Keccak-f[1600] permutations in the three state layouts (one
    state per lane / per two lanes / per 16-lane row): latency of one wave
    alone, and the whole card's throughput (2,048 workgroups of 256 threads)."""
    import ctypes
    from cudilithium_amd import api
    L = api.L()
    api._lib.register("dil_keccak_microbench", ctypes.c_int,
                      [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)])
    nperm = max(50, args.steps * 20)
    forms = {}
    for form, name in enumerate(("lane", "pair", "row")):
        r = {}
        for label, blocks, threads in (("one_wave", 1, 64), ("card", 2048, 256)):
            us, rate = ctypes.c_double(), ctypes.c_double()
            api._lib.check(L.dil_keccak_microbench(form, blocks, threads, nperm, ctypes.byref(us), ctypes.byref(rate)),
                           "keccak_microbench")
            r[label] = {"us_per_chained_perm": us.value, "perms_per_s": rate.value}
        forms[name] = r
        log(f"[bench] keccak {name}: one wave {r['one_wave']['us_per_chained_perm']:.2f} us per permutation; "
            f"card {r['card']['perms_per_s'] / 1e9:.2f} G permutations/s")
    return {"metric": "Keccak-f[1600] permutation latency / throughput", "value": forms["row"]["one_wave"]
            ["us_per_chained_perm"], "unit": "us per chained permutation (row form, one wave)",
            "higher_is_better": False, "ms_per_step": None, "dtype": "u32", "parity": None,
            "config": {"workload": f"keccak_chains_{nperm}_permutations", "forms": forms},
            "roofline": None, "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="sign3")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--slots", type=int, default=0, help="attempt slots (default: batch)")
    ap.add_argument("--spb", type=int, default=0, help="slots per workgroup (0 = autotuned)")
    ap.add_argument("--occupancy", type=int, default=0, help="waves per SIMD of the loop kernel (0 = autotuned)")
    ap.add_argument("--claim", type=int, default=0, help="items per workgroup of the loop kernel (0 = its slots)")
    ap.add_argument("--claim-guided", action="store_true", help="guided claims near the end of the queue")
    ap.add_argument("--check", type=int, default=128, help="items checked against the oracle")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-perstage", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="sign3: skip the serving summary")
    ap.add_argument("--no-dropin", action="store_true", help="sign3: skip the host-buffer drop-in summary")
    ap.add_argument("--no-extra", action="store_true",
                    help="sign3: skip the sign2 / sign5 / verify5 / ntt / reference-conditions summaries")
    ap.add_argument("--requests", type=int, default=40000, help="serve: requests per run")
    ap.add_argument("--budget-gib", type=float, default=4.0, help="serve: HBM budget of the in-flight batches")
    ap.add_argument("--lanes", type=int, default=1, help="serve: batches in flight per (op, mode) group")
    ap.add_argument("--max-batch", type=int, default=4096, help="serve: requests per batch")
    ap.add_argument("--window-us", type=int, default=0, help="serve: batching window")
    ap.add_argument("--long-bytes", type=int, default=4096, help="serve: messages above this get lanes of their own")
    ap.add_argument("--long-lanes", type=int, default=0, help="serve: lanes per long-message group (0: --lanes)")
    ap.add_argument("--fractions", type=float, nargs="*", default=[0.3, 0.6, 0.9],
                    help="serve: offered loads as fractions of the burst saturation")
    ap.add_argument("--rates", type=float, nargs="*", default=[25000, 50000, 100000],
                    help="serve: fixed offered loads (requests/s) measured after the saturation fractions")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(shard batches across devices by item range instead)", file=sys.stderr)
        sys.exit(2)
    from cudilithium_amd.device import device_name
    log(f"[bench] device: {device_name()}; host threads {cpu_threads()}")
    w = args.workload
    if w.startswith("sign"):
        res = bench_sign(args, int(w[4:] or 3))
        if w == "sign3" and not args.no_extra:
            # the rest of the metric, each on its own configuration: Dilithium2
            # and Dilithium5 signing (16,384 items, same message law),
            # Dilithium5 verify (65,536, config 4), the standalone NTT
            # pipeline (config 2) and the reference's own benchmark conditions
            import copy
            xargs = copy.copy(args)
            xargs.no_perstage, xargs.no_cpu, xargs.steps = True, True, max(3, args.steps // 2)
            for m in (2, 5):
                r = bench_sign(xargs, m)
                res[f"sign{m}"] = {"workload": r["config"]["workload"], "value": r["value"], "unit": "signs/s",
                                   "ms_per_step": r["ms_per_step"], "parity": r["parity"],
                                   "roofline": {k: r["roofline"][k] for k in ("bound", "achieved", "peak", "unit",
                                                                              "frac", "kernel", "kernel_ms")},
                                   "step": r["step"]}
                if r["parity"] != "bit-exact":
                    res["parity"] = "MISMATCH"
            vargs = copy.copy(args)
            vargs.batch = 0
            r = bench_verify(vargs, 5)
            res["verify5"] = {k: r[k] for k in ("value", "unit", "ms_per_step", "parity", "config", "kernel_ms",
                                                "roofline", "eager_value")}
            r = bench_ntt(vargs)
            res["ntt"] = {k: r[k] for k in ("value", "unit", "ms_per_step", "parity", "config", "roofline")}
            for k in ("verify5", "ntt"):
                if res[k]["parity"] != "bit-exact":
                    res["parity"] = "MISMATCH"
            rc = bench_reference_conditions(args)
            res["reference_conditions"] = rc
            if rc["parity"] != "bit-exact":
                res["parity"] = "MISMATCH"
        if w == "sign3" and not args.no_serve:
            # the serving mix beside the headline: 10,000 requests at two fixed
            # offered loads (the full sweep is --workload serve)
            import copy
            sargs = copy.copy(args)
            sargs.requests, sargs.fractions, sargs.rates = 10000, [], [25000.0, 50000.0]
            srv = bench_serve(sargs)
            res["serving"] = {"workload": srv["config"]["workload"], "parity": srv["parity"],
                              "saturation_rps": srv["value"], "hbm_budget_bytes": srv["config"]["hbm_budget_bytes"],
                              "loads": [{k: r[k] for k in ("offered_rps", "throughput_rps", "p50_us", "p99_us",
                                                           "p50_us_short", "p99_us_short", "p50_us_long",
                                                           "p99_us_long", "long_fraction")} for r in srv["loads"]]}
            if srv["parity"] != "bit-exact":
                res["parity"] = "MISMATCH"
        if w == "sign3" and not args.no_dropin:
            # the reference's host-buffer call surface on the same keys (PCIe-inclusive)
            dr = bench_dropin(args, 3)
            res["dropin_host_buffers"] = {"workload": dr["config"]["workload"], "parity": dr["parity"],
                                          "unit": "signs/s", "calls": dr["config"]["calls"]}
            if dr["parity"] != "bit-exact":
                res["parity"] = "MISMATCH"
    elif w.startswith("dropin"):
        res = bench_dropin(args, int(w[6:] or 3))
    elif w.startswith("verify"):
        res = bench_verify(args, int(w[6:] or 5))
    elif w == "ntt":
        res = bench_ntt(args)
    elif w == "copy":
        res = bench_copy(args)
    elif w == "keccak":
        res = bench_keccak(args)
    elif w == "serve":
        res = bench_serve(args)
    elif w == "refcond":
        res = bench_reference_conditions(args)
        res.update({"metric": METRIC, "value": res["forms"]["batch"]["sign"]["value"], "unit": "signs/s",
                    "ms_per_step": res["forms"]["batch"]["sign"]["ms_median"], "dtype": "int32",
                    "config": {"workload": res["workload"]}, "roofline": None, "cpu_baseline": None})
    else:
        raise SystemExit(f"unknown workload {w}")
    res.setdefault("higher_is_better", True)
    res.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "scaling": "weak", "vs_baseline": None, "data": "synthetic"})
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
