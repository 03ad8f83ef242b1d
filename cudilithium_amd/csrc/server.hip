// Serving loop: continuous batching of single keygen / sign / verify
// requests under a fixed HBM budget.
//
// Requests arrive from any host thread (dil_server_submit) into one FIFO
// per group, a group being an (op, mode) pair (9 groups).  Each group owns
// `lanes` lanes; a lane is one batch in flight with its own stream, pinned
// staging buffers and device buffers sized for its group's op.  One worker
// thread per lane runs the lane's cycle:
//   launch   when its group has waiting requests, take them in FIFO order,
//            up to max_batch and the lane's message area, stage the inputs
//            into the pinned buffer, copy them in one transfer, enqueue the
//            group's pipeline (persistent sign with resident keys / resident
//            verify / keygen) on the lane's stream, copy the outputs back in
//            one transfer and record an event
//   harvest  wait for the event (the thread sleeps until it signals), scatter the
//            outputs into the requests (signature bytes, verdict, pk || sk),
//            stamp t_done and publish done = 1
// Lanes never wait for each other's host work.
// Groups run concurrently on their own streams, so a request waits at most
// for its own group's batch in flight (with lanes = 1), not for other
// groups'.  A batch lasts as long as its longest message takes to hash (one
// lane absorbs a message serially, ~10 us per 136-byte block), so sign and
// verify groups keep messages longer than long_message_bytes in queues and
// lanes of their own: short requests are never held up by a long one.  Requests that arrive while batches run join the next batch of
// their group as soon as a lane frees (continuous batching).  All device
// memory -- every lane's buffers and signing context and the message areas
// -- is carved from the budget at creation, so admission never allocates; a
// message larger than its group's message area is rejected (DIL_ERANGE).
// Every lane's stream gets a hardware queue of its own (create_queue_stream):
// streams that share a queue serialise, and measured with plain streams a
// short message's batch waited behind a long message's hash.  For the same
// reason the staging copies are kernels on the lane's stream that read and
// write the pinned buffers directly (stage_copy_kernel): the runtime sends
// larger host<->device copies to DMA engines shared by every stream, where a
// short batch's copy queued behind a long batch's result copy, which waits
// for that batch's hash (gaps of ~15 ms on the short lanes in a kernel
// trace of the serving mix).
#include <hip/hip_runtime.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/cudilithium_amd.h"
#include "internal.hpp"
#include "layout.hpp"
#include "params.hpp"

using namespace dil;

namespace {

uint64_t mono_ns() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

int mode_slot(int mode) { return mode == 2 ? 0 : (mode == 3 ? 1 : (mode == 5 ? 2 : -1)); }
constexpr int MODES[3] = {2, 3, 5};
constexpr int NGROUPS = 9;  // group g = 3 * op + mode slot

constexpr size_t al8(size_t x) { return (x + 7) & ~size_t(7); }
constexpr size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

size_t out_item_bytes(int op, int mode) {
    if (op == DIL_OP_SIGN) return dil_signature_bytes(mode);
    if (op == DIL_OP_VERIFY) return 4;
    return dil_public_key_bytes(mode) + dil_secret_key_bytes(mode);
}

// fixed-size input per item: the signature to verify, or the keygen seed
size_t payload_item_bytes(int op, int mode) {
    if (op == DIL_OP_VERIFY) return dil_signature_bytes(mode);
    if (op == DIL_OP_KEYGEN) return 32;
    return 0;
}

// input area of a batch of n: key_idx u32[n] | moff u64[n] | mlen u32[n] |
// payload (n fixed-size items) | messages
struct InLayout {
    size_t idx, moff, mlen, payload, msgs;
    InLayout(size_t n, size_t payload_bytes) {
        idx = 0;
        moff = al8(4 * n);
        mlen = moff + 8 * n;
        payload = al8(mlen + 4 * n);
        msgs = al8(payload + payload_bytes);
    }
};

size_t sign_ctx_bytes(int mode, size_t b) {
    const size_t row = mode == 2 ? al256(SignLayout<Params<2>>::ints(0))
                                 : (mode == 3 ? al256(SignLayout<Params<3>>::ints(0))
                                              : al256(SignLayout<Params<5>>::ints(0)));
    return b * (row + al256(dil_sign_temp_row_bytes(mode)));
}

struct Lane {
    int group = 0, op = 0, mode = 0, cls = 0;  // cls 1: messages above long_message_bytes
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr, ev0 = nullptr;  // end (completion), start (timing)
    uint8_t *h_in = nullptr, *h_out = nullptr;        // pinned staging (host view)
    uint8_t *h_in_dev = nullptr, *h_out_dev = nullptr;  // the same buffers, device view
    uint8_t *d_in = nullptr, *d_out = nullptr, *d_work = nullptr;
    size_t in_fixed = 0, msg_cap = 0, work_pitch = 0;
    dil_sign_ctx *ctx = nullptr;  // sign lanes
    std::vector<dil_request *> reqs;
    uint32_t batch_id = 0;
    bool backlog = false;  // requests of the group were still waiting after fill()
    std::condition_variable *cv = nullptr;  // woken when its queue gets work
};


// Copies n16 16-byte words (both pointers 16-byte aligned); one side is a
// pinned host buffer's device view.  Grid-stride, so any grid size covers n16.
__global__ __launch_bounds__(256) void stage_copy_kernel(uint4 *dst, const uint4 *src, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int stage_copy(uint8_t *dst, const uint8_t *src, size_t bytes, hipStream_t st) {
    const size_t n16 = (bytes + 15) / 16;
    if (n16 == 0) return 0;
    const unsigned blocks = (unsigned)std::min<size_t>((n16 + 255) / 256, 1024);
    hipLaunchKernelGGL(stage_copy_kernel, dim3(blocks), dim3(256), 0, st, reinterpret_cast<uint4 *>(dst),
                       reinterpret_cast<const uint4 *>(src), n16);
    return (int)hipGetLastError();
}

void free_lane(Lane &ln) {
    if (ln.ctx) dil_sign_ctx_destroy(ln.ctx);
    (void)hipFree(ln.d_in);
    (void)hipFree(ln.d_out);
    (void)hipFree(ln.d_work);
    (void)hipHostFree(ln.h_in);
    (void)hipHostFree(ln.h_out);
    if (ln.ev) (void)hipEventDestroy(ln.ev);
    if (ln.ev0) (void)hipEventDestroy(ln.ev0);
    if (ln.st) (void)hipStreamDestroy(ln.st);
}

}  // namespace

struct dil_server {
    dil_server_config cfg{};
    int device = 0;
    dil_keystore *ks[3] = {nullptr, nullptr, nullptr};
    std::vector<Lane> lanes;
    size_t lane_bytes_max = 0, msg_cap = 0, long_bytes = 0;

    std::mutex mu;  // queues, stop flag, counters below
    // one condition variable per lane; a submit wakes only the lanes of the
    // queues it filled (waking every lane thread per request floods the CPUs)
    std::vector<std::unique_ptr<std::condition_variable>> cvs;
    std::vector<int> lanes_of[NGROUPS][2];
    std::deque<dil_request *> q[NGROUPS][2];  // [group][short / long message]
    size_t npending = 0;
    bool stop = false;
    uint64_t submitted = 0, batches = 0, rejected = 0, by_op[3] = {0, 0, 0};
    // per-op timing sums (written by the worker, read by dil_server_get_stats)
    std::atomic<uint64_t> nb_op[3] = {}, stage_ns[3] = {}, device_ns[3] = {}, finish_ns[3] = {};

    std::mutex done_mu;
    std::condition_variable cv_done;
    std::atomic<uint64_t> completed{0};

    std::vector<std::thread> workers;  // one per lane
    uint32_t next_batch = 0;

    void lane_loop(Lane &ln);
    bool fill(Lane &ln);
    int enqueue(Lane &ln);
    void finish(Lane &ln, int err);
    void publish(dil_request *r, int status) {
        r->status = status;
        r->t_done_ns = mono_ns();
        __atomic_store_n(&r->done, 1, __ATOMIC_RELEASE);
    }
};

// Takes the next batch of the lane's group (called with mu held): FIFO, up
// to max_batch and the lane's message area.  Returns false when nothing is
// ready.
bool dil_server::fill(Lane &ln) {
    auto &dq = q[ln.group][ln.cls];
    if (dq.empty()) return false;
    // hold a partial batch for the window (it fills while other lanes run)
    if (cfg.batch_window_us && dq.size() < cfg.max_batch &&
        mono_ns() - dq.front()->t_submit_ns < 1000ull * cfg.batch_window_us)
        return false;
    ln.reqs.clear();
    size_t bytes = 0;
    while (!dq.empty() && ln.reqs.size() < cfg.max_batch) {
        dil_request *r = dq.front();
        const size_t need = ln.op == DIL_OP_KEYGEN ? 0 : al8(r->mlen);
        if (bytes + need > ln.msg_cap) break;
        bytes += need;
        ln.reqs.push_back(r);
        dq.pop_front();
    }
    npending -= ln.reqs.size();
    ln.backlog = !dq.empty();
    ln.batch_id = next_batch++;
    ++batches;
    by_op[ln.op] += ln.reqs.size();
    return !ln.reqs.empty();
}

// Stages inputs and enqueues the group's pipeline for the lane's batch.
int dil_server::enqueue(Lane &ln) {
    const size_t n = ln.reqs.size();
    const int mode = ln.mode, ms = mode_slot(mode);
    const size_t pkb = dil_public_key_bytes(mode), skb = dil_secret_key_bytes(mode);
    const size_t sigb = dil_signature_bytes(mode), pay = payload_item_bytes(ln.op, mode);
    const InLayout L(n, n * pay);
    uint32_t *idx = reinterpret_cast<uint32_t *>(ln.h_in + L.idx);
    uint64_t *moff = reinterpret_cast<uint64_t *>(ln.h_in + L.moff);
    uint32_t *mlen = reinterpret_cast<uint32_t *>(ln.h_in + L.mlen);
    size_t pos = 0;
    const uint64_t t = mono_ns();
    for (size_t i = 0; i < n; ++i) {
        dil_request *r = ln.reqs[i];
        r->t_launch_ns = t;
        r->batch_id = ln.batch_id;
        idx[i] = r->key;
        moff[i] = pos;
        mlen[i] = ln.op == DIL_OP_KEYGEN ? 0 : r->mlen;
        if (ln.op == DIL_OP_VERIFY) std::memcpy(ln.h_in + L.payload + i * pay, r->sig, pay);
        if (ln.op == DIL_OP_KEYGEN) std::memcpy(ln.h_in + L.payload + i * pay, r->seed, pay);
        if (ln.op != DIL_OP_KEYGEN) {
            if (r->mlen) std::memcpy(ln.h_in + L.msgs + pos, r->msg, r->mlen);
            pos += al8(r->mlen);
        }
    }
    int e = (int)hipEventRecord(ln.ev0, ln.st);
    if (e) return e;
    // (L.msgs + pos <= in_fixed + msg_cap; the buffers hold that rounded up to 16 bytes)
    if ((e = stage_copy(ln.d_in, ln.h_in_dev, L.msgs + std::max<size_t>(pos, 8), ln.st))) return e;
    const uint32_t *d_idx = reinterpret_cast<const uint32_t *>(ln.d_in + L.idx);
    const uint64_t *d_moff = reinterpret_cast<const uint64_t *>(ln.d_in + L.moff);
    const uint32_t *d_mlen = reinterpret_cast<const uint32_t *>(ln.d_in + L.mlen);
    const uint8_t *d_msg = ln.d_in + L.msgs, *d_pay = ln.d_in + L.payload;
    if (ln.op == DIL_OP_SIGN) {
        // small batches: fewer slots per workgroup.  Without a backlog the
        // batch is spread over all workgroups, so that each item gets
        // spb / claim nonces per round and finishes in fewer rounds (latency);
        // with one, every slot is a distinct item's attempt (throughput).
        int spb = n >= 8192 ? 32 : (n >= 2048 ? 16 : 8);
        if ((size_t)spb > cfg.max_batch) spb = 1;
        if ((e = dil_sign_ctx_set_tiling(ln.ctx, spb, 0, 0))) return e;
        const size_t blocks = cfg.max_batch / (size_t)spb;
        const int claim = ln.backlog ? spb
                                     : (int)std::min<size_t>((size_t)spb,
                                                             std::max<size_t>(1, (n + blocks - 1) / blocks));
        if ((e = dil_sign_ctx_set_claim(ln.ctx, claim))) return e;
        e = dil_sign_device_resident(ln.ctx, ks[ms], d_idx, ln.d_out, sigb, d_msg, d_moff, d_mlen, n,
                                     DIL_ENGINE_PERSISTENT, ln.st);
    } else if (ln.op == DIL_OP_VERIFY) {
        e = dil_verify_device_resident_ex(ks[ms], d_idx, reinterpret_cast<int32_t *>(ln.d_out), d_pay, sigb, d_msg,
                                          d_moff, d_mlen, n, ln.d_work, ln.work_pitch,
                                          ln.cls ? DIL_HASH_LONG_MESSAGES : 0, ln.st);
    } else {
        e = dil_keypair_device(mode, ln.d_out, pkb + skb, ln.d_out + pkb, pkb + skb, d_pay, 32, 0, n, ln.d_work,
                               ln.work_pitch, ln.st);
    }
    if (e) return e;
    if ((e = stage_copy(ln.h_out_dev, ln.d_out, n * out_item_bytes(ln.op, mode), ln.st))) return e;
    return (int)hipEventRecord(ln.ev, ln.st);
}

void dil_server::finish(Lane &ln, int err) {
    const uint64_t t0 = mono_ns();
    const size_t n = ln.reqs.size(), sz = out_item_bytes(ln.op, ln.mode);
    float ms = 0.f;
    if (!err && hipEventElapsedTime(&ms, ln.ev0, ln.ev) == hipSuccess) device_ns[ln.op] += (uint64_t)(ms * 1e6f);
    for (size_t i = 0; i < n; ++i) {
        dil_request *r = ln.reqs[i];
        if (!err) {
            if (ln.op == DIL_OP_VERIFY)
                r->verdict = reinterpret_cast<const int32_t *>(ln.h_out)[i];
            else
                std::memcpy(r->out, ln.h_out + i * sz, sz);
        }
        publish(r, err);
    }
    ln.reqs.clear();
    finish_ns[ln.op] += mono_ns() - t0;
    completed.fetch_add(n, std::memory_order_acq_rel);
    std::lock_guard<std::mutex> g(done_mu);
    cv_done.notify_all();
}

void dil_server::lane_loop(Lane &ln) {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        auto &dq = q[ln.group][ln.cls];
        if (dq.empty()) {
            if (stop) break;
            ln.cv->wait(lk);
            continue;
        }
        if (!fill(ln)) {  // holding a partial batch for the window
            ln.cv->wait_for(lk, std::chrono::microseconds(std::max<uint32_t>(cfg.batch_window_us, 4) / 4));
            continue;
        }
        lk.unlock();
        const uint64_t t0 = mono_ns();
        int e = enqueue(ln);
        stage_ns[ln.op] += mono_ns() - t0;
        ++nb_op[ln.op];
        // the completion event is a blocking-sync event: the thread sleeps
        // until the event signals instead of polling the runtime (15 polling
        // threads contend with each other's launches)
        if (!e) e = (int)hipEventSynchronize(ln.ev);
        finish(ln, e);
        lk.lock();
    }
}

extern "C" {

uint64_t dil_monotonic_ns(void) { return mono_ns(); }

int dil_server_create(dil_server **out, const dil_server_config *cfg) {
    if (!out || !cfg || cfg->lanes == 0 || cfg->lanes > 16 || cfg->long_lanes > 16 || cfg->max_batch == 0 ||
        cfg->max_batch > 65536)
        return DIL_EINVAL;
    dil_server *s = new dil_server();
    s->cfg = *cfg;
    (void)hipGetDevice(&s->device);
    const size_t b = cfg->max_batch;
    // every lane's fixed part (index arrays, payload, outputs, work rows,
    // signing context) first; the rest of the budget is split evenly into
    // the message areas of the sign and verify lanes
    s->long_bytes = cfg->long_message_bytes;
    std::vector<std::pair<int, int>> kinds;  // (group, class) of each lane
    const uint32_t long_lanes = cfg->long_lanes ? cfg->long_lanes : cfg->lanes;
    for (uint32_t k = 0; k < cfg->lanes; ++k)
        for (int g = 0; g < NGROUPS; ++g) kinds.push_back({g, 0});
    for (uint32_t k = 0; s->long_bytes && k < long_lanes; ++k)
        for (int g = 3; g < NGROUPS; ++g) kinds.push_back({g, 1});
    s->lanes.resize(kinds.size());
    size_t fixed = 0, msg_lanes = 0, short_msg = 0;
    for (size_t i = 0; i < s->lanes.size(); ++i) {
        Lane &ln = s->lanes[i];
        ln.group = kinds[i].first;
        ln.cls = kinds[i].second;
        ln.op = ln.group / 3;
        ln.mode = MODES[ln.group % 3];
        ln.in_fixed = InLayout(b, b * payload_item_bytes(ln.op, ln.mode)).msgs;
        if (ln.op != DIL_OP_SIGN) {
            const size_t w = ln.op == DIL_OP_VERIFY ? dil_verify_work_row_bytes(ln.mode)
                                                    : dil_keypair_work_row_bytes(ln.mode);
            ln.work_pitch = al256(w);
        }
        fixed += ln.in_fixed + 8 + b * out_item_bytes(ln.op, ln.mode) + 1024 /* staging rounding */ +
                 b * ln.work_pitch +
                 (ln.op == DIL_OP_SIGN ? sign_ctx_bytes(ln.mode, b) : 0);
        // short-message lanes of a split server hold at most b messages of
        // long_message_bytes; the long-message (or unsplit) lanes share the rest
        if (ln.op != DIL_OP_KEYGEN) {
            if (s->long_bytes && ln.cls == 0)
                short_msg += b * al8(s->long_bytes);
            else
                ++msg_lanes;
        }
    }
    int e = 0;
    const size_t budget = cfg->hbm_budget_bytes;
    if (budget <= fixed + short_msg + msg_lanes * 4096) e = DIL_ENOMEM;
    s->msg_cap = e ? 0 : ((budget - fixed - short_msg) / msg_lanes) & ~size_t(255);
    for (int i = 0; i < 3 && !e; ++i)
        if (cfg->max_keys[i]) e = dil_keystore_create(&s->ks[i], MODES[i], cfg->max_keys[i]);
    for (auto &ln : s->lanes) {
        if (e) break;
        auto chk = [&](int r) {
            if (r && !e) e = r;
        };
        ln.msg_cap = ln.op == DIL_OP_KEYGEN ? 0 : (s->long_bytes && ln.cls == 0 ? b * al8(s->long_bytes) : s->msg_cap);
        // staging sizes rounded up for the 16-byte copy words
        const size_t in_bytes = al256(ln.in_fixed + ln.msg_cap + 8) + 256,
                     out_bytes = al256(b * out_item_bytes(ln.op, ln.mode)) + 256;
        s->lane_bytes_max = std::max(s->lane_bytes_max, in_bytes + out_bytes + b * ln.work_pitch +
                                                            (ln.op == DIL_OP_SIGN ? sign_ctx_bytes(ln.mode, b) : 0));
        chk(create_queue_stream(&ln.st));
        chk((int)hipEventCreateWithFlags(&ln.ev, hipEventBlockingSync));
        chk((int)hipEventCreate(&ln.ev0));
        // coherent: the copy kernels' host writes are visible once the lane's event completes
        chk((int)hipHostMalloc((void **)&ln.h_in, in_bytes, hipHostMallocMapped | hipHostMallocCoherent));
        chk((int)hipHostMalloc((void **)&ln.h_out, out_bytes, hipHostMallocMapped | hipHostMallocCoherent));
        if (!e) chk((int)hipHostGetDevicePointer((void **)&ln.h_in_dev, ln.h_in, 0));
        if (!e) chk((int)hipHostGetDevicePointer((void **)&ln.h_out_dev, ln.h_out, 0));
        if (!e && (!ln.h_in_dev || !ln.h_out_dev)) e = DIL_EINVAL;
        chk((int)hipMalloc((void **)&ln.d_in, in_bytes));
        chk((int)hipMalloc((void **)&ln.d_out, out_bytes));
        if (ln.work_pitch) chk((int)hipMalloc((void **)&ln.d_work, b * ln.work_pitch));
        if (ln.op == DIL_OP_SIGN) {
            chk(dil_sign_ctx_create_resident(&ln.ctx, ln.mode, b, b));
            // long-message lanes hash each message with a 16-lane row
            if (ln.ctx && ln.cls) chk(dil_sign_ctx_set_message_hashing(ln.ctx, DIL_HASH_LONG_MESSAGES));
        }
    }
    if (e) {
        for (auto &ln : s->lanes) free_lane(ln);
        for (auto *k : s->ks) dil_keystore_destroy(k);
        delete s;
        return e;
    }
    for (size_t i = 0; i < s->lanes.size(); ++i) {
        s->cvs.emplace_back(new std::condition_variable());
        s->lanes[i].cv = s->cvs.back().get();
        s->lanes_of[s->lanes[i].group][s->lanes[i].cls].push_back((int)i);
    }
    for (auto &ln : s->lanes) s->workers.emplace_back([s, &ln] { s->lane_loop(ln); });
    *out = s;
    return 0;
}

int dil_server_destroy(dil_server *s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    for (auto &cv : s->cvs) cv->notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &ln : s->lanes) free_lane(ln);
    for (auto *k : s->ks) dil_keystore_destroy(k);
    delete s;
    return 0;
}

dil_keystore *dil_server_keystore(dil_server *s, int mode) {
    const int ms = mode_slot(mode);
    return s && ms >= 0 ? s->ks[ms] : nullptr;
}

int dil_server_submit(dil_server *s, dil_request *const *reqs, size_t n) {
    if (!s || (n && !reqs)) return DIL_EINVAL;
    const uint64_t t = mono_ns();
    size_t queued = 0, rejected = 0;
    uint32_t touched = 0;  // bit 2 * group + class of every queue that got work
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (size_t i = 0; i < n; ++i) {
            dil_request *r = reqs[i];
            r->t_submit_ns = t;
            r->t_launch_ns = 0;
            r->done = 0;
            r->status = 0;
            ++s->submitted;
            const int ms = mode_slot(r->mode);
            int bad = (ms < 0 || r->op < 0 || r->op > 2) ? DIL_EINVAL : 0;
            if (!bad && r->op == DIL_OP_KEYGEN && (!r->seed || !r->out)) bad = DIL_EINVAL;
            if (!bad && r->op != DIL_OP_KEYGEN) {
                if (!s->ks[ms] || r->key >= dil_keystore_capacity(s->ks[ms])) bad = DIL_EINVAL;
                else if (r->op == DIL_OP_SIGN && !dil_keystore_has_secret(s->ks[ms], r->key)) bad = DIL_EINVAL;
                else if (r->mlen && !r->msg) bad = DIL_EINVAL;
                else if (al8(r->mlen) > s->msg_cap) bad = DIL_ERANGE;
                else if (r->op == DIL_OP_SIGN ? !r->out : !r->sig) bad = DIL_EINVAL;
            }
            if (bad) {
                ++s->rejected;
                ++rejected;
                s->publish(r, bad);
                continue;
            }
            const int cls = (s->long_bytes && r->op != DIL_OP_KEYGEN && r->mlen > s->long_bytes) ? 1 : 0;
            s->q[r->op * 3 + ms][cls].push_back(r);
            touched |= 1u << (2 * (r->op * 3 + ms) + cls);
            ++queued;
        }
        s->npending += queued;
    }
    if (rejected) {
        s->completed.fetch_add(rejected, std::memory_order_acq_rel);
        std::lock_guard<std::mutex> g(s->done_mu);
        s->cv_done.notify_all();
    }
    for (int k = 0; k < 2 * NGROUPS; ++k)
        if (touched & (1u << k))
            for (int li : s->lanes_of[k / 2][k % 2]) s->lanes[li].cv->notify_one();
    return 0;
}

int dil_server_wait(dil_server *s, dil_request *r, int64_t timeout_us) {
    if (!s || !r) return DIL_EINVAL;
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::microseconds(timeout_us < 0 ? 0 : timeout_us);
    std::unique_lock<std::mutex> lk(s->done_mu);
    while (!__atomic_load_n(&r->done, __ATOMIC_ACQUIRE)) {
        if (timeout_us < 0) {
            s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
        } else {
            if (std::chrono::steady_clock::now() >= deadline) return DIL_ERANGE;
            s->cv_done.wait_until(lk, deadline);
        }
    }
    return 0;
}

int dil_server_drain(dil_server *s) {
    if (!s) return DIL_EINVAL;
    uint64_t target;
    {
        std::lock_guard<std::mutex> g(s->mu);
        target = s->submitted;
    }
    std::unique_lock<std::mutex> lk(s->done_mu);
    while (s->completed.load(std::memory_order_acquire) < target) s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
    return 0;
}

int dil_server_get_stats(dil_server *s, dil_server_stats *st) {
    if (!s || !st) return DIL_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    st->submitted = s->submitted;
    st->completed = s->completed.load();
    st->batches = s->batches;
    st->rejected = s->rejected;
    for (int i = 0; i < 3; ++i) st->items_by_op[i] = s->by_op[i];
    st->lane_bytes = s->lane_bytes_max;
    st->max_msg_bytes = s->msg_cap;
    for (int i = 0; i < 3; ++i) {
        st->batches_by_op[i] = s->nb_op[i];
        st->stage_ns[i] = s->stage_ns[i];
        st->device_ns[i] = s->device_ns[i];
        st->finish_ns[i] = s->finish_ns[i];
    }
    return 0;
}

int dil_server_replay(dil_server *s, dil_request *reqs, const uint64_t *arrival_ns, size_t n) {
    if (!s || (n && (!reqs || !arrival_ns))) return DIL_EINVAL;
    std::vector<dil_request *> due;
    const uint64_t t0 = mono_ns();
    size_t i = 0;
    while (i < n) {
        const uint64_t now = mono_ns() - t0;
        due.clear();
        while (i < n && arrival_ns[i] <= now) due.push_back(&reqs[i++]);
        if (!due.empty()) {
            int e = dil_server_submit(s, due.data(), due.size());
            if (e) return e;
            continue;
        }
        const uint64_t gap = arrival_ns[i] - now;
        if (gap > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(gap - 100000));
        else if (gap > 20000) std::this_thread::yield();
    }
    return 0;
}

}  // extern "C"
