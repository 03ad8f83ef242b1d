// Serving loop: continuous batching of single keygen / sign / verify
// requests onto in-flight batches ("lanes") under a fixed HBM budget.
//
// Requests arrive from any host thread (dil_server_submit) into one FIFO
// per (op, mode).  One worker thread owns the device:
//   harvest  a lane whose completion event has fired scatters its outputs
//            into the requests (signature bytes, verdict, pk || sk), stamps
//            t_done and publishes done = 1
//   launch   while a lane is free and requests wait, a batch takes the
//            oldest requests across all (op, mode) groups, up to max_batch
//            and the lane's message area; inputs are staged into the lane's
//            pinned buffer and copied in one transfer, each group's pipeline
//            runs on its own stream (forked from the lane's stream and joined
//            back), the outputs come back in one transfer and an event marks
//            completion
// Each lane holds everything one batch needs (inputs, outputs, verify /
// keygen work rows and one resident-only signing context per mode), carved
// from the budget at creation, so admission never allocates.  Sign and
// verify requests name resident keys (dil_keystore, one per mode), whose
// expanded A, s1^, s2^, t0^, t1^ and tr are not recomputed per request.
#include <hip/hip_runtime.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/cudilithium_amd.h"
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

constexpr size_t al8(size_t x) { return (x + 7) & ~size_t(7); }
constexpr size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

// per-item byte maxima over the three modes
constexpr size_t MAX_SIG = Sizes<Params<5>>::SIG_BYTES;
constexpr size_t MAX_KEYPAIR = Sizes<Params<5>>::PK_BYTES + Sizes<Params<5>>::SK_BYTES;

// groups g = 3 * op + mode slot; a lane's batch holds any mix of them
struct Lane {
    hipStream_t st = nullptr;  // copies, fork and join
    hipStream_t sub[9] = {};   // one stream per group's pipeline
    hipEvent_t ev = nullptr, fork = nullptr, join[9] = {};
    uint8_t *h_in = nullptr, *h_out = nullptr;
    uint8_t *d_in = nullptr, *d_out = nullptr, *d_work = nullptr;
    dil_sign_ctx *ctx[3] = {nullptr, nullptr, nullptr};
    std::vector<dil_request *> reqs;  // ordered by group
    size_t g0[10] = {};               // group g is reqs[g0[g], g0[g + 1])
    size_t out_off[9] = {};           // group g's outputs in h_out / d_out
    bool busy = false;
    uint32_t batch_id = 0;
};

// input area of a batch of n: key_idx u32[n] | moff u64[n] | mlen u32[n] |
// fixed-size payload (verify: signatures, keygen: seeds) | messages
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

}  // namespace

struct dil_server {
    dil_server_config cfg{};
    int device = 0;
    dil_keystore *ks[3] = {nullptr, nullptr, nullptr};
    std::vector<Lane> lanes;
    size_t in_fixed = 0, msg_cap = 0, out_cap = 0, work_pitch = 0, lane_bytes = 0;

    std::mutex mu;  // queues, stop flag, counters below
    std::condition_variable cv_sub;
    std::deque<dil_request *> q[3][3];  // [op][mode slot]
    size_t npending = 0;
    bool stop = false;
    uint64_t submitted = 0, batches = 0, rejected = 0, by_op[3] = {0, 0, 0};

    std::mutex done_mu;
    std::condition_variable cv_done;
    std::atomic<uint64_t> completed{0};

    std::thread worker;
    uint32_t next_batch = 0;

    void run();
    bool fill(Lane &ln);
    int enqueue(Lane &ln);
    void finish(Lane &ln, int err);
    void publish(dil_request *r, int status) {
        r->status = status;
        r->t_done_ns = mono_ns();
        __atomic_store_n(&r->done, 1, __ATOMIC_RELEASE);
    }
};

namespace {

void free_lane(Lane &ln) {
    for (auto *c : ln.ctx)
        if (c) dil_sign_ctx_destroy(c);
    (void)hipFree(ln.d_in);
    (void)hipFree(ln.d_out);
    (void)hipFree(ln.d_work);
    (void)hipHostFree(ln.h_in);
    (void)hipHostFree(ln.h_out);
    if (ln.ev) (void)hipEventDestroy(ln.ev);
    if (ln.fork) (void)hipEventDestroy(ln.fork);
    for (int g = 0; g < 9; ++g) {
        if (ln.join[g]) (void)hipEventDestroy(ln.join[g]);
        if (ln.sub[g]) (void)hipStreamDestroy(ln.sub[g]);
    }
    if (ln.st) (void)hipStreamDestroy(ln.st);
}

size_t sign_ctx_bytes(int mode, size_t b) {
    const size_t row = mode == 2 ? al256(SignLayout<Params<2>>::ints(0))
                                 : (mode == 3 ? al256(SignLayout<Params<3>>::ints(0))
                                              : al256(SignLayout<Params<5>>::ints(0)));
    return b * (row + al256(dil_sign_temp_row_bytes(mode)));
}

}  // namespace

// Takes the next batch for a free lane (called with mu held): requests of
// every (op, mode) group, oldest first across groups, up to max_batch and the
// lane's message area; then ordered by group so that each group is one
// contiguous range.  Returns false when nothing is ready.
bool dil_server::fill(Lane &ln) {
    const uint64_t now = mono_ns();
    uint64_t oldest = ~0ull;
    size_t waiting = 0;
    for (int o = 0; o < 3; ++o)
        for (int m = 0; m < 3; ++m)
            if (!q[o][m].empty()) {
                oldest = std::min<uint64_t>(oldest, q[o][m].front()->t_submit_ns);
                waiting += q[o][m].size();
            }
    if (waiting == 0) return false;
    // hold a partial batch for the window (it fills while other lanes run)
    if (cfg.batch_window_us && waiting < cfg.max_batch && now - oldest < 1000ull * cfg.batch_window_us) return false;
    ln.reqs.clear();
    size_t bytes = 0;
    for (;;) {
        int bo = -1, bm = -1;
        uint64_t t = ~0ull;
        for (int o = 0; o < 3; ++o)
            for (int m = 0; m < 3; ++m)
                if (!q[o][m].empty() && q[o][m].front()->t_submit_ns < t) {
                    t = q[o][m].front()->t_submit_ns;
                    bo = o;
                    bm = m;
                }
        if (bo < 0 || ln.reqs.size() >= cfg.max_batch) break;
        dil_request *r = q[bo][bm].front();
        const size_t need = bo == DIL_OP_KEYGEN ? 0 : al8(r->mlen);
        if (bytes + need > msg_cap) break;
        bytes += need;
        ln.reqs.push_back(r);
        q[bo][bm].pop_front();
    }
    std::stable_sort(ln.reqs.begin(), ln.reqs.end(), [](const dil_request *a, const dil_request *b) {
        return a->op * 3 + mode_slot(a->mode) < b->op * 3 + mode_slot(b->mode);
    });
    for (int g = 0; g <= 9; ++g) ln.g0[g] = 0;
    for (const auto *r : ln.reqs) ++ln.g0[r->op * 3 + mode_slot(r->mode) + 1];
    for (int g = 0; g < 9; ++g) ln.g0[g + 1] += ln.g0[g];
    npending -= ln.reqs.size();
    ln.batch_id = next_batch++;
    ++batches;
    for (const auto *r : ln.reqs) ++by_op[r->op];
    return !ln.reqs.empty();
}

namespace {
size_t out_item_bytes(int op, int mode) {
    if (op == DIL_OP_SIGN) return dil_signature_bytes(mode);
    if (op == DIL_OP_VERIFY) return 4;
    return dil_public_key_bytes(mode) + dil_secret_key_bytes(mode);
}
}  // namespace

// Stages inputs and enqueues every group's pipeline of the lane's batch, each
// on its own stream forked from and joined back into the lane's stream.
int dil_server::enqueue(Lane &ln) {
    const size_t n = ln.reqs.size();
    const InLayout L(n, n * MAX_SIG);
    uint32_t *idx = reinterpret_cast<uint32_t *>(ln.h_in + L.idx);
    uint64_t *moff = reinterpret_cast<uint64_t *>(ln.h_in + L.moff);
    uint32_t *mlen = reinterpret_cast<uint32_t *>(ln.h_in + L.mlen);
    uint8_t *pay = ln.h_in + L.payload;
    size_t pos = 0;
    const uint64_t t = mono_ns();
    for (size_t i = 0; i < n; ++i) {
        dil_request *r = ln.reqs[i];
        r->t_launch_ns = t;
        r->batch_id = ln.batch_id;
        idx[i] = r->key;
        moff[i] = pos;
        mlen[i] = r->op == DIL_OP_KEYGEN ? 0 : r->mlen;
        if (r->op == DIL_OP_VERIFY) std::memcpy(pay + i * MAX_SIG, r->sig, dil_signature_bytes(r->mode));
        if (r->op == DIL_OP_KEYGEN) std::memcpy(pay + i * MAX_SIG, r->seed, 32);
        if (r->op != DIL_OP_KEYGEN && r->mlen) std::memcpy(ln.h_in + L.msgs + pos, r->msg, r->mlen);
        if (r->op != DIL_OP_KEYGEN) pos += al8(r->mlen);
    }
    int e = (int)hipMemcpyAsync(ln.d_in, ln.h_in, L.msgs + std::max<size_t>(pos, 8), hipMemcpyHostToDevice, ln.st);
    if (e) return e;
    if ((e = (int)hipEventRecord(ln.fork, ln.st))) return e;
    const uint8_t *d_msg = ln.d_in + L.msgs;
    size_t out = 0;
    for (int g = 0; g < 9; ++g) {
        const size_t g0 = ln.g0[g], ng = ln.g0[g + 1] - g0;
        ln.out_off[g] = out;
        if (ng == 0) continue;
        const int op = g / 3, ms = g % 3, mode = MODES[ms];
        const size_t pkb = dil_public_key_bytes(mode), skb = dil_secret_key_bytes(mode);
        const size_t sigb = dil_signature_bytes(mode);
        hipStream_t gs = ln.sub[g];
        if ((e = (int)hipStreamWaitEvent(gs, ln.fork, 0))) return e;
        const uint32_t *d_idx = reinterpret_cast<const uint32_t *>(ln.d_in + L.idx) + g0;
        const uint64_t *d_moff = reinterpret_cast<const uint64_t *>(ln.d_in + L.moff) + g0;
        const uint32_t *d_mlen = reinterpret_cast<const uint32_t *>(ln.d_in + L.mlen) + g0;
        const uint8_t *d_pay = ln.d_in + L.payload + g0 * MAX_SIG;
        uint8_t *d_out = ln.d_out + out;
        uint8_t *d_work = ln.d_work + g0 * work_pitch;
        if (op == DIL_OP_SIGN) {
            dil_sign_ctx *c = ln.ctx[ms];
            // small batches: fewer slots per workgroup, spread over all of them,
            // so that each item gets spb / claim nonces per round
            int spb = ng >= 8192 ? 32 : (ng >= 2048 ? 16 : 8);
            if ((size_t)spb > cfg.max_batch) spb = 1;
            if ((e = dil_sign_ctx_set_tiling(c, spb, 0, 0))) return e;
            const size_t blocks = cfg.max_batch / (size_t)spb;
            const int claim = (int)std::min<size_t>((size_t)spb, std::max<size_t>(1, (ng + blocks - 1) / blocks));
            if ((e = dil_sign_ctx_set_claim(c, claim))) return e;
            e = dil_sign_device_resident(c, ks[ms], d_idx, d_out, sigb, d_msg, d_moff, d_mlen, ng,
                                         DIL_ENGINE_PERSISTENT, gs);
        } else if (op == DIL_OP_VERIFY) {
            e = dil_verify_device_resident(ks[ms], d_idx, reinterpret_cast<int32_t *>(d_out), d_pay, MAX_SIG, d_msg,
                                           d_moff, d_mlen, ng, d_work, work_pitch, gs);
        } else {
            e = dil_keypair_device(mode, d_out, pkb + skb, d_out + pkb, pkb + skb, d_pay, MAX_SIG, 0, ng, d_work,
                                   work_pitch, gs);
        }
        if (e) return e;
        out += al8(ng * out_item_bytes(op, mode));
        if ((e = (int)hipEventRecord(ln.join[g], gs))) return e;
        if ((e = (int)hipStreamWaitEvent(ln.st, ln.join[g], 0))) return e;
    }
    if ((e = (int)hipMemcpyAsync(ln.h_out, ln.d_out, out, hipMemcpyDeviceToHost, ln.st))) return e;
    return (int)hipEventRecord(ln.ev, ln.st);
}

void dil_server::finish(Lane &ln, int err) {
    const size_t n = ln.reqs.size();
    for (int g = 0; g < 9 && !err; ++g) {
        const int op = g / 3, mode = MODES[g % 3];
        const size_t sz = out_item_bytes(op, mode);
        for (size_t i = ln.g0[g]; i < ln.g0[g + 1]; ++i) {
            dil_request *r = ln.reqs[i];
            const uint8_t *src = ln.h_out + ln.out_off[g] + (i - ln.g0[g]) * sz;
            if (op == DIL_OP_VERIFY)
                r->verdict = *reinterpret_cast<const int32_t *>(src);
            else
                std::memcpy(r->out, src, sz);
        }
    }
    for (size_t i = 0; i < n; ++i) publish(ln.reqs[i], err);
    ln.reqs.clear();
    ln.busy = false;
    completed.fetch_add(n, std::memory_order_acq_rel);
    std::lock_guard<std::mutex> g(done_mu);
    cv_done.notify_all();
}

void dil_server::run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool progress = false;
        // harvest finished lanes (without the queue lock)
        lk.unlock();
        for (auto &ln : lanes)
            if (ln.busy) {
                const hipError_t q = hipEventQuery(ln.ev);
                if (q == hipSuccess) {
                    finish(ln, 0);
                    progress = true;
                } else if (q != hipErrorNotReady) {
                    finish(ln, (int)q);
                    progress = true;
                }
            }
        lk.lock();
        // launch onto free lanes
        for (auto &ln : lanes) {
            if (ln.busy || npending == 0) continue;
            if (!fill(ln)) break;
            ln.busy = true;
            lk.unlock();
            const int e = enqueue(ln);
            if (e) finish(ln, e);
            lk.lock();
            progress = true;
        }
        bool any_busy = false;
        for (auto &ln : lanes) any_busy |= ln.busy;
        if (stop && npending == 0 && !any_busy) break;
        if (progress) continue;
        if (any_busy || npending)
            cv_sub.wait_for(lk, std::chrono::microseconds(any_busy ? 20 : std::max<uint32_t>(cfg.batch_window_us, 1) / 4 + 1));
        else
            cv_sub.wait(lk, [&] { return npending > 0 || stop; });
    }
}

extern "C" {

uint64_t dil_monotonic_ns(void) { return mono_ns(); }

int dil_server_create(dil_server **out, const dil_server_config *cfg) {
    if (!out || !cfg || cfg->lanes == 0 || cfg->max_batch == 0 || cfg->max_batch > 65536) return DIL_EINVAL;
    dil_server *s = new dil_server();
    s->cfg = *cfg;
    (void)hipGetDevice(&s->device);
    const size_t b = cfg->max_batch;
    // per lane: outputs, fixed inputs, work rows, signing contexts, messages
    s->out_cap = b * std::max(MAX_SIG, MAX_KEYPAIR) + 9 * 8;
    s->in_fixed = InLayout(b, b * MAX_SIG).msgs;
    size_t wrow = 0;
    for (int m : MODES)
        wrow = std::max({wrow, dil_verify_work_row_bytes(m), al256(dil_keypair_work_row_bytes(m))});
    s->work_pitch = al256(wrow);
    size_t ctx_bytes = 0;
    for (int m : MODES) ctx_bytes += sign_ctx_bytes(m, b);
    const size_t fixed = s->out_cap + s->in_fixed + b * s->work_pitch + ctx_bytes;
    const size_t per_lane = cfg->hbm_budget_bytes / cfg->lanes;
    int e = 0;
    if (per_lane <= fixed + 4096) e = DIL_ENOMEM;
    s->msg_cap = e ? 0 : per_lane - fixed;
    s->lane_bytes = per_lane;
    for (int i = 0; i < 3 && !e; ++i)
        if (cfg->max_keys[i]) e = dil_keystore_create(&s->ks[i], MODES[i], cfg->max_keys[i]);
    s->lanes.resize(cfg->lanes);
    for (auto &ln : s->lanes) {
        if (e) break;
        auto chk = [&](int r) {
            if (r && !e) e = r;
        };
        chk((int)hipStreamCreateWithFlags(&ln.st, hipStreamNonBlocking));
        chk((int)hipEventCreateWithFlags(&ln.ev, hipEventDisableTiming));
        chk((int)hipEventCreateWithFlags(&ln.fork, hipEventDisableTiming));
        for (int g = 0; g < 9; ++g) {
            chk((int)hipStreamCreateWithFlags(&ln.sub[g], hipStreamNonBlocking));
            chk((int)hipEventCreateWithFlags(&ln.join[g], hipEventDisableTiming));
        }
        chk((int)hipHostMalloc((void **)&ln.h_in, s->in_fixed + s->msg_cap, hipHostMallocDefault));
        chk((int)hipHostMalloc((void **)&ln.h_out, s->out_cap, hipHostMallocDefault));
        chk((int)hipMalloc((void **)&ln.d_in, s->in_fixed + s->msg_cap));
        chk((int)hipMalloc((void **)&ln.d_out, s->out_cap));
        chk((int)hipMalloc((void **)&ln.d_work, b * s->work_pitch));
        for (int i = 0; i < 3 && !e; ++i) chk(dil_sign_ctx_create_resident(&ln.ctx[i], MODES[i], b, b));
    }
    if (e) {
        for (auto &ln : s->lanes) free_lane(ln);
        for (auto *k : s->ks) dil_keystore_destroy(k);
        delete s;
        return e;
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return 0;
}

int dil_server_destroy(dil_server *s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_sub.notify_all();
    s->worker.join();
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
            s->q[r->op][ms].push_back(r);
            ++queued;
        }
        s->npending += queued;
    }
    if (rejected) {
        s->completed.fetch_add(rejected, std::memory_order_acq_rel);
        std::lock_guard<std::mutex> g(s->done_mu);
        s->cv_done.notify_all();
    }
    if (queued) s->cv_sub.notify_one();
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
    st->lane_bytes = s->lane_bytes;
    st->max_msg_bytes = s->msg_cap;
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
