// xc_serve.hip -- serving loop: continuous batching of pair-correlation
// requests over resident tiles, admission by an HBM budget, resident row
// spectra, batches overlapping on several streams, request deadlines and
// cancellation.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "xc_exec.hpp"
#include "xc_kernels.hpp"
#include "xc_sched.hpp"

using namespace xc;
using xc::sched::ArenaAlloc;
using xc::sched::Req;

namespace {

using Clock = xc::sched::Clock;

struct SpecEntry {
    size_t off = 0, bytes = 0;
    int refs = 0;                 // batches in flight that read or write it
    hipEvent_t ready = nullptr;   // recorded after the producing batch's row pass
    uint64_t last_use = 0;
};

struct Batch {
    int slot = -1;
    std::vector<uint64_t> ids;
    std::vector<std::pair<size_t, size_t>> blocks;  // arena blocks to release
    std::vector<std::pair<uint64_t, int>> specs;  // spectra to deref
    hipEvent_t done = nullptr;
};

struct Slot {
    hipStream_t stream = nullptr;
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    bool busy = false;
};

}  // namespace

struct xc_server {
    xc_ctx *ctx = nullptr;  // private context: twiddles, error text, launch stream
    hipStream_t own_stream = nullptr;  // the private context's stream (slots borrow ctx->stream)
    xc_serve_cfg cfg;
    void *arena_mem = nullptr;
    ArenaAlloc arena;
    std::mutex mu;
    std::condition_variable cv;
    xc::sched::Queue q;  // requests by ticket; those not yet launched in arrival order
    std::map<std::pair<uint64_t, int>, SpecEntry> specs;  // (tile uid, Nx) -> row spectrum
    std::vector<Slot> slots;
    std::vector<std::unique_ptr<Batch>> inflight;
    uint64_t clock = 0;
    bool stop = false;
    std::thread thr;
    xc_serve_stats st;
    int err = XC_OK;
    int fail_pinned = 0;  // fault injection (XC_SERVE_FAIL_PINNED): pinned allocations still to refuse
};

namespace {

size_t pair_work_bytes(const PairDev &D)
{
    return ArenaAlloc::round(sizeof(cpx<float>) * (size_t)(D.Nx / 2 + 1) * D.hR) + ArenaAlloc::round(sizeof(RowRec) * D.hR);
}

size_t spec_bytes(const xc_tile *t, int Nx) { return sizeof(cpx<float>) * (size_t)(Nx / 2 + 1) * t->h; }

int make_pair_dev(xc_server *S, const xc_pair &q, PairDev &D)
{
    memset(&D, 0, sizeof D);
    const int mode = q.mode == XC_MODE_NCC ? PM_NCC : (q.mode == XC_MODE_R ? PM_R : PM_S);
    setup_pair_geometry(D, q.a->w, q.a->h, q.b->w, q.b->h, mode);
    D.Ox = q.Ox;
    D.Oy = q.Oy;
    D.radius = q.radius;
    D.Rx = q.Rx;
    D.Ry = q.Ry;
    D.offx = q.b->ox - q.a->ox;  // B2.L - B1.L of the views' bounding boxes
    D.offy = q.b->oy - q.a->oy;
    D.rule = make_rule(&q.rgn, &q.cnt, nullptr);
    D.need_cnt = D.rule.cnt_enabled;
    D.tabk = q.a->dev.tabk;
    narrow_rows(D);
    D.A = q.a->dev.v;
    D.B = q.b->dev.v;
    D.SQA = q.a->dev.SQ;
    D.ZA = q.a->dev.Z;
    D.SQB = q.b->dev.SQ;
    D.ZB = q.b->dev.Z;
    (void)S;
    return XC_OK;
}

// Evict least-recently-used idle spectra until `need` bytes can be allocated.
bool alloc_evicting(xc_server *S, size_t need, size_t &off)
{
    while (!S->arena.alloc(need, off)) {
        auto victim = S->specs.end();
        for (auto it = S->specs.begin(); it != S->specs.end(); ++it)
            if (it->second.refs == 0 && (victim == S->specs.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == S->specs.end()) return false;
        S->arena.release(victim->second.off, victim->second.bytes);
        if (victim->second.ready) (void)hipEventDestroy(victim->second.ready);
        S->specs.erase(victim);
    }
    return true;
}

// A batch whose requests have left the queue failed before it was complete
// on the device: wait for whatever it enqueued, complete its requests with
// the error, give back its arena blocks and spectrum references, and drop the
// spectra it was to produce (never computed, so never to be read).  Lock held.
void fail_batch(xc_server *S, Batch &B, const std::map<std::pair<uint64_t, int>, bool> &fresh, int st)
{
    Slot &slot = S->slots[B.slot];
    (void)hipStreamSynchronize(slot.stream);
    (void)hipGetLastError();
    const auto now = Clock::now();
    for (uint64_t id : B.ids) S->q.complete(id, st, now);
    S->st.failed += (int64_t)B.ids.size();
    for (auto &b : B.blocks)
        if (b.first != SIZE_MAX) S->arena.release(b.first, b.second);
    for (auto &key : B.specs) S->specs[key].refs--;
    for (auto &kv : fresh) {
        auto it = S->specs.find(kv.first);
        if (it == S->specs.end()) continue;
        S->arena.release(it->second.off, it->second.bytes);
        if (it->second.ready) (void)hipEventDestroy(it->second.ready);
        S->specs.erase(it);
    }
    if (B.done) (void)hipEventDestroy(B.done);
    slot.busy = false;
    S->cv.notify_all();
}

// Descriptors, work lists and launches of a batch whose requests are taken
// (lock held).  XC_E_NOTREADY: its descriptor block does not fit yet (the
// requests are given back to the front of the queue); any other error leaves
// the cleanup to the caller.
int enqueue_batch(xc_server *S, int slot_i, std::unique_ptr<Batch> &B, std::vector<PairDev> &P,
                  const std::vector<bool> &ownA, const std::vector<bool> &ownB, const std::vector<hipEvent_t> &waits,
                  std::map<std::pair<uint64_t, int>, bool> &fresh)
{
    Slot &slot = S->slots[slot_i];
    const int n = (int)B->ids.size();
    std::vector<int> order;
    std::vector<Bucket> buckets;
    make_buckets(P, 0, order, buckets);
    Layout lay;
    size_t i_pairs = lay.add(sizeof(PairDev) * n), i_res = lay.add(sizeof(xc_result) * n);
    std::vector<size_t> i_wr(buckets.size()), i_wc(buckets.size()), i_wc2(buckets.size());
    std::vector<int> starts;
    for (size_t bi = 0; bi < buckets.size(); ++bi) {
        Bucket &b = buckets[bi];
        const int G = row_group<float>(b.NH), C = col_group<float>(b.NY), GR = r2c_rows<float>(b.NH);
        int nr = 0, nc = 0, nc2 = 0;
        for (int k = b.first; k < b.first + b.count; ++k) {
            const int i = order[k];
            const PairDev &D = P[i];
            if (ownA[i]) nr += (D.h1 + GR - 1) / GR;
            if (ownB[i]) nr += (D.h2 + GR - 1) / GR;
            nc += (D.Nx / 2 + 1 + C - 1) / C;
            live_row_groups(D, G, starts);
            nc2 += (int)starts.size();
        }
        b.n_r2c = nr;
        b.n_col = nc;
        b.n_c2r = nc2;
        i_wr[bi] = lay.add(sizeof(int4) * nr + 16);
        i_wc[bi] = lay.add(sizeof(int2) * nc);
        i_wc2[bi] = lay.add(sizeof(int2) * nc2);
    }
    size_t off_d;
    if (!alloc_evicting(S, lay.total, off_d)) {
        // give the requests back (front of the queue, same order)
        for (int k = n - 1; k >= 0; --k) {
            S->q.reqs[B->ids[k]].launched = false;
            S->q.pending.push_front(B->ids[k]);
        }
        for (auto &b : B->blocks) S->arena.release(b.first, b.second);
        for (auto &key : B->specs) S->specs[key].refs--;
        for (auto &kv : fresh) {  // never computed: drop them
            S->arena.release(S->specs[kv.first].off, S->specs[kv.first].bytes);
            S->specs.erase(kv.first);
        }
        B->ids.clear();
        B->blocks.clear();
        B->specs.clear();
        fresh.clear();
        return XC_E_NOTREADY;
    }
    B->blocks.push_back(std::make_pair(off_d, lay.total));
    char *dbase = S->arena.base + off_d;
    if (slot.pinned_bytes < lay.total) {
        if (slot.pinned) (void)hipHostFree(slot.pinned);
        slot.pinned = nullptr;
        slot.pinned_bytes = 0;
        if (S->fail_pinned > 0) {
            --S->fail_pinned;
            return set_err(S->ctx, XC_E_NOMEM, "pinned staging refused (fault injection)");
        }
        if (hipHostMalloc(&slot.pinned, lay.total * 2, hipHostMallocDefault) != hipSuccess) {
            slot.pinned = nullptr;
            (void)hipGetLastError();
            return set_err(S->ctx, XC_E_NOMEM, "pinned staging of %zu bytes", lay.total * 2);
        }
        slot.pinned_bytes = lay.total * 2;
    }
    char *h = (char *)slot.pinned;
    PairDev *hp = (PairDev *)(h + lay.off[i_pairs]);
    xc_result *dres = (xc_result *)(dbase + lay.off[i_res]);
    for (int k = 0; k < n; ++k) {
        const int i = order[k];
        PairDev D = P[i];
        D.out = dres + i;
        hp[k] = D;
    }
    for (size_t bi = 0; bi < buckets.size(); ++bi) {
        Bucket &b = buckets[bi];
        const int G = row_group<float>(b.NH), C = col_group<float>(b.NY), GR = r2c_rows<float>(b.NH);
        int4 *wr = (int4 *)(h + lay.off[i_wr[bi]]);
        int2 *wc = (int2 *)(h + lay.off[i_wc[bi]]);
        int2 *wc2 = (int2 *)(h + lay.off[i_wc2[bi]]);
        int a = 0, c = 0, d = 0;
        for (int k = b.first; k < b.first + b.count; ++k) {
            const int i = order[k];
            const PairDev &D = hp[k];
            if (ownA[i])
                for (int y = 0; y < D.h1; y += GR) wr[a++] = make_int4(k, 0, y, 0);
            if (ownB[i])
                for (int y = 0; y < D.h2; y += GR) wr[a++] = make_int4(k, 1, y, 0);
            for (int x = 0; x < D.Nx / 2 + 1; x += C) wc[c++] = make_int2(k, x);
            live_row_groups(D, G, starts);
            for (int y : starts) wc2[d++] = make_int2(k, y);
        }
        b.d_r2c = (const int4 *)(dbase + lay.off[i_wr[bi]]);
        b.d_col = (const int2 *)(dbase + lay.off[i_wc[bi]]);
        b.d_c2r = (const int2 *)(dbase + lay.off[i_wc2[bi]]);
    }
    xc_ctx *ctx = S->ctx;
    ctx->stream = slot.stream;
    for (hipEvent_t e : waits) XC_HIP(ctx, hipStreamWaitEvent(slot.stream, e, 0));
    XC_HIP(ctx, hipMemcpyAsync(dbase, h, lay.total, hipMemcpyHostToDevice, slot.stream));
    Exec ex;
    ex.fp64 = false;
    ex.d_pairs = (const PairDev *)(dbase + lay.off[i_pairs]);
    ex.buckets = buckets;
    int st = enqueue(ctx, ex, ST_R2C, ST_R2C + 1);
    if (st) return st;
    // spectra produced here become readable by later batches after this point
    for (auto &kv : fresh) {
        SpecEntry &e = S->specs[kv.first];
        if (!e.ready) XC_HIP(ctx, hipEventCreateWithFlags(&e.ready, hipEventDisableTiming));
        XC_HIP(ctx, hipEventRecord(e.ready, slot.stream));
    }
    if ((st = enqueue(ctx, ex, ST_COL, ST_COUNT))) return st;
    // results back into the slot's pinned area (after the uploaded block, reused only when the slot is free)
    XC_HIP(ctx, hipMemcpyAsync(h + lay.off[i_res], dres, sizeof(xc_result) * n, hipMemcpyDeviceToHost, slot.stream));
    XC_HIP(ctx, hipEventCreateWithFlags(&B->done, hipEventDisableTiming));
    XC_HIP(ctx, hipEventRecord(B->done, slot.stream));
    // the result block's host copy lives at h + lay.off[i_res]; remember it through the Batch
    B->blocks.push_back(std::make_pair(SIZE_MAX, lay.off[i_res]));
    return XC_OK;
}

// Build and launch one batch from the head of the queue (lock held).
// XC_OK: a batch was launched, or its requests completed with an error;
// XC_E_NOTREADY: nothing fits yet.
int launch_batch(xc_server *S, int slot_i)
{
    Slot &slot = S->slots[slot_i];
    auto B = std::make_unique<Batch>();
    B->slot = slot_i;
    std::vector<PairDev> P;
    std::vector<bool> ownA, ownB;
    std::vector<hipEvent_t> waits;
    std::map<std::pair<uint64_t, int>, bool> fresh;  // produced by this batch
    while (!S->q.pending.empty() && (int)B->ids.size() < S->cfg.max_batch) {
        const uint64_t id = S->q.pending.front();
        Req &r = S->q.reqs[id];
        PairDev D;
        make_pair_dev(S, r.pair, D);
        // memory: workspace + spectra not yet resident
        const auto ka = std::make_pair(r.pair.a->uid, D.Nx), kb = std::make_pair(r.pair.b->uid, D.Nx);
        size_t off_w;
        if (!alloc_evicting(S, pair_work_bytes(D), off_w)) break;
        bool ok = true;
        bool oa = false, ob = false;
        std::vector<std::pair<uint64_t, int>> mine, held;  // spectra this request created / found
        for (int side = 0; side < 2 && ok; ++side) {
            const auto key = side ? kb : ka;
            auto it = S->specs.find(key);
            if (it == S->specs.end()) {
                SpecEntry e;
                e.bytes = spec_bytes(side ? r.pair.b : r.pair.a, key.second);
                if (!alloc_evicting(S, e.bytes, e.off)) {
                    ok = false;
                    break;
                }
                e.refs = 1;  // held from now on, so later allocations cannot evict it
                S->specs[key] = e;
                fresh[key] = true;
                mine.push_back(key);
                (side ? ob : oa) = true;
                S->st.spectra_misses++;
            } else {
                it->second.refs++;
                held.push_back(key);
                if (!fresh.count(key) && it->second.ready) waits.push_back(it->second.ready);
                S->st.spectra_hits++;
            }
        }
        if (!ok) {
            for (auto &key : held) S->specs[key].refs--;
            for (auto &key : mine) {
                S->arena.release(S->specs[key].off, S->specs[key].bytes);
                S->specs.erase(key);
                fresh.erase(key);
            }
            S->arena.release(off_w, pair_work_bytes(D));
            break;
        }
        for (const auto &key : {ka, kb}) {
            S->specs[key].last_use = ++S->clock;
            B->specs.push_back(key);
        }
        B->blocks.push_back(std::make_pair(off_w, pair_work_bytes(D)));
        D.lag = S->arena.base + off_w;
        D.rows = (RowRec *)(S->arena.base + off_w + ArenaAlloc::round(sizeof(cpx<float>) * (size_t)(D.Nx / 2 + 1) * D.hR));
        D.specA = S->arena.base + S->specs[ka].off;
        D.specB = S->arena.base + S->specs[kb].off;
        P.push_back(D);
        ownA.push_back(oa);
        ownB.push_back(ob);
        B->ids.push_back(id);
        r.launched = true;
        S->q.pending.pop_front();
    }
    if (B->ids.empty()) {
        for (auto &b : B->blocks) S->arena.release(b.first, b.second);
        return XC_E_NOTREADY;
    }
    const int st = enqueue_batch(S, slot_i, B, P, ownA, ownB, waits, fresh);
    if (st == XC_E_NOTREADY) return st;
    if (st != XC_OK) {
        fail_batch(S, *B, fresh, st);
        return XC_OK;
    }
    slot.busy = true;
    S->st.batches++;
    S->inflight.push_back(std::move(B));
    return XC_OK;
}

// Deliver finished batches (lock held).  Returns true if any finished.
bool reap(xc_server *S)
{
    bool any = false;
    for (size_t k = 0; k < S->inflight.size();) {
        Batch &B = *S->inflight[k];
        const hipError_t q = hipEventQuery(B.done);
        if (q == hipErrorNotReady) {
            ++k;
            continue;
        }
        Slot &slot = S->slots[B.slot];
        size_t res_off = 0;
        for (auto &b : B.blocks)
            if (b.first == SIZE_MAX) res_off = b.second;
        const xc_result *hr = (const xc_result *)((char *)slot.pinned + res_off);
        const auto now = Clock::now();
        for (size_t i = 0; i < B.ids.size(); ++i) {
            Req &r = S->q.reqs[B.ids[i]];
            if (q == hipSuccess) {
                r.res = hr[i];
                r.latency_us = sched::us_between(r.t_submit, now);
                r.done = true;
                S->st.completed++;
            } else {
                S->q.complete(B.ids[i], XC_E_DEVICE, now);  // the batch's work failed on the device
                S->st.failed++;
            }
        }
        if (q != hipSuccess) (void)hipGetLastError();
        for (auto &b : B.blocks)
            if (b.first != SIZE_MAX) S->arena.release(b.first, b.second);
        for (auto &key : B.specs) S->specs[key].refs--;
        (void)hipEventDestroy(B.done);
        slot.busy = false;
        S->inflight.erase(S->inflight.begin() + k);
        any = true;
    }
    return any;
}

void run(xc_server *S)
{
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    Clock::time_point oldest_wait = Clock::now();
    bool waiting = false;
    for (;;) {
        bool progress = reap(S);
        const int expired = S->q.expire(Clock::now());
        if (expired) {
            S->st.timed_out += expired;
            progress = true;
        }
        if (progress) S->cv.notify_all();
        if (!S->q.pending.empty()) {
            if (!waiting) {
                waiting = true;
                oldest_wait = Clock::now();
            }
            const bool full = (int)S->q.pending.size() >= S->cfg.max_batch;
            const bool aged = sched::us_between(oldest_wait, Clock::now()) >= S->cfg.max_wait_us;
            if (full || aged || S->stop) {
                for (size_t si = 0; si < S->slots.size() && !S->q.pending.empty(); ++si) {
                    if (S->slots[si].busy) continue;
                    const int st = launch_batch(S, (int)si);
                    if (st == XC_OK) {
                        progress = true;
                        S->cv.notify_all();
                    } else if (st == XC_E_NOTREADY && S->inflight.empty()) {
                        // nothing in flight and the head request still does not fit: it never will
                        const uint64_t id = S->q.pending.front();
                        S->q.pending.pop_front();
                        S->q.complete(id, XC_E_BUDGET, Clock::now());
                        S->st.rejected++;
                        S->cv.notify_all();
                        progress = true;
                    }
                    break;
                }
                if (S->q.pending.empty()) waiting = false;
                else oldest_wait = Clock::now();
            }
        } else {
            waiting = false;
        }
        if (S->stop && S->q.pending.empty() && S->inflight.empty()) break;
        if (!progress) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(S->inflight.empty() ? 50 : 10));
            lk.lock();
        }
    }
}

}  // namespace

extern "C" int xc_server_create(xc_ctx *ctx, const xc_serve_cfg *cfg, xc_server **out)
{
    if (!ctx || !cfg || !out || cfg->hbm_budget <= 0 || cfg->max_batch <= 0 || cfg->streams <= 0 ||
        cfg->deadline_us < 0)
        return XC_E_ARG;
    *out = nullptr;
    auto S = std::make_unique<xc_server>();
    S->cfg = *cfg;
    memset(&S->st, 0, sizeof S->st);
    if (const char *f = getenv("XC_SERVE_FAIL_PINNED")) S->fail_pinned = atoi(f);
    int st = xc_ctx_create(ctx->device, &S->ctx);
    if (st) return st;
    S->own_stream = S->ctx->stream;
    if (hipMalloc(&S->arena_mem, (size_t)cfg->hbm_budget) != hipSuccess) {
        (void)hipGetLastError();
        xc_ctx_destroy(S->ctx);
        return set_err(ctx, XC_E_NOMEM, "server arena of %lld bytes", (long long)cfg->hbm_budget);
    }
    S->arena.init((char *)S->arena_mem, (size_t)cfg->hbm_budget);
    S->slots.resize(cfg->streams);
    for (auto &sl : S->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            for (auto &o : S->slots)
                if (o.stream) (void)hipStreamDestroy(o.stream);
            (void)hipFree(S->arena_mem);
            xc_ctx_destroy(S->ctx);
            return XC_E_DEVICE;
        }
    }
    xc_server *p = S.release();
    p->thr = std::thread(run, p);
    *out = p;
    return XC_OK;
}

extern "C" int xc_server_submit(xc_server *S, const xc_pair *pair, uint64_t *ticket)
{
    if (!S || !pair || !ticket || !pair->a || !pair->b) return XC_E_ARG;
    if (pair->angle_deg != 0.0 || (pair->rgn.enabled && pair->rgn.table)) return XC_E_UNSUPPORTED;
    if (pair->mode != XC_MODE_NCC && pair->mode != XC_MODE_R && pair->mode != XC_MODE_S) return XC_E_ARG;
    if (pair->a->dev.tabk != pair->b->dev.tabk) return XC_E_ARG;  // one table form per pair
    PairDev D;
    make_pair_dev(S, *pair, D);
    if (D.Nx / 2 > max_nh<float>() || D.Ny > max_ny<float>()) return XC_E_SHAPE;
    if (D.mode != PM_NCC && (D.nnegx < 6 || D.nposx < 6 || D.nnegy < 6 || D.nposy < 6)) return XC_E_SHAPE;
    const size_t need = pair_work_bytes(D) + ArenaAlloc::round(spec_bytes(pair->a, D.Nx)) +
                        ArenaAlloc::round(spec_bytes(pair->b, D.Nx)) + 65536;
    std::lock_guard<std::mutex> lk(S->mu);
    S->st.submitted++;
    if ((int64_t)need > S->cfg.hbm_budget) {
        S->st.rejected++;
        return XC_E_BUDGET;
    }
    *ticket = S->q.submit(*pair, Clock::now(), S->cfg.deadline_us);
    return XC_OK;
}

extern "C" int xc_server_cancel(xc_server *S, uint64_t ticket)
{
    if (!S) return XC_E_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    const int st = S->q.cancel(ticket, Clock::now());
    if (st == XC_OK) {
        S->st.cancelled++;
        S->cv.notify_all();
    }
    return st;
}

extern "C" int xc_server_poll(xc_server *S, uint64_t ticket, xc_result *res)
{
    if (!S) return XC_E_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->q.reqs.find(ticket);
    if (it == S->q.reqs.end()) return XC_E_ARG;
    if (!it->second.done) return XC_E_NOTREADY;
    if (res) *res = it->second.res;
    return XC_OK;
}

extern "C" int xc_server_wait(xc_server *S, uint64_t ticket, xc_result *res, double *latency_us)
{
    if (!S) return XC_E_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->q.reqs.find(ticket);
    if (it == S->q.reqs.end()) return XC_E_ARG;
    S->cv.wait(lk, [&] { return it->second.done; });
    if (res) *res = it->second.res;
    if (latency_us) *latency_us = it->second.latency_us;
    const int st = it->second.res.status;
    S->q.reqs.erase(it);
    return st < 0 ? st : XC_OK;
}

extern "C" int xc_server_stats(xc_server *S, xc_serve_stats *out)
{
    if (!S || !out) return XC_E_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->arena_bytes = (int64_t)S->arena.used;
    out->arena_peak = (int64_t)S->arena.peak;
    return XC_OK;
}

extern "C" int xc_server_destroy(xc_server *S)
{
    if (!S) return XC_E_ARG;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    if (S->thr.joinable()) S->thr.join();
    (void)hipSetDevice(S->ctx->device);
    for (auto &sl : S->slots) {
        if (sl.stream) {
            (void)hipStreamSynchronize(sl.stream);
            (void)hipStreamDestroy(sl.stream);
        }
        if (sl.pinned) (void)hipHostFree(sl.pinned);
    }
    for (auto &kv : S->specs)
        if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
    if (S->arena_mem) (void)hipFree(S->arena_mem);
    xc_ctx *c = S->ctx;
    c->stream = S->own_stream;
    delete S;
    xc_ctx_destroy(c);
    return XC_OK;
}
