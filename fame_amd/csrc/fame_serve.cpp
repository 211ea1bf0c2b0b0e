// Serving loop over one resident index: requests (chunks of single-end reads)
// are queued as they arrive, admitted into device batches while a batch fits
// the HBM budget (continuous batching), copied to the device on a stream of
// their own (three staging slots), matched by one of two compute workers
// (the context and a second context over the same index and CpG
// accumulators, each with its own batch buffers and stream: requests that
// arrive while a batch runs start at once on the other worker, and one
// worker's host waits overlap the other's kernels), and completed with their
// per-read results, counters and latency.  No
// counterpart in the reference (FAME matches a file chunk by chunk,
// main.cpp:363-398); each read's result is the one matchReads gives it,
// whatever batch it runs in.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fame_hip.h"
#include "fame_internal.h"

namespace {

using clk = std::chrono::steady_clock;

double ms_between(clk::time_point a, clk::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Req {
    uint64_t id = 0;
    const uint8_t* seqs = nullptr;
    const uint64_t* offs = nullptr;
    const uint32_t* lens = nullptr;
    uint32_t n = 0;
    uint64_t bytes = 0;
    bool contiguous = false; // reads back to back from offs[0]
    uint8_t* valid = nullptr;
    uint64_t* mat = nullptr;
    uint64_t cnt[3] = {0, 0, 0};
    clk::time_point t_arrive, t_start, t_done;
    uint32_t batch_reads = 0;
    int rc = FAME_OK;
    bool done = false;
};

struct Slot {
    uint8_t* h_seqs = nullptr; // pinned staging
    uint64_t* h_offs = nullptr;
    uint32_t* h_lens = nullptr;
    uint8_t* h_status = nullptr; // pinned results
    uint8_t* h_valid = nullptr;
    uint64_t* h_mat = nullptr;
    uint8_t* d_seqs = nullptr;
    uint64_t* d_offs = nullptr;
    uint32_t* d_lens = nullptr;
    hipEvent_t h2d = nullptr;
    std::vector<Req*> reqs;
    uint32_t n = 0;
    uint64_t bytes = 0;
};

} // namespace

constexpr int kSlots = 3;   // staging slots: one filling while two compute
constexpr int kWorkers = 2; // compute workers (contexts)

struct fame_server {
    fame_ctx* ctx = nullptr;
    fame_ctx* wctx[kWorkers] = {nullptr, nullptr}; // wctx[0] = ctx; wctx[1] owned
    int device = 0;
    hipStream_t copy = nullptr;
    uint32_t max_reads = 0;
    uint64_t max_bytes = 0;
    Slot slot[kSlots];
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req*> pending;
    std::deque<int> free_slots, ready;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread admit, compute[kWorkers];
    uint64_t batches = 0, reads = 0, requests = 0;
    double busy_ms = 0;
    std::string err;
};

namespace {

// device bytes one read of a batch may need: its letters and reverse
// complement, the 2 x 256-byte reduced rows, candidate and verify slots
// (four pools of 4 reads each, 8 bytes, plus match records), results
constexpr uint64_t BYTES_PER_READ = 1536;

void admit_loop(fame_server* s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        int si;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || (!s->pending.empty() && !s->free_slots.empty()); });
            if (s->stop) return;
            si = s->free_slots.front();
            s->free_slots.pop_front();
            Slot& sl = s->slot[si];
            sl.reqs.clear();
            sl.n = 0;
            sl.bytes = 0;
            // first come first served, as many whole requests as fit
            while (!s->pending.empty()) {
                Req* r = s->pending.front();
                if (!sl.reqs.empty() && (sl.n + r->n > s->max_reads || sl.bytes + r->bytes > s->max_bytes)) break;
                s->pending.pop_front();
                sl.reqs.push_back(r);
                sl.n += r->n;
                sl.bytes += r->bytes;
            }
        }
        Slot& sl = s->slot[si];
        uint64_t o = 0;
        uint32_t k = 0;
        hipError_t e = hipSuccess;
        for (Req* r : sl.reqs) {
            if (r->contiguous && r->n) {
                // a chunk's letters go straight from the caller's buffer
                if (e == hipSuccess)
                    e = hipMemcpyAsync(sl.d_seqs + o, r->seqs + r->offs[0], r->bytes, hipMemcpyHostToDevice,
                                       s->copy);
                const uint64_t b = o - r->offs[0];
                for (uint32_t i = 0; i < r->n; ++i) {
                    sl.h_offs[k + i] = r->offs[i] + b;
                    sl.h_lens[k + i] = r->lens[i];
                }
                o += r->bytes;
                k += r->n;
            } else {
                // scattered reads are gathered into the pinned staging area
                const uint64_t o0 = o;
                for (uint32_t i = 0; i < r->n; ++i) {
                    memcpy(sl.h_seqs + o, r->seqs + r->offs[i], r->lens[i]);
                    sl.h_offs[k] = o;
                    sl.h_lens[k] = r->lens[i];
                    o += r->lens[i];
                    ++k;
                }
                if (e == hipSuccess && o > o0)
                    e = hipMemcpyAsync(sl.d_seqs + o0, sl.h_seqs + o0, o - o0, hipMemcpyHostToDevice, s->copy);
            }
        }
        if (e == hipSuccess && k) e = hipMemcpyAsync(sl.d_offs, sl.h_offs, 8ull * k, hipMemcpyHostToDevice, s->copy);
        if (e == hipSuccess && k) e = hipMemcpyAsync(sl.d_lens, sl.h_lens, 4ull * k, hipMemcpyHostToDevice, s->copy);
        if (e == hipSuccess) e = hipEventRecord(sl.h2d, s->copy);
        std::lock_guard<std::mutex> lk(s->mu);
        if (e != hipSuccess) {
            for (Req* r : sl.reqs) {
                r->rc = FAME_ERR_HIP;
                r->done = true;
            }
            s->free_slots.push_back(si);
        } else {
            s->ready.push_back(si);
        }
        s->cv.notify_all();
    }
}

void compute_loop(fame_server* s, int wi)
{
    (void)hipSetDevice(s->device);
    fame_ctx* ctx = s->wctx[wi];
    hipStream_t st = fame::serve_stream(ctx);
    for (;;) {
        int si;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->ready.empty(); });
            if (s->stop) return;
            si = s->ready.front();
            s->ready.pop_front();
        }
        Slot& sl = s->slot[si];
        const auto t0 = clk::now();
        for (Req* r : sl.reqs) r->t_start = t0;
        int rc = hipStreamWaitEvent(st, sl.h2d, 0) == hipSuccess ? FAME_OK : FAME_ERR_HIP;
        const uint8_t* d_status = nullptr;
        const uint8_t* d_valid = nullptr;
        const uint64_t* d_mat = nullptr;
        if (rc == FAME_OK)
            rc = fame::serve_run(ctx, sl.d_seqs, sl.d_offs, sl.d_lens, sl.n, sl.bytes, st, &d_status, &d_valid,
                                 &d_mat);
        if (rc == FAME_OK && sl.n) {
            if (hipMemcpyAsync(sl.h_status, d_status, sl.n, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(sl.h_valid, d_valid, sl.n, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(sl.h_mat, d_mat, 8ull * sl.n, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                rc = FAME_ERR_HIP;
        }
        const auto t1 = clk::now();
        uint32_t k = 0;
        for (Req* r : sl.reqs) {
            r->rc = rc;
            r->batch_reads = sl.n;
            if (rc == FAME_OK) {
                for (uint32_t i = 0; i < r->n; ++i) {
                    const uint8_t stt = sl.h_status[k + i];
                    if (stt < 3) ++r->cnt[stt];
                }
                if (r->valid) memcpy(r->valid, sl.h_valid + k, r->n);
                if (r->mat) memcpy(r->mat, sl.h_mat + k, 8ull * r->n);
            }
            k += r->n;
        }
        std::lock_guard<std::mutex> lk(s->mu);
        for (Req* r : sl.reqs) {
            r->t_done = t1;
            r->done = true;
        }
        s->batches += 1;
        s->reads += sl.n;
        s->requests += sl.reqs.size();
        s->busy_ms += ms_between(t0, t1);
        s->free_slots.push_back(si);
        s->cv.notify_all();
    }
}

void release_slots(fame_server* s)
{
    for (Slot& sl : s->slot) {
        if (sl.h_seqs) (void)hipHostFree(sl.h_seqs);
        if (sl.h_offs) (void)hipHostFree(sl.h_offs);
        if (sl.h_lens) (void)hipHostFree(sl.h_lens);
        if (sl.h_status) (void)hipHostFree(sl.h_status);
        if (sl.h_valid) (void)hipHostFree(sl.h_valid);
        if (sl.h_mat) (void)hipHostFree(sl.h_mat);
        if (sl.d_seqs) (void)hipFree(sl.d_seqs);
        if (sl.d_offs) (void)hipFree(sl.d_offs);
        if (sl.d_lens) (void)hipFree(sl.d_lens);
        if (sl.h2d) (void)hipEventDestroy(sl.h2d);
        sl = Slot();
    }
    if (s->copy) (void)hipStreamDestroy(s->copy);
    s->copy = nullptr;
}

} // namespace

extern "C" {

int fame_serve_open(fame_ctx* ctx, uint64_t hbm_budget, uint32_t max_batch_reads, fame_server** out)
{
    if (!ctx || !out) return FAME_ERR_ARG;
    *out = nullptr;
    auto s = std::make_unique<fame_server>();
    s->ctx = ctx;
    s->device = fame::serve_device(ctx);
    if (hipSetDevice(s->device) != hipSuccess) return FAME_ERR_HIP;
    if (!hbm_budget) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return FAME_ERR_HIP;
        hbm_budget = fr / 2;
    }
    // the staging slots and the two workers' pipelines share the budget;
    // letters budgeted at FAME_MAX_MATE per read (longer reads take more of
    // the byte budget)
    uint64_t by_budget = hbm_budget / (kSlots * 2 * FAME_MAX_MATE + kWorkers * BYTES_PER_READ);
    uint64_t mr = max_batch_reads ? max_batch_reads : 2000000u;
    mr = std::min<uint64_t>(mr, by_budget);
    if (mr < 1024) {
        fame::serve_set_err(ctx, "HBM budget below 1024 reads per batch");
        return FAME_ERR_CAPACITY;
    }
    s->max_reads = (uint32_t)mr;
    s->max_bytes = mr * FAME_MAX_MATE;
    bool ok = hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) == hipSuccess;
    for (Slot& sl : s->slot) {
        ok = ok && hipHostMalloc((void**)&sl.h_seqs, s->max_bytes + 64, 0) == hipSuccess;
        ok = ok && hipHostMalloc((void**)&sl.h_offs, 8ull * mr, 0) == hipSuccess;
        ok = ok && hipHostMalloc((void**)&sl.h_lens, 4ull * mr, 0) == hipSuccess;
        ok = ok && hipHostMalloc((void**)&sl.h_status, mr, 0) == hipSuccess;
        ok = ok && hipHostMalloc((void**)&sl.h_valid, mr, 0) == hipSuccess;
        ok = ok && hipHostMalloc((void**)&sl.h_mat, 8ull * mr, 0) == hipSuccess;
        ok = ok && hipMalloc((void**)&sl.d_seqs, s->max_bytes + 64) == hipSuccess;
        ok = ok && hipMalloc((void**)&sl.d_offs, 8ull * mr) == hipSuccess;
        ok = ok && hipMalloc((void**)&sl.d_lens, 4ull * mr) == hipSuccess;
        ok = ok && hipEventCreateWithFlags(&sl.h2d, hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
        release_slots(s.get());
        fame::serve_set_err(ctx, "cannot allocate the serving slots");
        return FAME_ERR_HIP;
    }
    s->free_slots = {0, 1, 2};
    s->wctx[0] = ctx;
    s->wctx[1] = fame::serve_worker_open(ctx);
    fame_server* p = s.get();
    p->admit = std::thread(admit_loop, p);
    for (int w = 0; w < kWorkers; ++w)
        if (p->wctx[w]) p->compute[w] = std::thread(compute_loop, p, w);
    *out = s.release();
    return FAME_OK;
}

int fame_serve_submit(fame_server* s, const uint8_t* seqs, const uint64_t* offs, const uint32_t* lens, uint32_t n,
                      uint8_t* valid, uint64_t* mat, uint64_t* ticket)
{
    if (!s || !ticket || (n && (!seqs || !offs || !lens))) return FAME_ERR_ARG;
    auto r = std::make_unique<Req>();
    r->t_arrive = clk::now();
    r->seqs = seqs;
    r->offs = offs;
    r->lens = lens;
    r->n = n;
    r->valid = valid;
    r->mat = mat;
    bool contig = true;
    const uint32_t winlen = fame::serve_winlen(s->ctx);
    for (uint32_t i = 0; i < n; ++i) {
        if (lens[i] > FAME_MAX_READ && lens[i] <= winlen + 12) {
            fame::serve_set_err(s->ctx, "a read longer than FAME_MAX_READ could match a window");
            return FAME_ERR_ARG;
        }
        r->bytes += lens[i];
        if (i && offs[i] != offs[i - 1] + lens[i - 1]) contig = false;
    }
    r->contiguous = contig;
    if (n > s->max_reads || r->bytes > s->max_bytes) {
        fame::serve_set_err(s->ctx, "a request holds more reads or letters than one batch (max_batch_reads, "
                                    "255 letters per read)");
        return FAME_ERR_CAPACITY;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    *ticket = r->id;
    s->pending.push_back(r.get());
    s->reqs.emplace(r->id, std::move(r));
    s->cv.notify_all();
    return FAME_OK;
}

int fame_serve_wait(fame_server* s, uint64_t ticket, uint64_t counters[3], double times_ms[4])
{
    if (!s) return FAME_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return FAME_ERR_ARG;
    Req* r = it->second.get();
    s->cv.wait(lk, [&] { return r->done || s->stop; });
    if (!r->done) return FAME_ERR_ARG;
    const int rc = r->rc;
    if (counters)
        for (int k = 0; k < 3; ++k) counters[k] += r->cnt[k];
    if (times_ms) {
        times_ms[0] = ms_between(r->t_arrive, r->t_start); // queued (admission + copy)
        times_ms[1] = ms_between(r->t_start, r->t_done);   // on the GPU with its batch
        times_ms[2] = ms_between(r->t_arrive, r->t_done);  // latency
        times_ms[3] = (double)r->batch_reads;
    }
    s->reqs.erase(it);
    return rc;
}

int fame_serve_stats(fame_server* s, uint64_t out[4], double* busy_ms)
{
    if (!s || !out) return FAME_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = s->batches;
    out[1] = s->requests;
    out[2] = s->reads;
    out[3] = s->max_reads;
    if (busy_ms) *busy_ms = s->busy_ms;
    return FAME_OK;
}

int fame_serve_close(fame_server* s)
{
    if (!s) return FAME_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->admit.joinable()) s->admit.join();
    for (auto& t : s->compute)
        if (t.joinable()) t.join();
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    if (s->wctx[1]) fame::serve_worker_close(s->wctx[1]);
    release_slots(s);
    delete s;
    return FAME_OK;
}

} // extern "C"
