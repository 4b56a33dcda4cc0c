// Serving engine: continuous batching of beamformer blocks from several
// feeds (subarrays) on one GPU.
//
// A dispatcher thread admits queued blocks in arrival order whenever a slot
// (one HIP stream) is free and the block's device buffers fit in the HBM
// arena; each admitted block runs H2D -> split transpose -> coherent beams
// (-> incoherent beam) -> D2H on its slot's stream, so blocks of different
// feeds overlap on the device.  A host callback at the end of each block's
// stream work returns its slot and memory and records its latency; no HIP
// call is made from the callback.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fbf_internal.hpp"

using namespace fbf;

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
constexpr size_t kMaxPending = 4096;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device allocation, with coalescing.
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    size_t used = 0;
    size_t peak = 0;
    std::map<size_t, size_t> free_;   // offset -> length

    void init(char* b, size_t n)
    {
        base = b;
        size = n;
        free_.clear();
        free_[0] = n;
        used = peak = 0;
    }
    bool alloc(size_t n, size_t& off)
    {
        n = align_up(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                used += n;
                peak = std::max(peak, used);
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = align_up(n);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Feed {
    fbf_feed_desc d{};
    fbf_cb_plan* cb = nullptr;
    fbf_ib_plan* ib = nullptr;
    fbf_weights_plan* wplan = nullptr;
    size_t block_bytes = 0, cb_bytes = 0, ib_bytes = 0, weights_bytes = 0;
    void* weights[2] = {nullptr, nullptr};
    int current = -1;          // weights buffer new blocks use; -1 = none yet
    int inflight[2] = {0, 0};  // running blocks per weights buffer
};

struct Request {
    fbf_server* server = nullptr;
    int feed = 0;
    const void* input = nullptr;
    bool on_device = false;
    void* cb_out = nullptr;
    void* ib_out = nullptr;
    uint64_t tag = 0;
    Clock::time_point t_submit;
    int slot = -1;
    size_t arena_off = 0, arena_bytes = 0;
    bool has_arena = false;
    int wbuf = -1;
    int status = FBF_OK;
    double latency_ms = 0.0;
};

}  // namespace

struct fbf_server {
    int nslots = 0;
    size_t budget = 0;
    Arena arena;
    void* arena_mem = nullptr;
    std::vector<Feed> feeds;
    std::vector<hipStream_t> slot_streams;
    std::vector<bool> slot_busy;
    hipStream_t aux = nullptr;   // weights uploads / generation
    std::deque<Request*> pending;
    std::deque<Request*> done;
    std::mutex mu;
    std::condition_variable cv;
    std::thread dispatcher;
    bool started = false;
    bool stopping = false;
    uint64_t submitted = 0, completed = 0, failed = 0, in_flight = 0;
    double input_bytes = 0.0;
    std::vector<double> latencies;
    Clock::time_point t_start;
};

namespace {

void finish_locked(fbf_server* s, Request* r)
{
    Feed& f = s->feeds[r->feed];
    if (r->slot >= 0) s->slot_busy[r->slot] = false;
    if (r->has_arena) s->arena.release(r->arena_off, r->arena_bytes);
    if (r->wbuf >= 0) --f.inflight[r->wbuf];
    r->latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - r->t_submit).count();
    --s->in_flight;
    if (r->status == FBF_OK) {
        ++s->completed;
        s->latencies.push_back(r->latency_ms);
        s->input_bytes += (double)f.block_bytes;
    } else {
        ++s->failed;
    }
    s->done.push_back(r);
}

void block_done(void* p)
{
    Request* r = static_cast<Request*>(p);
    fbf_server* s = r->server;
    std::lock_guard<std::mutex> lock(s->mu);
    finish_locked(s, r);
    s->cv.notify_all();
}

// Enqueues one admitted block on its slot stream (called without the lock).
int enqueue_block(fbf_server* s, Request* r)
{
    Feed& f = s->feeds[r->feed];
    hipStream_t st = s->slot_streams[r->slot];
    char* mem = s->arena.base + r->arena_off;
    const void* input = r->input;
    size_t off = 0;
    if (!r->on_device) {
        FBF_HIP_TRY(hipMemcpyAsync(mem, r->input, f.block_bytes, hipMemcpyHostToDevice, st));
        input = mem;
        off += align_up(f.block_bytes);
    }
    void* ftpa = mem + off;
    off += align_up(f.block_bytes);
    void* cb_dev = mem + off;
    off += align_up(f.cb_bytes);
    void* cb_stage = mem + off;
    off += align_up(f.cb_bytes);
    void* ib_dev = mem + off;
    const bool fused = f.ib && f.d.ib_fscrunch == 1 && 32 % f.d.ib_tscrunch == 0;
    int e = fused ? fbf_split_transpose_ib(input, f.block_bytes, ftpa, f.block_bytes, f.d.nchans,
                                           f.d.nsamples_per_heap, f.ib, ib_dev, f.ib_bytes, st)
                  : fbf_split_transpose(input, f.block_bytes, ftpa, f.block_bytes, f.d.nantennas, f.d.nantennas, 0,
                                        f.d.nchans, f.d.nsamples_per_heap, st);
    if (e) return e;
    e = fbf_cb_beamform_ws(f.cb, ftpa, f.block_bytes, f.weights[r->wbuf], f.weights_bytes, cb_dev, f.cb_bytes,
                           cb_stage, f.cb_bytes, st);
    if (e) return e;
    if (f.ib) {
        if (!fused) {
            e = fbf_ib_beamform(f.ib, input, f.block_bytes, ib_dev, f.ib_bytes, st);
            if (e) return e;
        }
        if (r->ib_out) FBF_HIP_TRY(hipMemcpyAsync(r->ib_out, ib_dev, f.ib_bytes, hipMemcpyDeviceToHost, st));
    }
    FBF_HIP_TRY(hipMemcpyAsync(r->cb_out, cb_dev, f.cb_bytes, hipMemcpyDeviceToHost, st));
    FBF_HIP_TRY(hipLaunchHostFunc(st, block_done, r));
    return FBF_OK;
}

size_t block_device_bytes(const Feed& f, bool on_device)
{
    return (on_device ? 0 : align_up(f.block_bytes)) + align_up(f.block_bytes) + 2 * align_up(f.cb_bytes) +
           align_up(f.ib_bytes);
}

void dispatch_loop(fbf_server* s)
{
    std::unique_lock<std::mutex> lock(s->mu);
    while (true) {
        Request* r = nullptr;
        int slot = -1;
        s->cv.wait(lock, [&] {
            if (s->stopping) return true;
            if (s->pending.empty()) return false;
            for (int i = 0; i < s->nslots; ++i)
                if (!s->slot_busy[i]) {
                    slot = i;
                    break;
                }
            if (slot < 0) return false;
            Request* front = s->pending.front();
            const Feed& f = s->feeds[front->feed];
            size_t off = 0;
            if (!s->arena.alloc(block_device_bytes(f, front->on_device), off)) return false;
            front->arena_off = off;
            front->arena_bytes = block_device_bytes(f, front->on_device);
            front->has_arena = true;
            return true;
        });
        if (s->stopping) break;
        r = s->pending.front();
        s->pending.pop_front();
        Feed& f = s->feeds[r->feed];
        r->slot = slot;
        s->slot_busy[slot] = true;
        r->wbuf = f.current;
        ++f.inflight[r->wbuf];
        ++s->in_flight;
        lock.unlock();
        const int e = enqueue_block(s, r);
        lock.lock();
        if (e != FBF_OK) {
            // The stream may hold part of the block's work: wait for it before
            // the buffers go back to the arena.
            lock.unlock();
            (void)hipStreamSynchronize(s->slot_streams[slot]);
            lock.lock();
            r->status = e;
            finish_locked(s, r);
            s->cv.notify_all();
        }
    }
}

int check_feed(const fbf_server* s, int feed)
{
    FBF_REQUIRE(s && feed >= 0 && feed < (int)s->feeds.size(), FBF_ERR_ARG, "fbf_server: bad feed");
    return FBF_OK;
}

}  // namespace

extern "C" int fbf_server_create(int nslots, size_t hbm_budget, fbf_server** out)
{
    FBF_REQUIRE(out, FBF_ERR_ARG, "fbf_server_create: null output");
    *out = nullptr;
    FBF_REQUIRE(nslots > 0 && nslots <= 64 && hbm_budget > 0, FBF_ERR_ARG, "fbf_server_create: bad arguments");
    auto* s = new fbf_server;
    s->nslots = nslots;
    s->budget = hbm_budget;
    *out = s;
    return FBF_OK;
}

extern "C" int fbf_server_add_feed(fbf_server* s, const fbf_feed_desc* d, const float* cb_offsets,
                                   const float* cb_scales, const float* ib_offsets, const float* ib_scales, int* feed)
{
    FBF_REQUIRE(s && d && feed && cb_offsets && cb_scales, FBF_ERR_ARG, "fbf_server_add_feed: null argument");
    FBF_REQUIRE(!s->started, FBF_ERR_STATE, "fbf_server_add_feed: add feeds before fbf_server_start");
    FBF_REQUIRE(d->nheap_groups > 0 && d->nsamples_per_heap > 0, FBF_ERR_ARG, "fbf_server_add_feed: bad block");
    Feed f;
    f.d = *d;
    fbf_cb_desc cd{d->nantennas, d->nbeams, d->nchans, d->tscrunch, d->fscrunch, d->nsamples_per_heap_out,
                   d->quant_mode};
    int st = fbf_cb_plan_create(&cd, cb_offsets, cb_scales, &f.cb);
    if (st) return st;
    f.block_bytes = (size_t)4 * d->nantennas * d->nchans * d->nsamples_per_heap * d->nheap_groups;
    st = fbf_cb_output_bytes(f.cb, f.block_bytes, &f.cb_bytes);
    if (st == FBF_OK && d->ib_tscrunch > 0) {
        FBF_REQUIRE(ib_offsets && ib_scales, FBF_ERR_ARG, "fbf_server_add_feed: null incoherent-beam levels");
        fbf_ib_desc id{d->nantennas, d->nchans, d->nsamples_per_heap, d->ib_tscrunch, d->ib_fscrunch, d->quant_mode};
        st = fbf_ib_plan_create(&id, ib_offsets, ib_scales, &f.ib);
        if (st == FBF_OK) st = fbf_ib_output_bytes(f.ib, f.block_bytes, &f.ib_bytes);
    }
    f.weights_bytes = (size_t)2 * d->nchans * d->nbeams * d->nantennas;
    hipError_t e = hipSuccess;
    if (st == FBF_OK) e = hipMalloc(&f.weights[0], f.weights_bytes);
    if (st == FBF_OK && e == hipSuccess) e = hipMalloc(&f.weights[1], f.weights_bytes);
    if (st != FBF_OK || e != hipSuccess) {
        if (f.cb) fbf_cb_plan_destroy(f.cb);
        if (f.ib) fbf_ib_plan_destroy(f.ib);
        if (f.weights[0]) (void)hipFree(f.weights[0]);
        if (f.weights[1]) (void)hipFree(f.weights[1]);
        return st != FBF_OK ? st : hip_status(e, "fbf_server_add_feed");
    }
    s->feeds.push_back(f);
    *feed = (int)s->feeds.size() - 1;
    return FBF_OK;
}

extern "C" int fbf_server_feed_bytes(const fbf_server* s, int feed, size_t* block_bytes, size_t* cb_bytes,
                                     size_t* ib_bytes)
{
    int st = check_feed(s, feed);
    if (st) return st;
    const Feed& f = s->feeds[feed];
    if (block_bytes) *block_bytes = f.block_bytes;
    if (cb_bytes) *cb_bytes = f.cb_bytes;
    if (ib_bytes) *ib_bytes = f.ib_bytes;
    return FBF_OK;
}

// Picks the weights buffer that no running block uses, fills it with `fill`
// on the auxiliary stream, then makes it current.
template <class Fill>
static int update_weights(fbf_server* s, int feed, Fill fill)
{
    Feed& f = s->feeds[feed];
    int target;
    {
        std::unique_lock<std::mutex> lock(s->mu);
        target = f.current < 0 ? 0 : 1 - f.current;
        s->cv.wait(lock, [&] { return f.inflight[target] == 0; });
    }
    if (!s->aux) FBF_HIP_TRY(hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking));
    int st = fill(f.weights[target], s->aux);
    if (st) return st;
    FBF_HIP_TRY(hipStreamSynchronize(s->aux));
    std::lock_guard<std::mutex> lock(s->mu);
    f.current = target;
    return FBF_OK;
}

extern "C" int fbf_server_set_weights(fbf_server* s, int feed, const void* weights_host, size_t bytes)
{
    int st = check_feed(s, feed);
    if (st) return st;
    FBF_REQUIRE(weights_host && bytes == s->feeds[feed].weights_bytes, FBF_ERR_SIZE,
                "fbf_server_set_weights: weights size mismatch");
    return update_weights(s, feed, [&](void* dst, hipStream_t aux) -> int {
        FBF_HIP_TRY(hipMemcpyAsync(dst, weights_host, bytes, hipMemcpyHostToDevice, aux));
        return FBF_OK;
    });
}

extern "C" int fbf_server_set_delays(fbf_server* s, int feed, const void* delays_host, size_t bytes,
                                     const double* chan_freqs, double epoch)
{
    int st = check_feed(s, feed);
    if (st) return st;
    Feed& f = s->feeds[feed];
    FBF_REQUIRE(delays_host && chan_freqs, FBF_ERR_ARG, "fbf_server_set_delays: null argument");
    FBF_REQUIRE(bytes == (size_t)8 * f.d.nbeams * f.d.nantennas, FBF_ERR_SIZE, "fbf_server_set_delays: size mismatch");
    if (f.wplan) fbf_weights_plan_destroy(f.wplan);
    f.wplan = nullptr;
    st = fbf_weights_plan_create(f.d.nantennas, f.d.nbeams, f.d.nchans, chan_freqs, &f.wplan);
    if (st) return st;
    void* d_delays = nullptr;
    FBF_HIP_TRY(hipMalloc(&d_delays, bytes));
    st = update_weights(s, feed, [&](void* dst, hipStream_t aux) -> int {
        FBF_HIP_TRY(hipMemcpyAsync(d_delays, delays_host, bytes, hipMemcpyHostToDevice, aux));
        return fbf_weights_generate(f.wplan, d_delays, bytes, epoch, dst, f.weights_bytes, nullptr, aux);
    });
    (void)hipFree(d_delays);
    return st;
}

extern "C" int fbf_server_start(fbf_server* s)
{
    FBF_REQUIRE(s, FBF_ERR_ARG, "fbf_server_start: null server");
    FBF_REQUIRE(!s->started, FBF_ERR_STATE, "fbf_server_start: already started");
    FBF_REQUIRE(!s->feeds.empty(), FBF_ERR_STATE, "fbf_server_start: no feeds");
    size_t largest = 0;
    for (auto const& f : s->feeds) largest = std::max(largest, block_device_bytes(f, false));
    FBF_REQUIRE(largest <= s->budget, FBF_ERR_CAPACITY,
                "fbf_server_start: the HBM budget cannot hold one block of the largest feed (" +
                    std::to_string(largest) + " bytes)");
    size_t free_b = 0, total_b = 0;
    FBF_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    FBF_REQUIRE(s->budget + (size_t(1) << 30) <= free_b, FBF_ERR_CAPACITY,
                "fbf_server_start: budget exceeds free device memory");
    FBF_HIP_TRY(hipMalloc(&s->arena_mem, s->budget));
    s->arena.init(static_cast<char*>(s->arena_mem), s->budget);
    for (int i = 0; i < s->nslots; ++i) {
        hipStream_t st = nullptr;
        FBF_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        s->slot_streams.push_back(st);
        s->slot_busy.push_back(false);
    }
    s->t_start = Clock::now();
    s->started = true;
    s->dispatcher = std::thread(dispatch_loop, s);
    return FBF_OK;
}

extern "C" int fbf_server_submit(fbf_server* s, int feed, const void* input, size_t bytes, int input_on_device,
                                 void* cb_out, void* ib_out, uint64_t tag)
{
    int st = check_feed(s, feed);
    if (st) return st;
    FBF_REQUIRE(s->started, FBF_ERR_STATE, "fbf_server_submit: server not started");
    const Feed& f = s->feeds[feed];
    FBF_REQUIRE(input && cb_out, FBF_ERR_ARG, "fbf_server_submit: null buffer");
    FBF_REQUIRE(bytes == f.block_bytes, FBF_ERR_SIZE,
                "fbf_server_submit: block size " + std::to_string(bytes) + " != " + std::to_string(f.block_bytes));
    FBF_REQUIRE(!f.ib || ib_out, FBF_ERR_ARG, "fbf_server_submit: feed has an incoherent beam: ib_out needed");
    auto* r = new Request;
    r->server = s;
    r->feed = feed;
    r->input = input;
    r->on_device = input_on_device != 0;
    r->cb_out = cb_out;
    r->ib_out = ib_out;
    r->tag = tag;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lock(s->mu);
    if (f.current < 0) {
        delete r;
        return set_error(FBF_ERR_STATE, "fbf_server_submit: feed has no weights yet");
    }
    if (s->pending.size() >= kMaxPending) {
        delete r;
        return set_error(FBF_ERR_CAPACITY, "fbf_server_submit: queue full");
    }
    s->pending.push_back(r);
    ++s->submitted;
    s->cv.notify_all();
    return FBF_OK;
}

extern "C" int fbf_server_poll(fbf_server* s, uint64_t* tag, int* status, double* latency_ms, int timeout_ms)
{
    FBF_REQUIRE(s && tag && status && latency_ms, FBF_ERR_ARG, "fbf_server_poll: null argument");
    std::unique_lock<std::mutex> lock(s->mu);
    if (!s->cv.wait_for(lock, std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms),
                        [&] { return !s->done.empty(); }))
        return set_error(FBF_ERR_STATE, "fbf_server_poll: nothing completed");
    Request* r = s->done.front();
    s->done.pop_front();
    *tag = r->tag;
    *status = r->status;
    *latency_ms = r->latency_ms;
    delete r;
    return FBF_OK;
}

extern "C" int fbf_server_stats_get(fbf_server* s, fbf_server_stats* out)
{
    FBF_REQUIRE(s && out, FBF_ERR_ARG, "fbf_server_stats_get: null argument");
    std::vector<double> lat;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        lat = s->latencies;
        out->submitted = s->submitted;
        out->completed = s->completed;
        out->failed = s->failed;
        out->in_flight = s->in_flight;
        out->pending = s->pending.size();
        out->input_bytes = s->input_bytes;
        out->hbm_budget = s->budget;
        out->hbm_peak = s->arena.peak;
        out->wall_s = s->started ? std::chrono::duration<double>(Clock::now() - s->t_start).count() : 0.0;
    }
    auto pct = [&](double q) {
        if (lat.empty()) return 0.0;
        std::vector<double> v = lat;
        const size_t k = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
        std::nth_element(v.begin(), v.begin() + k, v.end());
        return v[k];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->max_ms = lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end());
    double sum = 0.0;
    for (double x : lat) sum += x;
    out->mean_ms = lat.empty() ? 0.0 : sum / (double)lat.size();
    return FBF_OK;
}

extern "C" int fbf_server_reset_stats(fbf_server* s)
{
    FBF_REQUIRE(s, FBF_ERR_ARG, "fbf_server_reset_stats: null server");
    std::lock_guard<std::mutex> lock(s->mu);
    s->latencies.clear();
    s->completed = s->failed = 0;
    s->submitted = s->in_flight + s->pending.size();
    s->input_bytes = 0.0;
    s->arena.peak = s->arena.used;
    s->t_start = Clock::now();
    return FBF_OK;
}

extern "C" int fbf_server_drain(fbf_server* s)
{
    FBF_REQUIRE(s, FBF_ERR_ARG, "fbf_server_drain: null server");
    std::unique_lock<std::mutex> lock(s->mu);
    s->cv.wait(lock, [&] { return s->pending.empty() && s->in_flight == 0; });
    return FBF_OK;
}

extern "C" int fbf_server_destroy(fbf_server* s)
{
    if (!s) return FBF_OK;
    if (s->started) {
        fbf_server_drain(s);
        {
            std::lock_guard<std::mutex> lock(s->mu);
            s->stopping = true;
            s->cv.notify_all();
        }
        s->dispatcher.join();
        for (auto st : s->slot_streams) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }
    for (Request* r : s->pending) delete r;
    for (Request* r : s->done) delete r;
    for (auto& f : s->feeds) {
        if (f.cb) fbf_cb_plan_destroy(f.cb);
        if (f.ib) fbf_ib_plan_destroy(f.ib);
        if (f.wplan) fbf_weights_plan_destroy(f.wplan);
        if (f.weights[0]) (void)hipFree(f.weights[0]);
        if (f.weights[1]) (void)hipFree(f.weights[1]);
    }
    if (s->aux) (void)hipStreamDestroy(s->aux);
    if (s->arena_mem) (void)hipFree(s->arena_mem);
    delete s;
    return FBF_OK;
}
