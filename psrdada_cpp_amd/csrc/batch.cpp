// Batched subarrays on one DADA block.
//
// One split transpose of the whole block into a full-antenna FTPA workspace,
// then every subarray's coherent beamformer reads its antenna range straight
// from that workspace (fbf_cb_beamform_subarray), and the incoherent beam
// reads the TAFTP block.  The beamformers are independent branches forked
// from the caller's stream and joined back, so uneven subarrays share the
// device; with use_graph the whole block is one hipGraph launch.
#include <cstring>
#include <string>
#include <vector>

#include "fbf_internal.hpp"

using namespace fbf;

struct fbf_batch {
    int A, F, S, nhg;
    std::vector<fbf_subarray_desc> subs;
    std::vector<fbf_cb_plan*> plans;
    fbf_ib_plan* ib = nullptr;
    bool ib_fused = false;         // incoherent beam computed by the split-transpose pass
    void* ftpa = nullptr;          // full-antenna FTPA workspace
    size_t ftpa_bytes = 0;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> join;
    hipEvent_t fork_ib = nullptr;   // block available: incoherent beam may start
    hipEvent_t fork_cb = nullptr;   // transpose done: coherent beams may start
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const void*> graph_ptrs;   // pointers the graph was captured with
    hipStream_t graph_stream = nullptr;
    // Every kernel of a block in turn on the caller's stream, instead of the
    // subarrays side by side on their own streams.  On this GPU a
    // cross-stream event wait costs 15-20 us before the waiting stream
    // moves on (measured between the transpose and the beamformers, and
    // between a block's join and the next block), while subarrays side by
    // side take about as long as in turn; fbf_batch_autotune times both.
    bool serial = true;
};

static size_t sub_out_bytes(const fbf_batch* b, const fbf_subarray_desc& d)
{
    const size_t T = (size_t)b->nhg * b->S;
    return T / d.tscrunch * d.nbeams * (b->F / d.fscrunch);
}

static void batch_release(fbf_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto* p : b->plans) fbf_cb_plan_destroy(p);
    if (b->ib) fbf_ib_plan_destroy(b->ib);
    if (b->ftpa) (void)hipFree(b->ftpa);
    for (auto s : b->streams) (void)hipStreamDestroy(s);
    for (auto e : b->join) (void)hipEventDestroy(e);
    if (b->fork_ib) (void)hipEventDestroy(b->fork_ib);
    if (b->fork_cb) (void)hipEventDestroy(b->fork_cb);
    delete b;
}

extern "C" int fbf_batch_create(int total_nantennas, int nchans, int nsamples_per_heap, int nheap_groups, int nsub,
                                const fbf_subarray_desc* subs, const float* const* offsets,
                                const float* const* scales, fbf_batch** out)
{
    FBF_REQUIRE(out, FBF_ERR_ARG, "fbf_batch_create: null output");
    *out = nullptr;
    FBF_REQUIRE(total_nantennas > 0 && nchans > 0 && nsamples_per_heap > 0 && nheap_groups > 0 && nsub > 0 &&
                    nsub <= 64 && subs && offsets && scales,
                FBF_ERR_ARG, "fbf_batch_create: bad argument");
    auto* b = new fbf_batch;
    b->A = total_nantennas;
    b->F = nchans;
    b->S = nsamples_per_heap;
    b->nhg = nheap_groups;
    b->subs.assign(subs, subs + nsub);
    for (int i = 0; i < nsub; ++i) {
        const fbf_subarray_desc& s = subs[i];
        if (s.start_antenna < 0 || s.nantennas <= 0 || s.start_antenna + s.nantennas > total_nantennas) {
            batch_release(b);
            return set_error(FBF_ERR_ARG, "fbf_batch_create: subarray " + std::to_string(i) +
                                              " outside the block's antennas");
        }
        if (s.nantennas % 8 == 0 && (2 * s.start_antenna) % 16 != 0) {
            batch_release(b);
            return set_error(FBF_ERR_UNSUPPORTED,
                             "fbf_batch_create: a subarray of a multiple of 8 antennas must start at "
                             "an antenna offset that is a multiple of 8");
        }
        fbf_cb_desc d{s.nantennas, s.nbeams, nchans, s.tscrunch, s.fscrunch, s.nsamples_per_heap_out, s.quant_mode};
        fbf_cb_plan* p = nullptr;
        int st = fbf_cb_plan_create(&d, offsets[i], scales[i], &p);
        if (st) {
            batch_release(b);
            return st;
        }
        b->plans.push_back(p);
        // The graph captures the plan's staging buffer: size it now.
        st = fbf_cb_plan_reserve(p, (long long)nsamples_per_heap * nheap_groups);
        if (st) {
            batch_release(b);
            return st;
        }
    }
    b->ftpa_bytes = (size_t)4 * total_nantennas * nchans * nsamples_per_heap * nheap_groups;
    hipError_t e = hipMalloc(&b->ftpa, b->ftpa_bytes);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork_ib, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork_cb, hipEventDisableTiming);
    for (int i = 0; i <= nsub && e == hipSuccess; ++i) {
        hipStream_t s = nullptr;
        hipEvent_t j = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess) b->streams.push_back(s);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
        if (e == hipSuccess) b->join.push_back(j);
    }
    if (e != hipSuccess) {
        batch_release(b);
        return hip_status(e, "fbf_batch_create");
    }
    *out = b;
    return FBF_OK;
}

extern "C" int fbf_batch_destroy(fbf_batch* b)
{
    if (b) batch_release(b);
    return FBF_OK;
}

extern "C" int fbf_batch_enable_ib(fbf_batch* b, int tscrunch, int fscrunch, int quant_mode, const float* offsets,
                                   const float* scales)
{
    FBF_REQUIRE(b, FBF_ERR_ARG, "fbf_batch_enable_ib: null batch");
    fbf_ib_desc d{b->A, b->F, b->S, tscrunch, fscrunch, quant_mode};
    fbf_ib_plan* p = nullptr;
    int st = fbf_ib_plan_create(&d, offsets, scales, &p);
    if (st) return st;
    if (b->ib) fbf_ib_plan_destroy(b->ib);
    b->ib = p;
    b->ib_fused = fscrunch == 1 && tscrunch > 0 && 32 % tscrunch == 0 && b->S % 32 == 0;
    b->graph_ptrs.clear();   // forces a new capture
    return FBF_OK;
}

extern "C" int fbf_batch_output_bytes(const fbf_batch* b, int index, size_t* bytes)
{
    FBF_REQUIRE(b && bytes, FBF_ERR_ARG, "fbf_batch_output_bytes: null argument");
    const int n = (int)b->subs.size();
    FBF_REQUIRE(index >= 0 && (index < n || (index == n && b->ib)), FBF_ERR_ARG,
                "fbf_batch_output_bytes: bad index");
    if (index < n) {
        *bytes = sub_out_bytes(b, b->subs[index]);
    } else {
        return fbf_ib_output_bytes(b->ib, b->ftpa_bytes, bytes);
    }
    return FBF_OK;
}

extern "C" int fbf_batch_weights_bytes(const fbf_batch* b, int sub, size_t* bytes)
{
    FBF_REQUIRE(b && bytes && sub >= 0 && sub < (int)b->subs.size(), FBF_ERR_ARG,
                "fbf_batch_weights_bytes: bad argument");
    *bytes = (size_t)2 * b->F * b->subs[sub].nbeams * b->subs[sub].nantennas;
    return FBF_OK;
}

extern "C" int fbf_batch_set_variant(fbf_batch* b, int sub, int variant, int chunk)
{
    FBF_REQUIRE(b && sub >= 0 && sub < (int)b->subs.size(), FBF_ERR_ARG, "fbf_batch_set_variant: bad argument");
    int st = fbf_cb_plan_set_variant(b->plans[sub], variant);
    if (st == FBF_OK) st = fbf_cb_plan_set_chunk(b->plans[sub], chunk);
    b->graph_ptrs.clear();
    return st;
}

extern "C" int fbf_batch_set_serial(fbf_batch* b, int serial)
{
    FBF_REQUIRE(b, FBF_ERR_ARG, "fbf_batch_set_serial: null batch");
    b->serial = serial != 0;
    b->graph_ptrs.clear();   // forces a new capture
    return FBF_OK;
}

extern "C" int fbf_batch_get_serial(const fbf_batch* b, int* serial)
{
    FBF_REQUIRE(b && serial, FBF_ERR_ARG, "fbf_batch_get_serial: null argument");
    *serial = b->serial ? 1 : 0;
    return FBF_OK;
}

// Tunes every subarray's kernel (fbf_cb_autotune), then times whole blocks
// as graph launches with the subarrays in turn and side by side, on
// synthetic data of the batch's geometry.  In turn is kept unless side by
// side is faster by more than 3 %: the synthetic bytes run at a higher
// clock than real voltages, and on real data in turn was the faster at
// every tscrunch measured.
extern "C" int fbf_batch_autotune(fbf_batch* b, int reps)
{
    FBF_REQUIRE(b && reps > 0, FBF_ERR_ARG, "fbf_batch_autotune: bad argument");
    for (auto* p : b->plans) {
        int st = fbf_cb_autotune(p, (long long)b->nhg * b->S, reps, nullptr, nullptr, nullptr);
        if (st) return st;
    }
    const int n = (int)b->subs.size();
    const bool serial0 = b->serial;   // kept if the timing fails
    std::vector<void*> bufs;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        for (void* p : bufs) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
        b->graph_ptrs.clear();   // the next run captures its own pointers
        b->graph_stream = nullptr;
    };
    auto alloc = [&](size_t bytes, int fill, void** p) {
        hipError_t e = hipMalloc(p, bytes);
        if (e == hipSuccess) {
            bufs.push_back(*p);
            e = hipMemsetAsync(*p, fill, bytes, s);
        }
        return e;
    };
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    void* taftp = nullptr;
    if (e == hipSuccess) e = alloc(b->ftpa_bytes, 0x11, &taftp);
    std::vector<const void*> w(n);
    std::vector<void*> o(n + 1, nullptr);
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        size_t wb = 0, ob = 0;
        fbf_batch_weights_bytes(b, i, &wb);
        fbf_batch_output_bytes(b, i, &ob);
        void* p = nullptr;
        e = alloc(wb, 0x5A, &p);
        w[i] = p;
        if (e == hipSuccess) e = alloc(ob, 0, &o[i]);
    }
    if (e == hipSuccess && b->ib) {
        size_t ob = 0;
        fbf_batch_output_bytes(b, n, &ob);
        e = alloc(ob, 0, &o[n]);
    }
    if (e != hipSuccess) {
        cleanup();
        return hip_status(e, "fbf_batch_autotune setup");
    }
    const int nrep = reps < 8 ? 8 : reps;
    float mode_ms[2] = {0.0f, 0.0f};   // [0] in turn, [1] side by side
    for (int mode = 0; mode < 2; ++mode) {
        b->serial = mode == 0;
        b->graph_ptrs.clear();
        int st = FBF_OK;
        for (int r = 0; r < 2 && st == FBF_OK; ++r) st = fbf_batch_run(b, taftp, b->ftpa_bytes, w.data(), o.data(), 1, s);
        if (st == FBF_OK) e = hipEventRecord(e0, s);
        for (int r = 0; r < nrep && st == FBF_OK && e == hipSuccess; ++r)
            st = fbf_batch_run(b, taftp, b->ftpa_bytes, w.data(), o.data(), 1, s);
        if (st == FBF_OK && e == hipSuccess) e = hipEventRecord(e1, s);
        if (st == FBF_OK && e == hipSuccess) e = hipEventSynchronize(e1);
        if (st != FBF_OK || e != hipSuccess) {
            (void)hipStreamSynchronize(s);
            b->serial = serial0;
            cleanup();
            return st != FBF_OK ? st : hip_status(e, "fbf_batch_autotune timing");
        }
        (void)hipEventElapsedTime(&mode_ms[mode], e0, e1);
    }
    b->serial = !(mode_ms[1] < 0.97f * mode_ms[0]);
    cleanup();
    return FBF_OK;
}

// Enqueues one block in turn on `stream`: incoherent beam (when not fused),
// transpose (with the fused incoherent beam), then each subarray.
static int batch_enqueue_serial(fbf_batch* b, const void* taftp, size_t taftp_bytes, const void* const* weights,
                                void* const* outputs, hipStream_t stream)
{
    const int n = (int)b->subs.size();
    int st = FBF_OK;
    if (b->ib) {
        size_t ob = 0;
        st = fbf_ib_output_bytes(b->ib, taftp_bytes, &ob);
        if (st) return st;
        st = b->ib_fused ? fbf_split_transpose_ib(taftp, taftp_bytes, b->ftpa, b->ftpa_bytes, b->F, b->S, b->ib,
                                                  outputs[n], ob, stream)
                         : fbf_ib_beamform(b->ib, taftp, taftp_bytes, outputs[n], ob, stream);
        if (st) return st;
    }
    if (!b->ib || !b->ib_fused) {
        st = fbf_split_transpose(taftp, taftp_bytes, b->ftpa, b->ftpa_bytes, b->A, b->A, 0, b->F, b->S, stream);
        if (st) return st;
    }
    for (int i = 0; i < n; ++i) {
        size_t wb = 0;
        fbf_batch_weights_bytes(b, i, &wb);
        st = fbf_cb_beamform_subarray(b->plans[i], b->ftpa, b->ftpa_bytes, b->A, b->subs[i].start_antenna, weights[i],
                                      wb, outputs[i], sub_out_bytes(b, b->subs[i]), stream);
        if (st) return st;
    }
    return FBF_OK;
}

// Enqueues one block: transpose, fork, beamformers, join.
static int batch_enqueue(fbf_batch* b, const void* taftp, size_t taftp_bytes, const void* const* weights,
                         void* const* outputs, hipStream_t stream)
{
    if (b->serial) return batch_enqueue_serial(b, taftp, taftp_bytes, weights, outputs, stream);
    const int n = (int)b->subs.size();
    // The incoherent beam comes out of the split-transpose pass when its
    // scrunch allows (fbf_split_transpose_ib), else it runs on its own
    // branch (it only reads the TAFTP block).
    bool ib_branch = false;
    int st = FBF_OK;
    if (b->ib) {
        size_t ob = 0;
        st = fbf_ib_output_bytes(b->ib, taftp_bytes, &ob);
        if (st) return st;
        if (b->ib_fused) {
            st = fbf_split_transpose_ib(taftp, taftp_bytes, b->ftpa, b->ftpa_bytes, b->F, b->S, b->ib, outputs[n],
                                        ob, stream);
            if (st) return st;
        } else {
            ib_branch = true;
            FBF_HIP_TRY(hipEventRecord(b->fork_ib, stream));
            hipStream_t s = b->streams[n];
            FBF_HIP_TRY(hipStreamWaitEvent(s, b->fork_ib, 0));
            st = fbf_ib_beamform(b->ib, taftp, taftp_bytes, outputs[n], ob, s);
            if (st) return st;
            FBF_HIP_TRY(hipEventRecord(b->join[n], s));
        }
    }
    if (!b->ib || !b->ib_fused) {
        st = fbf_split_transpose(taftp, taftp_bytes, b->ftpa, b->ftpa_bytes, b->A, b->A, 0, b->F, b->S, stream);
        if (st) return st;
    }
    FBF_HIP_TRY(hipEventRecord(b->fork_cb, stream));
    for (int i = 0; i < n; ++i) {
        hipStream_t s = b->streams[i];
        FBF_HIP_TRY(hipStreamWaitEvent(s, b->fork_cb, 0));
        size_t wb = 0;
        fbf_batch_weights_bytes(b, i, &wb);
        st = fbf_cb_beamform_subarray(b->plans[i], b->ftpa, b->ftpa_bytes, b->A, b->subs[i].start_antenna, weights[i],
                                      wb, outputs[i], sub_out_bytes(b, b->subs[i]), s);
        if (st) return st;
        FBF_HIP_TRY(hipEventRecord(b->join[i], s));
    }
    for (int i = 0; i < n + (ib_branch ? 1 : 0); ++i) FBF_HIP_TRY(hipStreamWaitEvent(stream, b->join[i], 0));
    return FBF_OK;
}

extern "C" int fbf_batch_run(fbf_batch* b, const void* taftp, size_t taftp_bytes, const void* const* weights,
                             void* const* outputs, int use_graph, fbf_stream_t stream_)
{
    FBF_REQUIRE(b && taftp && weights && outputs, FBF_ERR_ARG, "fbf_batch_run: null argument");
    FBF_REQUIRE(taftp_bytes == b->ftpa_bytes, FBF_ERR_SIZE,
                "fbf_batch_run: block size " + std::to_string(taftp_bytes) + " != expected " +
                    std::to_string(b->ftpa_bytes));
    const int n = (int)b->subs.size();
    for (int i = 0; i < n; ++i)
        FBF_REQUIRE(weights[i] && outputs[i], FBF_ERR_ARG, "fbf_batch_run: null weights or output");
    FBF_REQUIRE(!b->ib || outputs[n], FBF_ERR_ARG, "fbf_batch_run: null incoherent-beam output");
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    if (!use_graph) return batch_enqueue(b, taftp, taftp_bytes, weights, outputs, stream);

    std::vector<const void*> ptrs{taftp};
    for (int i = 0; i < n; ++i) {
        ptrs.push_back(weights[i]);
        ptrs.push_back(outputs[i]);
    }
    if (b->ib) ptrs.push_back(outputs[n]);
    if (!b->exec || ptrs != b->graph_ptrs || stream != b->graph_stream) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        b->graph_ptrs.clear();
        FBF_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int st = batch_enqueue(b, taftp, taftp_bytes, weights, outputs, stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return hip_status(e, "hipStreamEndCapture");
        b->graph = g;
        FBF_HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_ptrs = ptrs;
        b->graph_stream = stream;
    }
    FBF_HIP_TRY(hipGraphLaunch(b->exec, stream));
    return FBF_OK;
}
