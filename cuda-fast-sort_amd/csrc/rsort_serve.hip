// Serving loop: continuous batching of sort requests into segmented sorts.
//
// No counterpart in the reference, which sorts one caller array per call
// (its per-call cost: graph capture + instantiate + launch, radix_kernel.cuh:
// 583-600).  Here many small requests share one launch sequence:
//
//   submit()      a request = device keys (+ values) sorted in place, like
//                 rsort_sort.  Admission: it must fit one batch slot (keys and
//                 request count); the pending queue is bounded (BUSY when full).
//   scheduler     a host thread closes a batch when the slot is full, when its
//                 oldest request has waited `window_us`, or on flush, and
//                 launches it on a free slot (one stream per slot, `depth`
//                 slots in flight).
//   batch         direct form (default): the host knows every request's size,
//                 so it bins them itself; each non-empty LDS bin is one launch
//                 whose workgroups sort one request each, in place, reading
//                 the request list from host-mapped memory; a request longer
//                 than the LDS capacity gets its own rsort_sort on the slot's
//                 stream.  Batched form (RSORT_SERVE_BATCHED): k_serve_gather
//                 copies every request into the slot's device buffer and
//                 writes the segment offsets, the segmented sort sorts them
//                 all, k_serve_scatter copies the results back; the three
//                 steps are captured once per slot into a hipGraph (fixed
//                 capacity => fixed launch parameters) and replayed.
//   completion    a second host thread waits for each slot's event in launch
//                 order and stamps its requests' completion time.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/rsort_c.h"
#include "rsort_launch.hpp"
#include "rsort_tuning.hpp"

namespace rsa {

// One contiguous piece of one request (<= kServeChunk keys).
struct ServeChunk {
    uint64_t keys;    // device address of the piece's first key
    uint64_t vals;    // device address of its first value (0 without values)
    uint64_t len;     // keys in the piece
    uint64_t offset;  // position of the piece in the batch buffer
};

struct ServeHeader {
    uint64_t chunks;    // pieces in this batch
    uint64_t requests;  // requests in this batch
    uint64_t total;     // keys in this batch
    uint64_t pad;
};

// Pieces of 4096 keys: a batch of a few long requests still spreads over
// many workgroups (one piece per workgroup at a time).
constexpr uint64_t kServeChunk = 4096;

// Copy `count` elements of type E (a key or value word, or a slice of one):
// four loads in flight per thread before their stores.
template <typename E>
__device__ __forceinline__ void serve_copy(E* __restrict__ dst, const E* __restrict__ src, uint64_t count) {
    const uint64_t step = blockDim.x;
    uint64_t i = threadIdx.x;
    for (; i + 3 * step < count; i += 4 * step) {
        const E a = src[i], b = src[i + step], c = src[i + 2 * step], d = src[i + 3 * step];
        dst[i] = a;
        dst[i + step] = b;
        dst[i + 2 * step] = c;
        dst[i + 3 * step] = d;
    }
    for (; i < count; i += step) dst[i] = src[i];
}

// Batch copy in: pieces -> batch buffer, and the segment offsets
// offsets[r] = first key of request r, offsets[r >= requests] = total.
// KE: key word type; VE: value element type, vw elements per value (0: none).
template <typename KE, typename VE>
__global__ __launch_bounds__(256) void k_serve_gather(const ServeHeader* __restrict__ hdr,
                                                      const ServeChunk* __restrict__ ch,
                                                      const uint64_t* __restrict__ host_off, KE* bkeys, VE* bvals,
                                                      uint32_t vw, uint64_t* offsets, uint32_t cap_requests) {
    const uint64_t nch = hdr->chunks;
    const uint64_t nreq = hdr->requests;
    const uint64_t total = hdr->total;
    for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const ServeChunk p = ch[c];
        serve_copy<KE>(bkeys + p.offset, (const KE*)p.keys, p.len);
        if (vw) serve_copy<VE>(bvals + p.offset * vw, (const VE*)p.vals, p.len * vw);
    }
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= cap_requests;
         r += (uint64_t)gridDim.x * blockDim.x)
        offsets[r] = r < nreq ? host_off[r] : total;
}

// Batch copy out: batch buffer -> each request's own buffers.
template <typename KE, typename VE>
__global__ __launch_bounds__(256) void k_serve_scatter(const ServeHeader* __restrict__ hdr,
                                                       const ServeChunk* __restrict__ ch, const KE* bkeys,
                                                       const VE* bvals, uint32_t vw) {
    const uint64_t nch = hdr->chunks;
    for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const ServeChunk p = ch[c];
        serve_copy<KE>((KE*)p.keys, bkeys + p.offset, p.len);
        if (vw) serve_copy<VE>((VE*)p.vals, bvals + p.offset * vw, p.len * vw);
    }
}

template <int B>
struct Word;
template <> struct Word<1> { using T = uint8_t; };
template <> struct Word<2> { using T = uint16_t; };
template <> struct Word<4> { using T = uint32_t; };
template <> struct Word<8> { using T = uint64_t; };
template <> struct Word<16> { using T = uint4; };

struct CopyArgs {
    const ServeHeader* hdr;
    const ServeChunk* chunks;
    const uint64_t* host_off;
    void* bkeys;
    void* bvals;
    uint32_t vw;           // value elements per value
    uint64_t* offsets;
    uint32_t cap_requests;
    uint32_t grid;
};

template <int KB, int VE>
void launch_copy_t(const CopyArgs& c, bool in, hipStream_t s) {
    using K = typename Word<KB>::T;
    using V = typename Word<VE>::T;
    if (in)
        hipLaunchKernelGGL((k_serve_gather<K, V>), dim3(c.grid), dim3(256), 0, s, c.hdr, c.chunks, c.host_off,
                           (K*)c.bkeys, (V*)c.bvals, c.vw, c.offsets, c.cap_requests);
    else
        hipLaunchKernelGGL((k_serve_scatter<K, V>), dim3(c.grid), dim3(256), 0, s, c.hdr, c.chunks, (const K*)c.bkeys,
                           (const V*)c.bvals, c.vw);
}

template <int KB>
void launch_copy_k(const CopyArgs& c, uint32_t ve, bool in, hipStream_t s) {
    switch (ve) {
        case 1: launch_copy_t<KB, 1>(c, in, s); break;
        case 2: launch_copy_t<KB, 2>(c, in, s); break;
        case 4: launch_copy_t<KB, 4>(c, in, s); break;
        case 8: launch_copy_t<KB, 8>(c, in, s); break;
        default: launch_copy_t<KB, 16>(c, in, s); break;
    }
}

hipError_t launch_serve_copy(const CopyArgs& c, uint32_t kb, uint32_t ve, bool in, hipStream_t s) {
    switch (kb) {
        case 1: launch_copy_k<1>(c, ve, in, s); break;
        case 2: launch_copy_k<2>(c, ve, in, s); break;
        case 4: launch_copy_k<4>(c, ve, in, s); break;
        case 8: launch_copy_k<8>(c, ve, in, s); break;
        case 16: launch_copy_k<16>(c, ve, in, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

namespace {

using Clock = std::chrono::steady_clock;

uint64_t us_since(Clock::time_point a, Clock::time_point b) {
    return (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(b - a).count();
}

struct Request {
    void* keys;
    void* vals;
    uint64_t n;
    Clock::time_point t_submit, t_launch, t_done;
    uint64_t batch = 0;
    bool done = false;
};

// One batch slot: device buffers, host-mapped copy list, stream, graph.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    void* bkeys = nullptr;
    void* bvals = nullptr;
    uint64_t* offsets = nullptr;  // device, cap_requests + 1
    void* ws = nullptr;
    size_t ws_bytes = 0;
    uint8_t* host = nullptr;      // host-mapped: header | chunks | offsets
    uint8_t* host_dev = nullptr;  // its device address
    hipGraphExec_t exec = nullptr;
    uint8_t* reqs = nullptr;      // host-mapped request list (direct form)
    uint8_t* reqs_dev = nullptr;
    void* flat_ws = nullptr;      // workspace of rsort_sort for long requests (direct form)
    size_t flat_ws_bytes = 0;
    hipStream_t side[3] = {};     // direct form: the batch forks onto these and joins back
    hipEvent_t fork = nullptr, join[3] = {};
    std::vector<uint64_t> ids;    // requests of the batch in flight
    std::vector<SegReq> items;    // their buffers and sizes
    uint64_t batch = 0;
};

}  // namespace

struct Server {
    rsort_serve_config cfg{};
    int device = 0;
    uint32_t kb = 0, vb = 0, ve = 0, vw = 0;
    bool direct = true;            // direct form (else batched)
    uint32_t cap[kSegBins] = {};   // LDS bin capacities of the element type
    uint64_t max_chunks = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_slot, cv_done;
    std::deque<uint64_t> pending;           // request ids, FIFO
    uint64_t pending_keys = 0;
    std::deque<int> inflight;               // slot indices in launch order
    std::vector<int> free_slots;
    std::unordered_map<uint64_t, Request> req;
    uint64_t next_id = 1, batches = 0, completed = 0, completed_keys = 0;
    uint32_t launching = 0;                 // batches taken from `pending`, not yet in `inflight`
    bool stop = false, flush = false, sched_done = false;
    rsort_status error = RSORT_SUCCESS;
    char error_text[256] = "";
    std::thread scheduler, completer;

    size_t header_bytes() const { return sizeof(ServeHeader); }
    size_t chunks_bytes() const { return (size_t)max_chunks * sizeof(ServeChunk); }
    size_t host_bytes() const { return header_bytes() + chunks_bytes() + ((size_t)cfg.batch_requests + 1) * 8u; }

    // Copy list of requests sl.items[pick[i]] (sl.items[i] without `pick`),
    // i < count, into the slot's host-mapped header / pieces / offsets; the
    // slot is idle, so nothing reads them meanwhile.  Returns the keys listed.
    uint64_t fill_copy_list(Slot& sl, const size_t* pick, size_t count) {
        ServeHeader* h = (ServeHeader*)sl.host;
        ServeChunk* ch = (ServeChunk*)(sl.host + header_bytes());
        uint64_t* off = (uint64_t*)(sl.host + header_bytes() + chunks_bytes());
        uint64_t nch = 0, pos = 0;
        for (size_t i = 0; i < count; ++i) {
            const SegReq& r = sl.items[pick ? pick[i] : i];
            off[i] = pos;
            for (uint64_t b0 = 0; b0 < r.n; b0 += kServeChunk) {
                ServeChunk& c = ch[nch++];
                c.keys = r.keys + b0 * kb;
                c.vals = vb ? r.vals + b0 * vb : 0;
                c.len = r.n - b0 < kServeChunk ? r.n - b0 : kServeChunk;
                c.offset = pos + b0;
            }
            pos += r.n;
        }
        h->chunks = nch;
        h->requests = count;
        h->total = pos;
        return pos;
    }

    CopyArgs copy_args(Slot& sl) const {
        CopyArgs c{};
        c.hdr = (const ServeHeader*)sl.host_dev;
        c.chunks = (const ServeChunk*)(sl.host_dev + header_bytes());
        c.host_off = (const uint64_t*)(sl.host_dev + header_bytes() + chunks_bytes());
        c.bkeys = sl.bkeys;
        c.bvals = sl.bvals;
        c.vw = vw;
        c.offsets = sl.offsets;
        c.cap_requests = cfg.batch_requests;
        c.grid = 1024;
        return c;
    }

    hipError_t enqueue_batch(Slot& sl) {
        const CopyArgs c = copy_args(sl);
        hipError_t e = launch_serve_copy(c, kb, ve, true, sl.stream);
        if (e != hipSuccess) return e;
        size_t tb = sl.ws_bytes;
        const rsort_status st = rsort_segmented_sort(&cfg.desc, sl.bkeys, sl.bvals, cfg.batch_keys, sl.offsets, 8,
                                                     cfg.batch_requests, sl.ws, &tb, sl.stream);
        if (st != RSORT_SUCCESS) return hipErrorLaunchFailure;
        return launch_serve_copy(c, kb, ve, false, sl.stream);
    }

    static constexpr size_t kForkRequests = 32;
    static constexpr size_t kMergeRequests = 4;  // batches this small run their bins as one launch
#ifndef RSA_SERVE_LONG_EACH
#define RSA_SERVE_LONG_EACH 4
#endif
    // up to this many long requests in a batch are sorted one by one
    static constexpr size_t kLongEach = RSA_SERVE_LONG_EACH;

    hipError_t launch_req_on(const ReqJob& j, hipStream_t st) {
        switch (kb) {
            case 1: return launch_req_kb1(j, st);
            case 2: return launch_req_kb2(j, st);
            case 4: return launch_req_kb4(j, st);
            case 8: return launch_req_kb8(j, st);
            default: return launch_req_kb16(j, st);
        }
    }

    // Direct form: bin the slot's requests on the host, one launch per
    // non-empty bin.  A request longer than the LDS capacity is sorted in
    // place by rsort_sort; two or more are copied into the slot's batch
    // buffer back to back, sorted there by one rsort_segmented_sort (its long
    // segment path: a fixed number of launches for any number of requests)
    // and copied back.  The requests are disjoint, so a large batch forks onto
    // the slot's three side streams: the bins spread over the four streams as
    // the segmented sort spreads them, the long requests go on the one with
    // the least bin work, and the side streams join back before the slot's
    // completion event is recorded.
    hipError_t enqueue_direct(Slot& sl) {
        ReqJob j{};
        j.kb = kb;
        j.vb = vb;
        j.tw = make_tw_masks(cfg.desc.key_kind, kb, cfg.desc.descending != 0);
        j.reqs = sl.reqs_dev;
        const size_t nr = sl.items.size();
        std::vector<uint32_t> bin(nr);
        uint32_t in_bin[kSegBins] = {};
        std::vector<size_t> longs;
        for (size_t i = 0; i < nr; ++i) {
            const uint64_t n = sl.items[i].n;
            uint32_t b = kSegBins;  // longer than every bin (or nothing to do)
            if (n >= 2)
                for (uint32_t c = 0; c < (uint32_t)kSegBins; ++c)
                    if (n <= cap[c]) {
                        b = c;
                        break;
                    }
            bin[i] = b;
            if (b < (uint32_t)kSegBins) ++in_bin[b];
            else if (n >= 2) longs.push_back(i);
        }
        // A few requests in different bins share the largest of those bins:
        // one launch instead of one per bin (a bin's workgroup sorts any
        // request up to its capacity), which shortens a small batch on one
        // stream by a kernel per extra bin.
        if (nr <= kMergeRequests) {
            uint32_t top = 0, used = 0;
            for (int b = 0; b < kSegBins; ++b)
                if (in_bin[b]) top = (uint32_t)b, ++used;
            if (used > 1) {
                for (size_t i = 0; i < nr; ++i)
                    if (bin[i] < (uint32_t)kSegBins) bin[i] = top;
                for (int b = 0; b < kSegBins; ++b) in_bin[b] = 0;
                for (size_t i = 0; i < nr; ++i)
                    if (bin[i] < (uint32_t)kSegBins) ++in_bin[bin[i]];
            }
        }
        uint32_t at = 0;
        for (int b = 0; b < kSegBins; ++b) {
            j.first[b] = at;
            j.count[b] = 0;
            at += in_bin[b];
        }
        SegReq* list = (SegReq*)sl.reqs;  // host-mapped; the slot is idle
        for (size_t i = 0; i < nr; ++i)
            if (bin[i] < (uint32_t)kSegBins) list[j.first[bin[i]] + j.count[bin[i]]++] = sl.items[i];

        // fork only for a batch large enough to pay for the cross-stream
        // events, and only when at least two streams get work: batches of a
        // few requests (light load) stay on one stream.  Stream 1 takes the
        // long requests (its one bin, 5, is short).
        static constexpr int kBinStream[kSegBins] = {3, 2, 3, 2, 2, 1, 0, 0, 3};  // bins 7, 6 | 5 | 4, 3, 1 | 0, 2, 8
        static constexpr int kLongStream = 1;
        const bool long_batch = longs.size() > kLongEach;
        bool busy[4] = {false, false, false, false};
        for (int b = 0; b < kSegBins; ++b) busy[kBinStream[b]] = busy[kBinStream[b]] || j.count[b] != 0;
        if (long_batch) busy[kLongStream] = true;
        for (size_t q = 0; q < longs.size() && !long_batch; ++q) busy[(kLongStream + q) % 4] = true;
        const int nbusy = (int)busy[0] + (int)busy[1] + (int)busy[2] + (int)busy[3];
        const bool fork = sl.side[0] != nullptr && nbusy > 1 && (nr >= kForkRequests || longs.size() > 1);
        const hipStream_t st[4] = {sl.stream, fork ? sl.side[0] : sl.stream, fork ? sl.side[1] : sl.stream,
                                   fork ? sl.side[2] : sl.stream};
        hipError_t e = hipSuccess;
        if (fork) {
            e = hipEventRecord(sl.fork, sl.stream);
            for (int k = 0; k < 3 && e == hipSuccess; ++k)
                if (busy[k + 1]) e = hipStreamWaitEvent(sl.side[k], sl.fork, 0);
        }
        // long requests first: theirs is the longest chain of launches.  A
        // few: one rsort_sort each, round-robin over the streams from
        // kLongStream on, each stream's in its own part of the flat workspace
        // (all on one stream when the parts do not fit).  More: one
        // segmented sort in the batch buffer.
        if (e == hipSuccess && !long_batch && !longs.empty()) {
            int nst = fork ? 4 : 1;
            uint64_t maxn[4] = {0, 0, 0, 0};
            size_t off[4] = {0, 0, 0, 0}, bytes[4] = {0, 0, 0, 0}, total = 0;
            for (size_t q = 0; q < longs.size(); ++q) {
                const uint64_t n = sl.items[longs[q]].n;
                if (n > maxn[q % nst]) maxn[q % nst] = n;
            }
            for (int k = 0; k < nst; ++k) {
                if (!maxn[k]) continue;
                size_t need = 0;
                if (rsort_sort(&cfg.desc, nullptr, nullptr, nullptr, &need, maxn[k], nullptr) != RSORT_SUCCESS)
                    need = sl.flat_ws_bytes + 1u;
                off[k] = total;
                bytes[k] = need;
                total += (need + 255u) & ~(size_t)255u;
            }
            if (total > sl.flat_ws_bytes) {
                nst = 1;
                off[0] = 0;
                bytes[0] = sl.flat_ws_bytes;
            }
            for (size_t q = 0; q < longs.size() && e == hipSuccess; ++q) {
                const int k = (int)(q % nst);
                const SegReq& r = sl.items[longs[q]];
                size_t tb = bytes[k];
                if (rsort_sort(&cfg.desc, (void*)(uintptr_t)r.keys, (void*)(uintptr_t)r.vals,
                               (uint8_t*)sl.flat_ws + off[k], &tb, r.n, st[(kLongStream + k) % 4]) != RSORT_SUCCESS)
                    e = hipErrorLaunchFailure;
            }
        } else if (e == hipSuccess && long_batch) {
            const uint64_t total = fill_copy_list(sl, longs.data(), longs.size());
            const CopyArgs c = copy_args(sl);
            e = launch_serve_copy(c, kb, ve, true, st[kLongStream]);
            rsort_desc d = cfg.desc;
            d.flags &= ~RSORT_FLAG_GRAPH;  // sizes differ from batch to batch
            size_t tb = sl.ws_bytes;
            if (e == hipSuccess && rsort_segmented_sort(&d, sl.bkeys, sl.bvals, total, sl.offsets, 8, longs.size(),
                                                        sl.ws, &tb, st[kLongStream]) != RSORT_SUCCESS)
                e = hipErrorLaunchFailure;
            if (e == hipSuccess) e = launch_serve_copy(c, kb, ve, false, st[kLongStream]);
        }
        for (int k = 0; k < 4 && e == hipSuccess; ++k) {
            ReqJob jk = j;
            bool any = false;
            for (int b = 0; b < kSegBins; ++b) {
                if (kBinStream[b] != k) jk.count[b] = 0;
                any = any || jk.count[b] != 0;
            }
            if (any) e = launch_req_on(jk, st[k]);
        }
        if (fork) {
            for (int k = 0; k < 3 && e == hipSuccess; ++k) {
                if (!busy[k + 1]) continue;
                e = hipEventRecord(sl.join[k], sl.side[k]);
                if (e == hipSuccess) e = hipStreamWaitEvent(sl.stream, sl.join[k], 0);
            }
        }
        return e;
    }

    // Runs every kernel the direct form can use once, on the slot's own
    // buffers (their contents are not data yet): one request per LDS bin, and
    // rsort_sort at sizes from just above the LDS threshold up to batch_keys
    // by factors of 4, which covers the tuned tile classes.  A kernel's
    // first launch in a process loads its code object, which stalled the
    // first long request of a served stream by ~16 ms (p99 8 ms over 5000
    // Poisson requests) before this warm-up.
    hipError_t warm_direct(Slot& sl) {
        sl.items.clear();
        uint64_t at = 0, longest = 0;
        for (int b = 0; b < kSegBins; ++b) {
            if (!cap[b]) continue;
            longest = cap[b];
            const uint64_t n = cap[b] < cfg.batch_keys - at ? cap[b] : cfg.batch_keys - at;
            if (n < 2) break;
            SegReq r{};
            r.keys = (uint64_t)(uintptr_t)((uint8_t*)sl.bkeys + at * kb);
            r.vals = vb ? (uint64_t)(uintptr_t)((uint8_t*)sl.bvals + at * vb) : 0u;
            r.n = n;
            sl.items.push_back(r);
            at += n;
        }
        hipError_t e = enqueue_direct(sl);
        uint64_t n = longest + 1u > 2u ? longest + 1u : 2u;
        // enough long requests for one segmented sort: the copy and
        // segmented-sort kernels (each copy lands where it came from, the
        // requests lying back to back in the batch buffer)
        const uint64_t nlong = kLongEach + 1u;  // a batch can hold that many only if batch_requests allows
        if (e == hipSuccess && nlong <= cfg.batch_requests && nlong * n <= cfg.batch_keys) {
            sl.items.clear();
            for (uint64_t i = 0; i < nlong; ++i) {
                SegReq r{};
                r.keys = (uint64_t)(uintptr_t)((uint8_t*)sl.bkeys + i * n * kb);
                r.vals = vb ? (uint64_t)(uintptr_t)((uint8_t*)sl.bvals + i * n * vb) : 0u;
                r.n = n;
                sl.items.push_back(r);
            }
            e = enqueue_direct(sl);
        }
        while (n <= cfg.batch_keys && e == hipSuccess) {
            size_t tb = sl.flat_ws_bytes;
            if (rsort_sort(&cfg.desc, sl.bkeys, sl.bvals, sl.flat_ws, &tb, n, sl.stream) != RSORT_SUCCESS)
                e = hipErrorLaunchFailure;
            if (n == cfg.batch_keys) break;
            n = n > cfg.batch_keys / 4u ? cfg.batch_keys : n * 4u;
        }
        if (e == hipSuccess) e = hipStreamSynchronize(sl.stream);
        sl.items.clear();
        return e;
    }

    void fail_with(rsort_status st, const char* what, hipError_t e) {
        std::lock_guard<std::mutex> lk(mu);
        if (error == RSORT_SUCCESS) {
            error = st;
            snprintf(error_text, sizeof(error_text), "%s: %s", what, hipGetErrorString(e));
        }
        stop = true;
        cv_work.notify_all();
        cv_slot.notify_all();
        cv_done.notify_all();
    }

    void run_scheduler() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            if (stop && (pending.empty() || error != RSORT_SUCCESS)) break;
            if (pending.empty()) {
                cv_work.wait(lk);
                continue;
            }
            const auto oldest = req[pending.front()].t_submit;
            const auto deadline = oldest + std::chrono::microseconds(cfg.window_us);
            const bool full = pending_keys >= cfg.batch_keys || pending.size() >= cfg.batch_requests;
            if (!full && !flush && !stop && Clock::now() < deadline) {
                cv_work.wait_until(lk, deadline);
                continue;
            }
            if (free_slots.empty()) {
                cv_slot.wait(lk);
                continue;
            }
            const int si = free_slots.back();
            free_slots.pop_back();
            Slot& sl = slots[si];
            // admission into this batch: FIFO while the slot has room
            // (with a snapshot of each record: they are only read under the lock)
            sl.ids.clear();
            sl.items.clear();
            sl.batch = ++batches;
            const auto now = Clock::now();
            uint64_t keys = 0;
            while (!pending.empty() && sl.ids.size() < cfg.batch_requests) {
                Request& r = req.find(pending.front())->second;
                if (keys + r.n > cfg.batch_keys) break;
                keys += r.n;
                pending_keys -= r.n;
                r.t_launch = now;
                r.batch = sl.batch;
                sl.ids.push_back(pending.front());
                sl.items.push_back(SegReq{(uint64_t)(uintptr_t)r.keys, (uint64_t)(uintptr_t)r.vals, r.n, 0});
                pending.pop_front();
            }
            if (pending.empty()) flush = false;
            ++launching;
            lk.unlock();
            hipError_t e = hipSuccess;
            if (direct) {
                e = enqueue_direct(sl);
            } else {
                fill_copy_list(sl, nullptr, sl.items.size());
                e = sl.exec ? hipGraphLaunch(sl.exec, sl.stream) : enqueue_batch(sl);
            }
            if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
            lk.lock();
            --launching;
            if (e != hipSuccess) {
                lk.unlock();
                fail_with(RSORT_ERROR_DEVICE, "batch launch", e);
                lk.lock();
                break;
            }
            inflight.push_back(si);
            cv_done.notify_all();  // wakes the completion thread
        }
        sched_done = true;
        cv_done.notify_all();
    }

    void run_completer() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            if (inflight.empty()) {
                if (sched_done) break;
                cv_done.wait(lk);
                continue;
            }
            const int si = inflight.front();
            Slot& sl = slots[si];
            lk.unlock();
            const hipError_t e = hipEventSynchronize(sl.done);
            const auto now = Clock::now();
            lk.lock();
            inflight.pop_front();
            if (e != hipSuccess) {
                lk.unlock();
                fail_with(RSORT_ERROR_DEVICE, "batch completion", e);
                lk.lock();
            }
            for (uint64_t id : sl.ids) {
                Request& r = req[id];
                r.t_done = now;
                r.done = true;
                ++completed;
                completed_keys += r.n;
            }
            sl.ids.clear();
            free_slots.push_back(si);
            cv_slot.notify_all();
            cv_work.notify_all();
            cv_done.notify_all();
        }
    }
};

}  // namespace rsa

using rsa::Server;

namespace {

void destroy_slots(Server* s) {
    for (auto& sl : s->slots) {
        if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
        if (sl.done) (void)hipEventDestroy(sl.done);
        for (int k = 0; k < 3; ++k) {
            if (sl.side[k]) (void)hipStreamDestroy(sl.side[k]);
            if (sl.join[k]) (void)hipEventDestroy(sl.join[k]);
        }
        if (sl.fork) (void)hipEventDestroy(sl.fork);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.bkeys) (void)hipFree(sl.bkeys);
        if (sl.bvals) (void)hipFree(sl.bvals);
        if (sl.offsets) (void)hipFree(sl.offsets);
        if (sl.ws) (void)hipFree(sl.ws);
        if (sl.host) (void)hipHostFree(sl.host);
        if (sl.reqs) (void)hipHostFree(sl.reqs);
        if (sl.flat_ws) (void)hipFree(sl.flat_ws);
    }
    s->slots.clear();
}

}  // namespace

extern "C" {

rsort_status rsort_serve_create(const rsort_serve_config* cfg, rsort_server** out) {
    if (!cfg || !out) return RSORT_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (cfg->batch_keys < 2 || cfg->batch_requests < 1 || cfg->depth < 1 || cfg->depth > 8 || cfg->reserved)
        return RSORT_ERROR_INVALID_ARGUMENT;
    if (cfg->batch_requests > (1u << 24)) return RSORT_ERROR_UNSUPPORTED;
    size_t need = 0;
    rsort_status st = rsort_segmented_sort(&cfg->desc, nullptr, nullptr, cfg->batch_keys, nullptr, 8,
                                           cfg->batch_requests, nullptr, &need, nullptr);
    if (st != RSORT_SUCCESS) return st;

    Server* s = new Server();
    s->cfg = *cfg;
    s->kb = cfg->desc.key_bytes;
    s->vb = cfg->desc.value_bytes;
    s->ve = s->vb == 0 ? 1 : (s->vb % 16 == 0 ? 16 : s->vb % 8 == 0 ? 8 : s->vb % 4 == 0 ? 4 : s->vb % 2 == 0 ? 2 : 1);
    s->vw = s->vb / s->ve;
    s->max_chunks = cfg->batch_requests + cfg->batch_keys / rsa::kServeChunk + 1;
    const bool pow2v = s->vb == 0 || s->vb == 1 || s->vb == 2 || s->vb == 4 || s->vb == 8 || s->vb == 16;
    s->direct = !(cfg->flags & RSORT_SERVE_BATCHED) && pow2v;
    rsort_tuning tune{};
    const uint32_t thr = rsort_tuning_get(s->kb, s->vb, &tune) == RSORT_SUCCESS ? tune.lds_threshold : 0u;
    for (int b = 0; b < rsa::kSegBins; ++b) {
        const uint32_t c = rsa::seg_cap((int)s->kb, (int)s->vb, b);
        s->cap[b] = c <= thr ? c : 0u;  // bins above the tuned LDS threshold are not used
    }
    size_t flat_need = 0;
    if (s->direct) {
        // room for the long requests of a batch split over four streams: their
        // keys add up to at most a batch, each part adds its fixed overhead
        size_t fixed = 0;
        st = rsort_sort(&cfg->desc, nullptr, nullptr, nullptr, &flat_need, cfg->batch_keys, nullptr);
        if (st == RSORT_SUCCESS) st = rsort_sort(&cfg->desc, nullptr, nullptr, nullptr, &fixed, 1, nullptr);
        flat_need += 3 * (fixed + 256u);
        if (st != RSORT_SUCCESS) {
            delete s;
            return st;
        }
    }
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return RSORT_ERROR_DEVICE;
    }
    s->slots.resize(cfg->depth);
    s->req.reserve((size_t)cfg->batch_requests * (cfg->depth + 1u));  // no rehash while batches fill
    hipError_t e = hipSuccess;
    for (uint32_t i = 0; i < cfg->depth && e == hipSuccess; ++i) {
        auto& sl = s->slots[i];
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(&sl.bkeys, cfg->batch_keys * s->kb);
        if (e == hipSuccess && s->vb) e = hipMalloc(&sl.bvals, cfg->batch_keys * s->vb);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.offsets, ((size_t)cfg->batch_requests + 1) * 8u);
        sl.ws_bytes = need;
        if (e == hipSuccess) e = hipMalloc(&sl.ws, need);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&sl.host, s->host_bytes(), hipHostMallocMapped | hipHostMallocCoherent);
        if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&sl.host_dev, sl.host, 0);
        if (e == hipSuccess) memset(sl.host, 0, s->host_bytes());
        if (e == hipSuccess && s->direct) {
            const size_t rb = (size_t)cfg->batch_requests * sizeof(rsa::SegReq);
            e = hipHostMalloc((void**)&sl.reqs, rb, hipHostMallocMapped | hipHostMallocCoherent);
            if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&sl.reqs_dev, sl.reqs, 0);
            sl.flat_ws_bytes = flat_need;
            if (e == hipSuccess) e = hipMalloc(&sl.flat_ws, flat_need);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.fork, hipEventDisableTiming);
            for (int k = 0; k < 3 && e == hipSuccess; ++k) {
                e = hipStreamCreateWithFlags(&sl.side[k], hipStreamNonBlocking);
                if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.join[k], hipEventDisableTiming);
            }
        }
    }
    if (e == hipSuccess && s->direct) e = s->warm_direct(s->slots[0]);
    // one direct run per slot (empty batch: also initialises the launchers'
    // one-time queries), then capture the batch into a graph
    for (uint32_t i = 0; i < cfg->depth && e == hipSuccess && !s->direct; ++i) {
        auto& sl = s->slots[i];
        e = s->enqueue_batch(sl);
        if (e == hipSuccess) e = hipStreamSynchronize(sl.stream);
        if (e == hipSuccess && !(cfg->flags & RSORT_SERVE_NO_GRAPH)) {
            hipGraph_t g = nullptr;
            e = hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                const hipError_t e1 = s->enqueue_batch(sl);
                e = hipStreamEndCapture(sl.stream, &g);
                if (e == hipSuccess) e = e1;
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&sl.exec, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
        }
    }
    if (e != hipSuccess) {
        destroy_slots(s);
        delete s;
        return RSORT_ERROR_DEVICE;
    }
    for (uint32_t i = 0; i < cfg->depth; ++i) s->free_slots.push_back((int)(cfg->depth - 1 - i));
    s->scheduler = std::thread([s] { s->run_scheduler(); });
    s->completer = std::thread([s] { s->run_completer(); });
    *out = (rsort_server*)s;
    return RSORT_SUCCESS;
}

rsort_status rsort_serve_submit(rsort_server* srv, void* keys, void* values, uint64_t n, uint64_t* request_id) {
    Server* s = (Server*)srv;
    if (!s || !request_id) return RSORT_ERROR_INVALID_ARGUMENT;
    if (n > 0 && !keys) return RSORT_ERROR_INVALID_ARGUMENT;
    if (s->vb && n > 0 && !values) return RSORT_ERROR_INVALID_ARGUMENT;
    if ((uintptr_t)keys % s->kb || (s->vb && (uintptr_t)values % s->ve)) return RSORT_ERROR_INVALID_ARGUMENT;
    if (n > s->cfg.batch_keys) return RSORT_ERROR_UNSUPPORTED;  // never fits a batch
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->error != RSORT_SUCCESS) return s->error;
    if (s->stop) return RSORT_ERROR_BUSY;
    if (s->cfg.max_pending && s->pending.size() >= s->cfg.max_pending) return RSORT_ERROR_BUSY;
    const uint64_t id = s->next_id++;
    rsa::Request r{};
    r.keys = keys;
    r.vals = values;
    r.n = n;
    r.t_submit = rsa::Clock::now();
    s->req.emplace(id, r);
    s->pending.push_back(id);
    s->pending_keys += n;
    *request_id = id;
    s->cv_work.notify_one();
    return RSORT_SUCCESS;
}

rsort_status rsort_serve_submit_many(rsort_server* srv, void* const* keys, void* const* values, const uint64_t* n,
                                     uint64_t count, uint64_t* request_ids) {
    Server* s = (Server*)srv;
    if (!s || (count && (!keys || !n || !request_ids))) return RSORT_ERROR_INVALID_ARGUMENT;
    if (s->vb && count && !values) return RSORT_ERROR_INVALID_ARGUMENT;
    for (uint64_t i = 0; i < count; ++i) {
        if (n[i] > 0 && !keys[i]) return RSORT_ERROR_INVALID_ARGUMENT;
        if (s->vb && n[i] > 0 && !values[i]) return RSORT_ERROR_INVALID_ARGUMENT;
        if ((uintptr_t)keys[i] % s->kb || (s->vb && (uintptr_t)values[i] % s->ve)) return RSORT_ERROR_INVALID_ARGUMENT;
        if (n[i] > s->cfg.batch_keys) return RSORT_ERROR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->error != RSORT_SUCCESS) return s->error;
    if (s->stop) return RSORT_ERROR_BUSY;
    if (s->cfg.max_pending && s->pending.size() + count > s->cfg.max_pending) return RSORT_ERROR_BUSY;
    const auto now = rsa::Clock::now();
    for (uint64_t i = 0; i < count; ++i) {
        const uint64_t id = s->next_id++;
        rsa::Request r{};
        r.keys = keys[i];
        r.vals = s->vb ? values[i] : nullptr;
        r.n = n[i];
        r.t_submit = now;
        s->req.emplace(id, r);
        s->pending.push_back(id);
        s->pending_keys += n[i];
        request_ids[i] = id;
    }
    s->cv_work.notify_one();
    return RSORT_SUCCESS;
}

namespace {

// Waits, under the server's lock, until request `id` is done, the server has
// failed, or the deadline passes (timeout_us: 0 = poll, RSORT_SERVE_WAIT_FOREVER
// or more = no limit).  SUCCESS when done; BUSY on time-out; INVALID_ARGUMENT
// when the record is unknown (never submitted, or released by another wait).
rsort_status wait_done(Server* s, std::unique_lock<std::mutex>& lk, uint64_t id, uint64_t timeout_us,
                       rsa::Clock::time_point deadline) {
    while (true) {
        auto it = s->req.find(id);
        if (it == s->req.end()) return RSORT_ERROR_INVALID_ARGUMENT;
        if (it->second.done) return RSORT_SUCCESS;
        if (s->error != RSORT_SUCCESS) return s->error;
        if (timeout_us == 0) return RSORT_ERROR_BUSY;
        if (timeout_us >= RSORT_SERVE_WAIT_FOREVER) {
            s->cv_done.wait(lk);
        } else if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
            it = s->req.find(id);
            if (it == s->req.end()) return RSORT_ERROR_INVALID_ARGUMENT;
            return it->second.done ? RSORT_SUCCESS : RSORT_ERROR_BUSY;
        }
    }
}

rsa::Clock::time_point deadline_after(uint64_t timeout_us) {
    const uint64_t us = timeout_us < RSORT_SERVE_WAIT_FOREVER ? timeout_us : 0u;
    return rsa::Clock::now() + std::chrono::microseconds((int64_t)us);
}

// Fills `info` (if any) from a done record and releases it.
void release(Server* s, uint64_t id, rsort_request_info* info) {
    auto it = s->req.find(id);
    if (it == s->req.end()) return;
    if (info) {
        info->latency_us = rsa::us_since(it->second.t_submit, it->second.t_done);
        info->queue_us = rsa::us_since(it->second.t_submit, it->second.t_launch);
        info->batch = it->second.batch;
        info->n = it->second.n;
    }
    s->req.erase(it);
}

}  // namespace

rsort_status rsort_serve_wait(rsort_server* srv, uint64_t request_id, uint64_t timeout_us, rsort_request_info* info) {
    Server* s = (Server*)srv;
    if (!s) return RSORT_ERROR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    const rsort_status st = wait_done(s, lk, request_id, timeout_us, deadline_after(timeout_us));
    if (st == RSORT_SUCCESS) release(s, request_id, info);
    return st;
}

rsort_status rsort_serve_wait_many(rsort_server* srv, const uint64_t* request_ids, uint64_t count,
                                   uint64_t timeout_us, rsort_request_info* infos) {
    Server* s = (Server*)srv;
    if (!s || (count && !request_ids)) return RSORT_ERROR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    for (uint64_t i = 0; i < count; ++i)
        if (s->req.find(request_ids[i]) == s->req.end()) return RSORT_ERROR_INVALID_ARGUMENT;
    const auto deadline = deadline_after(timeout_us);
    for (uint64_t i = 0; i < count; ++i) {
        const rsort_status st = wait_done(s, lk, request_ids[i], timeout_us, deadline);
        if (st != RSORT_SUCCESS) return st;  // nothing released
    }
    for (uint64_t i = 0; i < count; ++i) release(s, request_ids[i], infos ? infos + i : nullptr);
    return RSORT_SUCCESS;
}

rsort_status rsort_serve_flush(rsort_server* srv) {
    Server* s = (Server*)srv;
    if (!s) return RSORT_ERROR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    s->flush = true;
    s->cv_work.notify_all();
    while ((!s->pending.empty() || s->launching || !s->inflight.empty()) && s->error == RSORT_SUCCESS)
        s->cv_done.wait(lk);
    return s->error;
}

rsort_status rsort_serve_stats(rsort_server* srv, rsort_serve_stats_info* out) {
    Server* s = (Server*)srv;
    if (!s || !out) return RSORT_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    out->batches = s->batches;
    out->completed = s->completed;
    out->completed_keys = s->completed_keys;
    out->pending = s->pending.size();
    out->in_flight = s->inflight.size();
    out->graphs = s->slots.empty() ? 0u : (s->slots[0].exec ? 1u : 0u);
    out->direct = s->direct ? 1u : 0u;
    out->workspace_bytes = 0;
    for (auto& sl : s->slots)
        out->workspace_bytes += sl.ws_bytes + s->cfg.batch_keys * (s->kb + s->vb) + (s->cfg.batch_requests + 1) * 8u;
    return s->error;
}

rsort_status rsort_serve_destroy(rsort_server* srv) {
    Server* s = (Server*)srv;
    if (!s) return RSORT_ERROR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->flush = true;
        s->cv_work.notify_all();
        s->cv_slot.notify_all();
        s->cv_done.notify_all();
    }
    if (s->scheduler.joinable()) s->scheduler.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->cv_done.notify_all();
    }
    if (s->completer.joinable()) s->completer.join();
    const rsort_status st = s->error;
    destroy_slots(s);
    delete s;
    return st;
}

}  // extern "C"
