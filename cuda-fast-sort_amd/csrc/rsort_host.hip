// C ABI of the sort: descriptor validation, plan and workspace layout,
// dispatch to the per-key-width launchers.
//
// Reference counterparts:
//   Sort_Workspace::build/bind        include/rsort/detail/radix_kernel.cuh:35-109
//   onesweep_byte_sort_impl (query)   include/rsort/detail/radix_kernel.cuh:486-515
//   lookback_policy_enforcer          include/rsort/detail/radix_kernel.cuh:525-559
//   long double -> 128-bit word       include/rsort/detail/radix_kernel.cuh:351-356
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <dlfcn.h>

#include <unistd.h>

#include <atomic>
#include <chrono>
#include <mutex>
#include <random>
#include <vector>

#include "../../include/rsort_c.h"
#include "rsort_kernels.hpp"
#include "rsort_launch.hpp"

namespace rsa {

namespace {

thread_local char g_last_error[512] = "";

rsort_status fail(rsort_status s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
rsort_status fail(rsort_status s, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return s;
}

rsort_status hip_fail(hipError_t e, const char* what) {
    return fail(RSORT_ERROR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

size_t reserve(size_t* off, size_t bytes, size_t align = 256) {
    const size_t at = (*off + align - 1) / align * align;
    *off = at + bytes;
    return at;
}

bool is_pow2_width(uint32_t b) { return b == 1 || b == 2 || b == 4 || b == 8 || b == 16; }

// Passes of at most this many tiles walk the lookback in segments, one round
// of loads for up to (threads / 256 - 1) * 4 groups: there every tile starts
// at once and the walks reach back to the first tiles before those have
// published inclusive prefixes.  RSORT_AMD_LB_WIDE_TILES overrides it (0 = never).
uint64_t lb_wide_max_tiles() {
    static const uint64_t v = [] {
        const char* e = getenv("RSORT_AMD_LB_WIDE_TILES");
        return e ? (uint64_t)strtoull(e, nullptr, 10) : (uint64_t)kLbWideMaxTiles;
    }();
    return v;
}

// u32/u32 passes of at least this many tiles publish tile aggregates before
// the rank.  RSORT_AMD_EARLY_AGG_TILES overrides it (0 = never).
uint64_t early_agg_min_tiles() {
    static const uint64_t v = [] {
        const char* e = getenv("RSORT_AMD_EARLY_AGG_TILES");
        return e ? (uint64_t)strtoull(e, nullptr, 10) : (uint64_t)kEarlyAggMinTiles;
    }();
    return v;
}

// Per-XCD tile tickets of the pass kernel (where the device allows them).
// RSORT_AMD_XCD_TICKETS=0 keeps the single global ticket everywhere.
bool xcd_tickets_allowed() {
    static const bool v = [] {
        const char* e = getenv("RSORT_AMD_XCD_TICKETS");
        return !(e && strcmp(e, "0") == 0);
    }();
    return v;
}

// The histogram kernel clears the control region itself (no fill kernel)
// where its grid allows; RSORT_AMD_HIST_CLEAR=0 keeps the fill kernel.
bool hist_clear_allowed() {
    static const bool v = [] {
        const char* e = getenv("RSORT_AMD_HIST_CLEAR");
        return !(e && strcmp(e, "0") == 0);
    }();
    return v;
}

// RSORT_AMD_HIST_CLAIM=late (a test setting): the histogram's workgroup 0
// does not claim the clearing when it starts, so the claim is always made
// where a workgroup first needs the counts.
bool hist_late_claim() {
    static const bool v = [] {
        const char* e = getenv("RSORT_AMD_HIST_CLAIM");
        return e && strcmp(e, "late") == 0;
    }();
    return v;
}

// A fresh nonzero 64-bit id per sort call: odd numbers counted up from a
// random start (random device, clock and process id), so ids never repeat in
// a process and one left in a workspace by another process matches by chance
// only.
uint64_t next_clear_id() {
    static std::atomic<uint64_t> next{[] {
        uint64_t v = (uint64_t)std::chrono::steady_clock::now().time_since_epoch().count() * 0x9E3779B97F4A7C15ull;
        v ^= (uint64_t)getpid() << 32;
        try {
            std::random_device rd;
            v ^= ((uint64_t)rd() << 32) ^ (uint64_t)rd();
        } catch (...) {
        }
        return v | 1u;
    }()};
    return next.fetch_add(2u, std::memory_order_relaxed);
}

// Histogram grids of more than one fold group and at most this many
// workgroups add their counts with device atomics directly instead of the
// two-level fold.  RSORT_AMD_HIST_DIRECT overrides it (0 = never).
uint32_t hist_direct_max() {
    static const uint32_t v = [] {
        const char* e = getenv("RSORT_AMD_HIST_DIRECT");
        return e ? (uint32_t)strtoul(e, nullptr, 10) : kHistDirectMaxGrid;
    }();
    return v;
}

// ---------------------------------------------------------------------------
// Tuning table: per (key width, value width) the pass kernel's tile
// candidate, the histogram workgroups per CU and the segmented sort's LDS
// threshold, plus the tile candidate of small sorts.  Built-in defaults,
// overridden by a text table ("key_bytes value_bytes tile_candidate
// hist_blocks_per_cu lds_threshold [small_tile_candidate small_log2n
// [tiny_tile_candidate tiny_log2n]]" per
// line, '#' comments) read once from $RSORT_AMD_TUNING_FILE or tuning_gfx950.txt next to
// this library, and by rsort_tuning_set.  The table is written by the
// autotuner (autotune.py).
// ---------------------------------------------------------------------------
struct TuneEntry {
    uint32_t tile_cand, hist_bpc, lds_threshold;
    uint32_t small_tile_cand, small_log2n;  // candidate of sorts below 2^small_log2n keys
    uint32_t tiny_tile_cand, tiny_log2n;    // candidate of sorts below 2^tiny_log2n keys (checked first)
};

int kb_index(uint32_t kb) { return kb == 1 ? 0 : kb == 2 ? 1 : kb == 4 ? 2 : kb == 8 ? 3 : kb == 16 ? 4 : -1; }
int vb_index(uint32_t vb) {
    return vb == 0 ? 0 : vb == 1 ? 1 : vb == 2 ? 2 : vb == 4 ? 3 : vb == 8 ? 4 : vb == 16 ? 5 : -1;
}
constexpr uint32_t kVbOf[6] = {0, 1, 2, 4, 8, 16};
constexpr uint32_t kKbOf[5] = {1, 2, 4, 8, 16};

std::mutex g_tune_mu;
TuneEntry g_tune[5][6];
bool g_tune_ready = false;

// Measured on MI355X (whole-call time, uniform keys): u32 keys sort faster
// with 1024 x 8 tiles below 2^22 keys; u32/u32 with 1024 x 16 tiles (two
// workgroups per CU) from 2^22 up and 1024 x 10 tiles below; u64/u64 with
// 1024 x 5 tiles below 2^21.
TuneEntry tune_default(uint32_t kb, uint32_t vb) {
    TuneEntry t{0u, hist_blocks_per_cu((int)kb), seg_cap((int)kb, (int)vb, kSegBins - 1), 0u, 0u, 0u, 0u};
    if (kb == 4 && vb == 0) t.small_tile_cand = 3u, t.small_log2n = 22u;
    if (kb == 4 && vb == 4) t.tile_cand = 1u, t.small_tile_cand = 3u, t.small_log2n = 22u;
    if (kb == 8 && vb == 8) t.small_tile_cand = 3u, t.small_log2n = 21u;
    return t;
}

bool tune_valid(uint32_t kb, uint32_t vb, const TuneEntry& t) {
    return t.tile_cand < (uint32_t)kTileCands && t.hist_bpc >= 1 && t.hist_bpc <= 8 &&
           t.lds_threshold <= seg_cap((int)kb, (int)vb, kSegBins - 1) && t.small_tile_cand < (uint32_t)kTileCands &&
           t.small_log2n <= 40u && t.tiny_tile_cand < (uint32_t)kTileCands && t.tiny_log2n <= 40u;
}

// Reads a table file into g_tune (caller holds g_tune_mu); returns entries read or -1.
int tune_read_file(const char* path) {
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    char line[256];
    int n = 0;
    while (fgets(line, sizeof(line), f)) {
        char* h = strchr(line, '#');
        if (h) *h = 0;
        unsigned kb, vb, tc, bpc, thr, stc = 0, sl = 0, ttc = 0, tl = 0;
        const int got = sscanf(line, "%u %u %u %u %u %u %u %u %u", &kb, &vb, &tc, &bpc, &thr, &stc, &sl, &ttc, &tl);
        if (got != 5 && got != 7 && got != 9) continue;
        const int ki = kb_index(kb), vi = vb_index(vb);
        const TuneEntry t{tc, bpc, thr, stc, sl, got == 9 ? ttc : 0u, got == 9 ? tl : 0u};
        if (ki < 0 || vi < 0 || !tune_valid(kb, vb, t)) continue;
        g_tune[ki][vi] = t;
        ++n;
    }
    fclose(f);
    return n;
}

void tune_reset_locked() {
    for (int ki = 0; ki < 5; ++ki)
        for (int vi = 0; vi < 6; ++vi) g_tune[ki][vi] = tune_default(kKbOf[ki], kVbOf[vi]);
}

// Defaults, then the persisted table ($RSORT_AMD_TUNING_FILE, else
// tuning_gfx950.txt next to this library); returns entries read or -1.
int tune_load_default_locked() {
    tune_reset_locked();
    g_tune_ready = true;
    const char* env = getenv("RSORT_AMD_TUNING_FILE");
    if (env && *env) return tune_read_file(env);
    Dl_info di;
    if (dladdr((const void*)&tune_default, &di) && di.dli_fname) {
        char path[1024];
        snprintf(path, sizeof(path), "%s", di.dli_fname);
        char* slash = strrchr(path, '/');
        if (slash) {
            snprintf(slash + 1, sizeof(path) - (size_t)(slash + 1 - path), "tuning_gfx950.txt");
            return tune_read_file(path);
        }
    }
    return -1;
}

void tune_ensure_locked() {
    if (!g_tune_ready) (void)tune_load_default_locked();
}

// Tuning of (kb, vb_core); value widths outside {0,1,2,4,8,16} use index mode (4-byte values).
TuneEntry tune_get(uint32_t kb, uint32_t vb_core) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_ensure_locked();
    const int ki = kb_index(kb), vi = vb_index(vb_core);
    return (ki < 0 || vi < 0) ? tune_default(kb, vb_core) : g_tune[ki][vi];
}

struct Plan {
    uint32_t tile_cand, hist_bpc;
    uint32_t kind, kb, vb_user, vb_core;
    bool desc, allow_skip, index_mode, wide, lb_wide;
    uint32_t passes;
    TileCfg cfg;
    uint64_t tile, num_tiles;
    size_t off_tmp_keys, off_tmp_vals, off_idx, off_vals_copy, off_ctrl, off_counts, off_gp, off_lb;
    size_t off_hist_partials;
    size_t lb_bytes, lb_group_off, total;
};

rsort_status check_desc(const rsort_desc* d) {
    if (!d) return fail(RSORT_ERROR_INVALID_ARGUMENT, "descriptor is NULL");
    if (d->reserved != 0) return fail(RSORT_ERROR_INVALID_ARGUMENT, "descriptor reserved field must be 0");
    if (d->key_kind > RSORT_KEY_FLOAT) return fail(RSORT_ERROR_INVALID_ARGUMENT, "unknown key kind %u", d->key_kind);
    if (!is_pow2_width(d->key_bytes))
        return fail(RSORT_ERROR_UNSUPPORTED, "key width %u bytes not supported (1,2,4,8,16)", d->key_bytes);
    if (d->key_kind == RSORT_KEY_FLOAT && d->key_bytes == 1)
        return fail(RSORT_ERROR_UNSUPPORTED, "no 1-byte floating key type");
    if (d->descending > 1) return fail(RSORT_ERROR_INVALID_ARGUMENT, "descending must be 0 or 1");
    if (d->flags & ~(RSORT_FLAG_NO_PASS_SKIP | RSORT_FLAG_GRAPH))
        return fail(RSORT_ERROR_INVALID_ARGUMENT, "unknown flags 0x%x", d->flags);
    return RSORT_SUCCESS;
}

rsort_status make_plan(const rsort_desc* d, uint64_t n, Plan* p) {
    const rsort_status st = check_desc(d);
    if (st != RSORT_SUCCESS) return st;
    memset(p, 0, sizeof(*p));
    p->kind = d->key_kind;
    p->kb = d->key_bytes;
    p->vb_user = d->value_bytes;
    p->desc = d->descending != 0;
    p->allow_skip = (d->flags & RSORT_FLAG_NO_PASS_SKIP) == 0;
    p->index_mode = d->value_bytes != 0 && !is_pow2_width(d->value_bytes);
    p->vb_core = p->index_mode ? 4u : d->value_bytes;
    if (p->index_mode && n > 0xFFFFFFFFull)
        return fail(RSORT_ERROR_UNSUPPORTED, "values of %u bytes need n < 2^32", d->value_bytes);
    p->wide = n > kNarrowMaxN;
    p->passes = p->kb;
    const TuneEntry te = tune_get(p->kb, p->vb_core);
    const uint32_t cand = (te.tiny_log2n && n < (1ull << te.tiny_log2n))     ? te.tiny_tile_cand
                          : (te.small_log2n && n < (1ull << te.small_log2n)) ? te.small_tile_cand
                                                                             : te.tile_cand;
    p->tile_cand = tile_cands_built((int)p->kb, (int)p->vb_core) ? cand : 0u;
    p->hist_bpc = te.hist_bpc;
    p->cfg = tile_cand((int)p->kb, (int)p->vb_core, (int)p->tile_cand);
    p->tile = (uint64_t)p->cfg.threads * p->cfg.items;
    p->num_tiles = n ? (n + p->tile - 1) / p->tile : 0;
    if (p->num_tiles > 0x7FFFFFFFull) return fail(RSORT_ERROR_UNSUPPORTED, "n too large");
    p->lb_wide = p->num_tiles <= lb_wide_max_tiles();

    const size_t off_sz = p->wide ? 8 : 4;
    size_t off = 0;
    p->off_tmp_keys = reserve(&off, (size_t)n * p->kb);
    p->off_tmp_vals = p->vb_core ? reserve(&off, (size_t)n * p->vb_core) : 0;
    if (p->index_mode) {
        p->off_idx = reserve(&off, (size_t)n * 4u);
        p->off_vals_copy = reserve(&off, (size_t)n * p->vb_user);
    }
    // control header, digit counts and lookback words are contiguous: one
    // fill kernel clears them at the start of every sort
    p->off_ctrl = reserve(&off, sizeof(Ctrl));
    p->off_counts = reserve(&off, (size_t)p->passes * kRadix * off_sz, 16);
    // lookback: the per-pass tile ticket counters, tile words, then per-pass
    // group words (kLbGroup tiles each)
    const size_t groups = (size_t)((p->num_tiles + kLbGroup - 1) / kLbGroup);
    p->lb_group_off = kPartTicketBytes + (size_t)p->num_tiles * kRadix * off_sz;
    p->lb_bytes = p->lb_group_off + (size_t)p->passes * groups * kRadix * off_sz;
    p->off_lb = reserve(&off, p->lb_bytes, 16);
    if (p->off_counts != p->off_ctrl + sizeof(Ctrl) || p->off_lb != p->off_counts + (size_t)p->passes * kRadix * off_sz)
        return fail(RSORT_ERROR_UNSUPPORTED, "workspace layout: cleared region is not contiguous");
    p->off_gp = reserve(&off, (size_t)p->passes * kRadix * off_sz, 4);
    {
        uint64_t g = (n + hist_keys_per_block_iter((int)p->kb) - 1) / hist_keys_per_block_iter((int)p->kb);
        if (g > kHistMaxGrid) g = kHistMaxGrid;
        if (g < 1) g = 1;
        p->off_hist_partials = reserve(&off, (size_t)g * p->passes * kRadix * 4u);
    }
    p->total = off + 256;  // slack so any 16-byte-aligned workspace pointer can be rounded up
    return RSORT_SUCCESS;
}

// Workspace of a segmented sort (no reference counterpart).  The long-segment
// tables are sized for the worst case of disjoint segments: at most
// n / (LDS capacity + 1) segments are too long for the LDS bins, and their
// onesweep tiles number at most n / tile plus one partial tile each.
struct SegPlan {
    uint32_t kind, kb, vb_user, vb_core;
    bool desc, index_mode, wide;
    uint32_t passes;
    uint64_t num_segments, chunks, max_large, max_tiles, tile;
    uint32_t lds_cap;  // longest segment sorted in LDS (the largest bin within the tuned threshold)
    size_t off_tmp_keys, off_tmp_vals, off_idx, off_vals_copy, off_ctrl, off_chunk, off_queue, off_desc;
    size_t off_tile_seg, off_seg_gp, off_lb, total;
};

rsort_status make_seg_plan(const rsort_desc* d, uint64_t n, uint32_t offset_bytes, uint64_t num_segments,
                           SegPlan* p) {
    const rsort_status st = check_desc(d);
    if (st != RSORT_SUCCESS) return st;
    if (offset_bytes != 4 && offset_bytes != 8)
        return fail(RSORT_ERROR_INVALID_ARGUMENT, "offset_bytes must be 4 or 8, got %u", offset_bytes);
    if (num_segments > 0xFFFFFFFEull) return fail(RSORT_ERROR_UNSUPPORTED, "more than 2^32-2 segments");
    if (offset_bytes == 4 && n > 0xFFFFFFFFull)
        return fail(RSORT_ERROR_UNSUPPORTED, "32-bit offsets cannot address n = %llu keys", (unsigned long long)n);
    memset(p, 0, sizeof(*p));
    p->kind = d->key_kind;
    p->kb = d->key_bytes;
    p->vb_user = d->value_bytes;
    p->desc = d->descending != 0;
    p->index_mode = d->value_bytes != 0 && !is_pow2_width(d->value_bytes);
    p->vb_core = p->index_mode ? 4u : d->value_bytes;
    if (p->index_mode && n > 0xFFFFFFFFull)
        return fail(RSORT_ERROR_UNSUPPORTED, "values of %u bytes need n < 2^32", d->value_bytes);
    p->wide = n > kNarrowMaxN;
    p->passes = p->kb;
    p->num_segments = num_segments;
    p->chunks = (num_segments + kSegChunk - 1) / kSegChunk;
    const TileCfg c = tile_cfg((int)p->kb, (int)p->vb_core);
    p->tile = (uint64_t)c.threads * c.items;
    const uint32_t thr = tune_get(p->kb, p->vb_core).lds_threshold;
    p->lds_cap = 0;
    for (int b = 0; b < kSegBins; ++b)
        if (seg_cap((int)p->kb, (int)p->vb_core, b) <= thr) p->lds_cap = seg_cap((int)p->kb, (int)p->vb_core, b);
    if (p->lds_cap < 1) p->lds_cap = 1;  // no LDS bin: every segment of 2 or more keys is long
    p->max_large = num_segments ? n / ((uint64_t)p->lds_cap + 1u) : 0;
    if (p->max_large > num_segments) p->max_large = num_segments;
    p->max_tiles = p->max_large ? (n + p->tile - 1) / p->tile + p->max_large : 0;
    if (p->max_tiles > 0x7FFFFFFFull) return fail(RSORT_ERROR_UNSUPPORTED, "n too large");

    const size_t off_sz = p->wide ? 8 : 4;
    size_t off = 0;
    if (p->max_large) {
        p->off_tmp_keys = reserve(&off, (size_t)n * p->kb);
        p->off_tmp_vals = p->vb_core ? reserve(&off, (size_t)n * p->vb_core) : 0;
    }
    if (p->index_mode) {
        p->off_idx = reserve(&off, (size_t)n * 4u);
        p->off_vals_copy = reserve(&off, (size_t)n * p->vb_user);
    }
    p->off_ctrl = reserve(&off, sizeof(SegCtrl));
    p->off_chunk = reserve(&off, (size_t)p->chunks * kSegRow * 4u);
    p->off_queue = reserve(&off, (size_t)num_segments * 4u);
    p->off_desc = reserve(&off, (size_t)p->max_large * 32u);
    p->off_tile_seg = reserve(&off, (size_t)p->max_tiles * 4u);
    p->off_seg_gp = reserve(&off, (size_t)p->max_large * p->passes * kRadix * off_sz);
    p->off_lb = reserve(&off, (size_t)p->max_tiles * kRadix * off_sz);
    p->total = off + 256;
    return RSORT_SUCCESS;
}

// The per-XCD ticket lock of each device: one word, zero when free.  The
// histogram of a sort whose passes may use per-XCD tile tickets takes it
// with a device-scope compare-and-swap, and the last tile of its last
// lookback pass clears it; a sort that finds it taken runs on the global
// ticket.  The variable's name is only a host-side handle: kernels get its
// device address, looked up once per device with that device current.
static __device__ uint32_t rsa_xcd_lock;

uint32_t* xcd_lock_word(int dev) {
    static std::mutex mu;
    static uint32_t* cache[64] = {nullptr};
    static bool done[64] = {false};
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!done[dev]) {
        void* p = nullptr;
        cache[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(rsa_xcd_lock)) == hipSuccess ? (uint32_t*)p : nullptr;
        (void)hipGetLastError();
        done[dev] = true;
    }
    return cache[dev];
}

int device_cus(int dev) {
    static std::mutex mu;
    static int cache[64] = {0};
    if (dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lk(mu);
    if (cache[dev] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cache[dev] = v;
    }
    return cache[dev];
}

// Phase-timing slots of debug builds (-DRSA_PHASE_TIMING): 64 copies x 32.
constexpr int kDbgSlots = 32;
unsigned long long* g_dbg = nullptr;

unsigned long long* debug_slots() {
#ifdef RSA_PHASE_TIMING
    static std::once_flag once;
    std::call_once(once, [] {
        if (hipMalloc(&g_dbg, 64 * kDbgSlots * sizeof(unsigned long long)) == hipSuccess)
            (void)hipMemset(g_dbg, 0, 64 * kDbgSlots * sizeof(unsigned long long));
        else
            g_dbg = nullptr;
    });
#endif
    return g_dbg;
}

hipError_t dispatch_core(const SortJob& j, hipStream_t s) {
    switch (j.kb) {
        case 1: return launch_core_kb1(j, s);
        case 2: return launch_core_kb2(j, s);
        case 4: return launch_core_kb4(j, s);
        case 8: return launch_core_kb8(j, s);
        case 16: return launch_core_kb16(j, s);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Graph replay (RSORT_FLAG_GRAPH).  The first call with a given set of
// arguments sorts directly and then captures the same enqueue, on a private
// stream of the device, into an executable graph; later calls with the same
// arguments replay it with one hipGraphLaunch on the caller's stream.  Every
// launch of the sequence is a kernel (the clears are fill kernels), so the
// graph holds kernel nodes only.  Up to kGraphCacheMax graphs are kept; the
// least recently used one is destroyed after its last replay has completed.
// (The reference captures and instantiates its enqueue on every call,
// radix_kernel.cuh:583-600.)
// ---------------------------------------------------------------------------
struct GraphKey {
    int dev;
    uint32_t entry;  // 0: rsort_sort, 1: rsort_segmented_sort
    rsort_desc desc;
    uint64_t keys, values, workspace, offsets;
    uint64_t n, num_segments, temp_bytes;
    uint32_t offset_bytes, pad;
};

struct GraphEntry {
    GraphKey key;
    hipGraphExec_t exec;  // null: the capture failed; these arguments run directly
    hipEvent_t done;      // recorded after every replay
    uint64_t used;
};

constexpr size_t kGraphCacheMax = 16;
std::mutex g_graph_mu;

std::vector<GraphEntry>& graph_cache() {
    static std::vector<GraphEntry>* v = new std::vector<GraphEntry>();  // kept until the process ends
    return *v;
}

hipStream_t capture_stream(int dev) {
    static hipStream_t cs[64] = {nullptr};
    if (dev < 0 || dev >= 64) return nullptr;
    if (!cs[dev] && hipStreamCreateWithFlags(&cs[dev], hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        cs[dev] = nullptr;
    }
    return cs[dev];
}

// `enqueue(s)` enqueues the whole sort on stream s; caller holds no lock.
template <typename Enqueue>
rsort_status run_graphed(const GraphKey& key, void* stream, Enqueue&& enqueue) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    static uint64_t tick = 0;
    auto& v = graph_cache();
    for (auto& g : v) {
        if (memcmp(&g.key, &key, sizeof(key)) != 0) continue;
        g.used = ++tick;
        if (!g.exec) return enqueue((hipStream_t)stream);
        hipError_t e = hipGraphLaunch(g.exec, (hipStream_t)stream);
        if (e == hipSuccess) e = hipEventRecord(g.done, (hipStream_t)stream);
        return e == hipSuccess ? RSORT_SUCCESS : hip_fail(e, "hipGraphLaunch");
    }
    const rsort_status st = enqueue((hipStream_t)stream);
    if (st != RSORT_SUCCESS) return st;
    GraphEntry ne{};
    ne.key = key;
    ne.used = ++tick;
    const hipStream_t cs = capture_stream(key.dev);
    if (cs) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            const rsort_status sc = enqueue(cs);
            e = hipStreamEndCapture(cs, &g);
            if (e == hipSuccess && sc != RSORT_SUCCESS) e = hipErrorUnknown;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&ne.exec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ne.done, hipEventDisableTiming);
        if (e != hipSuccess) {
            if (ne.exec) (void)hipGraphExecDestroy(ne.exec);
            ne.exec = nullptr;
            (void)hipGetLastError();
        }
    }
    if (v.size() >= kGraphCacheMax) {
        size_t lru = 0;
        for (size_t i = 1; i < v.size(); ++i)
            if (v[i].used < v[lru].used) lru = i;
        if (v[lru].exec) {
            (void)hipEventSynchronize(v[lru].done);
            (void)hipGraphExecDestroy(v[lru].exec);
            (void)hipEventDestroy(v[lru].done);
        }
        v.erase(v.begin() + (ptrdiff_t)lru);
    }
    v.push_back(ne);
    return RSORT_SUCCESS;
}

GraphKey graph_key(uint32_t entry, const rsort_desc* d, const void* keys, const void* values, const void* workspace,
                   size_t temp_bytes, uint64_t n, const void* offsets, uint32_t offset_bytes, uint64_t num_segments) {
    GraphKey k;
    memset(&k, 0, sizeof(k));
    (void)hipGetDevice(&k.dev);
    k.entry = entry;
    k.desc = *d;
    k.desc.flags &= ~RSORT_FLAG_GRAPH;
    k.keys = (uint64_t)(uintptr_t)keys;
    k.values = (uint64_t)(uintptr_t)values;
    k.workspace = (uint64_t)(uintptr_t)workspace;
    k.offsets = (uint64_t)(uintptr_t)offsets;
    k.n = n;
    k.num_segments = num_segments;
    k.temp_bytes = temp_bytes;
    k.offset_bytes = offset_bytes;
    return k;
}

// Control region of a plan: the control header, digit counts, tile tickets,
// lookback words, digit offsets and histogram partials -- everything from
// off_ctrl to the end of the layout -- in bytes.
size_t plan_ctrl_bytes(const Plan& p) { return p.total - 256u - p.off_ctrl; }

// Where one sort's buffers live: the ping-pong partners of the keys and
// values, the control region (plan offset off_ctrl maps to ctrl_region) and,
// for value widths moved by index, the workspace holding the index and the
// value copy (plan offsets off_idx, off_vals_copy from index_base).
struct SortBuffers {
    void* tmp_keys;
    void* tmp_vals;
    uint8_t* ctrl_region;
    uint8_t* index_base;
};

// Argument checks shared by the entry points (after the plan, before device work).
rsort_status check_sort_args(const Plan& p, const void* keys, const void* values) {
    if (!keys) return fail(RSORT_ERROR_INVALID_ARGUMENT, "keys is NULL");
    if (p.vb_user && !values) return fail(RSORT_ERROR_INVALID_ARGUMENT, "values is NULL but value_bytes = %u", p.vb_user);
    if ((uintptr_t)keys % p.kb) return fail(RSORT_ERROR_INVALID_ARGUMENT, "keys not aligned to %u bytes", p.kb);
    if (p.vb_user && is_pow2_width(p.vb_user) && (uintptr_t)values % p.vb_user)
        return fail(RSORT_ERROR_INVALID_ARGUMENT, "values not aligned to %u bytes", p.vb_user);
    return RSORT_SUCCESS;
}

// Enqueues one sort of plan p on `stream` (n >= 2, arguments checked).
rsort_status enqueue_sort(const Plan& p, void* keys, void* values, const SortBuffers& b, uint64_t n, void* stream,
                          unsigned long long* dbg, hipEvent_t* events, int max_events, int* n_events) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    const int cus = device_cus(dev);
    hipStream_t s = (hipStream_t)stream;
    // plan offsets from off_ctrl on map into the control region
    auto at = [&](size_t off) { return b.ctrl_region + (off - p.off_ctrl); };

    SortJob j{};
    j.tile_cand = p.tile_cand;
    j.hist_bpc = p.hist_bpc;
    j.kb = p.kb;
    j.vb = p.vb_core;
    j.wide = p.wide;
    j.allow_skip = p.allow_skip;
    j.lb_wide = p.lb_wide;
    j.xcd_tickets = xcd_tickets_allowed();
    j.xcd_lock = j.xcd_tickets ? xcd_lock_word(dev) : nullptr;
    j.early_agg = p.kb == 4 && p.vb_core == 4 && !p.wide && early_agg_min_tiles() > 0 &&
                  p.num_tiles >= early_agg_min_tiles();
    j.hist_direct_max = hist_direct_max();
    // the histogram may clear the control region itself unless the launch is
    // being captured (a graph's replays would all carry this one id)
    j.clear_id = 0u;
    if (hist_clear_allowed()) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusActive;
        if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) j.clear_id = next_clear_id();
        else (void)hipGetLastError();
    }
    j.late_claim = hist_late_claim();
    j.tw = make_tw_masks(p.kind, p.kb, p.desc);
    j.keys = keys;
    j.tmp_keys = b.tmp_keys;
    j.tmp_vals = p.vb_core ? b.tmp_vals : nullptr;
    j.vals = p.index_mode ? (void*)(b.index_base + p.off_idx) : values;
    j.ctrl = (Ctrl*)at(p.off_ctrl);
    j.counts = at(p.off_counts);
    j.gp = at(p.off_gp);
    j.hist_partials = (uint32_t*)at(p.off_hist_partials);
    j.part_tickets = (uint32_t*)at(p.off_lb);
    j.lookback = at(p.off_lb) + kPartTicketBytes;
    j.lb_bytes = p.lb_bytes;
    j.lb_groups = at(p.off_lb) + p.lb_group_off;
    j.n = n;
    j.num_tiles = p.num_tiles;
    j.passes = p.passes;
    j.num_cus = cus;
    j.dbg = dbg;
    j.events = events;
    j.max_events = max_events;
    j.n_events = n_events;

    if (p.index_mode) {
        e = launch_iota((uint32_t*)j.vals, n, cus, s);
        if (e != hipSuccess) return hip_fail(e, "index fill");
    }
    e = dispatch_core(j, s);
    if (e != hipSuccess) return hip_fail(e, "sort launch");
    if (p.index_mode) {
        void* copy = b.index_base + p.off_vals_copy;
        e = launch_gather(copy, values, (const uint32_t*)j.vals, n, p.vb_user, cus, s);
        if (e == hipSuccess) e = launch_copy(values, copy, n * p.vb_user, cus, s);
        if (e != hipSuccess) return hip_fail(e, "value gather");
    }
    return RSORT_SUCCESS;
}

// Plain sorts of at most the largest LDS bin's capacity (16384 keys for
// 4-byte keys and values) run as one workgroup of the segmented sort's LDS
// bins: one launch instead of the histogram and a launch per pass (the same
// stable LSD order, every pass in LDS).  RSORT_AMD_SMALL_LDS=0 keeps them on
// the onesweep path; so do the profiled entry (its stage events) and values
// moved by index.
bool small_lds_allowed() {
    static const bool v = [] {
        const char* e = getenv("RSORT_AMD_SMALL_LDS");
        return !(e && strcmp(e, "0") == 0);
    }();
    return v;
}

int small_lds_bin(const Plan& p, uint64_t n) {
    if (p.index_mode || n < 2 || !small_lds_allowed()) return -1;
    for (int b = 0; b < kSegBins; ++b)
        if (n <= seg_cap((int)p.kb, (int)p.vb_core, b)) return b;
    return -1;
}

hipError_t dispatch_req(const ReqJob& j, hipStream_t s) {
    switch (j.kb) {
        case 1: return launch_req_kb1(j, s);
        case 2: return launch_req_kb2(j, s);
        case 4: return launch_req_kb4(j, s);
        case 8: return launch_req_kb8(j, s);
        case 16: return launch_req_kb16(j, s);
        default: return hipErrorInvalidValue;
    }
}

rsort_status run_sort(const rsort_desc* d, void* keys, void* values, void* workspace, size_t* temp_bytes, uint64_t n,
                      void* stream, hipEvent_t* events, int max_events, int* n_events) {
    Plan p;
    rsort_status st = make_plan(d, n, &p);
    if (st != RSORT_SUCCESS) return st;
    if (!temp_bytes) return fail(RSORT_ERROR_INVALID_ARGUMENT, "temp_bytes is NULL");
    if (!workspace) {
        *temp_bytes = p.total;
        return RSORT_SUCCESS;
    }
    if (*temp_bytes < p.total)
        return fail(RSORT_ERROR_WORKSPACE_TOO_SMALL, "workspace of %zu bytes, %zu required", *temp_bytes, p.total);
    if (n <= 1) return RSORT_SUCCESS;
    st = check_sort_args(p, keys, values);
    if (st != RSORT_SUCCESS) return st;
    if ((uintptr_t)workspace % 16) return fail(RSORT_ERROR_INVALID_ARGUMENT, "workspace not 16-byte aligned");

    const int sbin = events ? -1 : small_lds_bin(p, n);
    if (sbin >= 0) {
        ReqJob j{};
        j.kb = p.kb;
        j.vb = p.vb_core;
        j.tw = make_tw_masks(p.kind, p.kb, p.desc);
        j.reqs = nullptr;
        j.count[sbin] = 1u;
        j.keys = keys;
        j.vals = values;
        j.n = n;
        auto enqueue_small = [&](hipStream_t s) -> rsort_status {
            const hipError_t e = dispatch_req(j, s);
            return e == hipSuccess ? RSORT_SUCCESS : hip_fail(e, "small sort launch");
        };
        if (d->flags & RSORT_FLAG_GRAPH) {
            const GraphKey k = graph_key(0u, d, keys, values, workspace, *temp_bytes, n, nullptr, 0u, 0u);
            return run_graphed(k, stream, enqueue_small);
        }
        return enqueue_small((hipStream_t)stream);
    }

    uint8_t* base = (uint8_t*)(((uintptr_t)workspace + 255u) & ~(uintptr_t)255u);
    SortBuffers b{};
    b.tmp_keys = base + p.off_tmp_keys;
    b.tmp_vals = p.vb_core ? base + p.off_tmp_vals : nullptr;
    b.ctrl_region = base + p.off_ctrl;
    b.index_base = base;
    if ((d->flags & RSORT_FLAG_GRAPH) && !events) {
        const GraphKey k = graph_key(0u, d, keys, values, workspace, *temp_bytes, n, nullptr, 0u, 0u);
        return run_graphed(k, stream, [&](hipStream_t s) {
            return enqueue_sort(p, keys, values, b, n, s, debug_slots(), nullptr, 0, nullptr);
        });
    }
    return enqueue_sort(p, keys, values, b, n, stream, debug_slots(), events, max_events, n_events);
}

// ---------------------------------------------------------------------------
// Allocation-free sort: the control region lives in a file-scope device
// array instead of a caller's workspace, and the ping-pong partners are the
// caller's own buffers.  Every device the library runs on holds its own copy
// of the array (kScratchBytes) until the process ends.  The array's name is
// a host-side handle only: kernels and copies get its device address, looked
// up once per device with that device current.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

uint8_t* scratch_ptr(int dev) {
    static std::mutex mu;
    static uint8_t* cache[64] = {nullptr};
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!cache[dev]) {
        uint8_t* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return nullptr;
        }
        cache[dev] = ptr;
    }
    return cache[dev];
}

rsort_status run_sort_noalloc(const rsort_desc* d, void* keys, void* values, void* keys_alt, void* values_alt,
                              uint64_t n, void* stream) {
    Plan p;
    rsort_status st = make_plan(d, n, &p);
    if (st != RSORT_SUCCESS) return st;
    if (p.index_mode)
        return fail(RSORT_ERROR_UNSUPPORTED,
                    "the allocation-free sort takes value widths 0, 1, 2, 4, 8 and 16 bytes, not %u", p.vb_user);
    const size_t need = plan_ctrl_bytes(p);
    if (need > kScratchBytes)
        return fail(RSORT_ERROR_UNSUPPORTED,
                    "n = %llu needs %zu bytes of sort state; the allocation-free sort holds %zu per device",
                    (unsigned long long)n, need, kScratchBytes);
    if (n <= 1) return RSORT_SUCCESS;
    st = check_sort_args(p, keys, values);
    if (st != RSORT_SUCCESS) return st;
    if (!keys_alt) return fail(RSORT_ERROR_INVALID_ARGUMENT, "keys_alt is NULL");
    if (p.vb_core && !values_alt) return fail(RSORT_ERROR_INVALID_ARGUMENT, "values_alt is NULL but value_bytes = %u",
                                              p.vb_user);
    if ((uintptr_t)keys_alt % p.kb) return fail(RSORT_ERROR_INVALID_ARGUMENT, "keys_alt not aligned to %u bytes", p.kb);
    if (p.vb_core && (uintptr_t)values_alt % p.vb_core)
        return fail(RSORT_ERROR_INVALID_ARGUMENT, "values_alt not aligned to %u bytes", p.vb_core);
    if (keys_alt == keys || (p.vb_core && values_alt == values))
        return fail(RSORT_ERROR_INVALID_ARGUMENT, "the alternate buffers must be distinct from keys and values");
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    uint8_t* sp = scratch_ptr(dev);
    if (!sp) return fail(RSORT_ERROR_DEVICE, "hipGetSymbolAddress of the sort state failed on device %d", dev);
    SortBuffers b{};
    b.tmp_keys = keys_alt;
    b.tmp_vals = p.vb_core ? values_alt : nullptr;
    b.ctrl_region = sp;
    b.index_base = nullptr;
    return enqueue_sort(p, keys, values, b, n, stream, nullptr, nullptr, 0, nullptr);
}

// Side streams of the segmented sort: per device, one set for direct calls,
// whose fork and join are enqueued under a per-device lock (so that calls
// from several host threads cannot interleave the events' records and
// waits); and, per host thread and device, one set for calls made while the
// caller's stream is being captured (the library's own graph capture, or a
// caller's): a side stream joined into a capture stays in it until that
// capture ends, so two threads capturing at once must not share one.
// Created on first use, non-blocking, at the default priority; kept until
// the process ends.  RSORT_AMD_SEG_STREAMS=0 keeps every launch on the
// caller's stream.
struct SideSet {
    hipStream_t s[3];
    hipEvent_t fork, join[3];
    bool ok;
};

bool seg_streams_allowed() {
    static const bool v = [] {
        const char* e = getenv("RSORT_AMD_SEG_STREAMS");
        return !(e && strcmp(e, "0") == 0);
    }();
    return v;
}

void side_set_init(SideSet& ss) {
    hipError_t e = hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipStreamCreateWithFlags(&ss.s[i], hipStreamNonBlocking);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&ss.join[i], hipEventDisableTiming);
    ss.ok = e == hipSuccess;
    (void)hipGetLastError();
}

SideSet* side_set(int dev, bool capture) {
    if (dev < 0 || dev >= 64 || !seg_streams_allowed()) return nullptr;
    if (capture) {
        thread_local SideSet tsets[64];
        thread_local bool tdone[64] = {false};
        if (!tdone[dev]) {
            tdone[dev] = true;
            // first use inside a capture: creating streams and events is not
            // capturable work, so this thread's capture mode is relaxed for it
            hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
            const bool swapped = hipThreadExchangeStreamCaptureMode(&mode) == hipSuccess;
            side_set_init(tsets[dev]);
            if (swapped) (void)hipThreadExchangeStreamCaptureMode(&mode);
            (void)hipGetLastError();
        }
        return tsets[dev].ok ? &tsets[dev] : nullptr;
    }
    static std::mutex mu;
    static SideSet sets[64];
    static bool done[64] = {false};
    std::lock_guard<std::mutex> lk(mu);
    if (!done[dev]) {
        done[dev] = true;
        side_set_init(sets[dev]);
    }
    return sets[dev].ok ? &sets[dev] : nullptr;
}

std::mutex& side_lock(int dev) {
    static std::mutex mu[64];
    return mu[(dev >= 0 && dev < 64) ? dev : 0];
}

hipError_t dispatch_seg(const SegJob& j, hipStream_t s) {
    switch (j.kb) {
        case 1: return launch_seg_kb1(j, s);
        case 2: return launch_seg_kb2(j, s);
        case 4: return launch_seg_kb4(j, s);
        case 8: return launch_seg_kb8(j, s);
        case 16: return launch_seg_kb16(j, s);
        default: return hipErrorInvalidValue;
    }
}

rsort_status run_segmented(const rsort_desc* d, void* keys, void* values, uint64_t n, const void* offsets,
                           uint32_t offset_bytes, uint64_t num_segments, void* workspace, size_t* temp_bytes,
                           void* stream, hipEvent_t* events, int max_events, int* n_events) {
    SegPlan p;
    rsort_status st = make_seg_plan(d, n, offset_bytes, num_segments, &p);
    if (st != RSORT_SUCCESS) return st;
    if (!temp_bytes) return fail(RSORT_ERROR_INVALID_ARGUMENT, "temp_bytes is NULL");
    if (!workspace) {
        *temp_bytes = p.total;
        return RSORT_SUCCESS;
    }
    if (*temp_bytes < p.total)
        return fail(RSORT_ERROR_WORKSPACE_TOO_SMALL, "workspace of %zu bytes, %zu required", *temp_bytes, p.total);
    if (n <= 1 || num_segments == 0) return RSORT_SUCCESS;
    if (!keys) return fail(RSORT_ERROR_INVALID_ARGUMENT, "keys is NULL");
    if (!offsets) return fail(RSORT_ERROR_INVALID_ARGUMENT, "offsets is NULL");
    if (p.vb_user && !values) return fail(RSORT_ERROR_INVALID_ARGUMENT, "values is NULL but value_bytes = %u", p.vb_user);
    if ((uintptr_t)keys % p.kb) return fail(RSORT_ERROR_INVALID_ARGUMENT, "keys not aligned to %u bytes", p.kb);
    if (p.vb_user && is_pow2_width(p.vb_user) && (uintptr_t)values % p.vb_user)
        return fail(RSORT_ERROR_INVALID_ARGUMENT, "values not aligned to %u bytes", p.vb_user);
    if ((uintptr_t)offsets % offset_bytes)
        return fail(RSORT_ERROR_INVALID_ARGUMENT, "offsets not aligned to %u bytes", offset_bytes);
    if ((uintptr_t)workspace % 16) return fail(RSORT_ERROR_INVALID_ARGUMENT, "workspace not 16-byte aligned");

    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    const int cus = device_cus(dev);

    uint8_t* base = (uint8_t*)(((uintptr_t)workspace + 255u) & ~(uintptr_t)255u);
    SegJob j{};
    j.kb = p.kb;
    j.vb = p.vb_core;
    j.wide = p.wide;
    j.tw = make_tw_masks(p.kind, p.kb, p.desc);
    j.keys = keys;
    j.vals = p.index_mode ? (void*)(base + p.off_idx) : values;
    j.tmp_keys = p.max_large ? base + p.off_tmp_keys : nullptr;
    j.tmp_vals = (p.max_large && p.vb_core) ? base + p.off_tmp_vals : nullptr;
    j.offsets = offsets;
    j.off64 = offset_bytes == 8 ? 1u : 0u;
    j.num_segments = num_segments;
    j.n = n;
    j.ctrl = (SegCtrl*)(base + p.off_ctrl);
    j.chunk = (uint32_t*)(base + p.off_chunk);
    j.queue = (uint32_t*)(base + p.off_queue);
    j.desc = base + p.off_desc;
    j.tile_seg = (uint32_t*)(base + p.off_tile_seg);
    j.seg_gp = base + p.off_seg_gp;
    j.lookback = base + p.off_lb;
    j.chunks = (uint32_t)p.chunks;
    j.max_large = p.max_large;
    j.max_tiles = p.max_tiles;
    j.lds_threshold = p.lds_cap;
    j.num_cus = cus;
    j.dbg = debug_slots();
    j.events = events;
    j.max_events = max_events;
    j.n_events = n_events;

    auto enqueue = [&](hipStream_t s) -> rsort_status {
        hipError_t e2;
        if (p.index_mode) {
            e2 = launch_iota((uint32_t*)j.vals, n, cus, s);
            if (e2 != hipSuccess) return hip_fail(e2, "index fill");
        }
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        (void)hipStreamIsCapturing(s, &cap);
        SideSet* ss = events ? nullptr : side_set(dev, cap != hipStreamCaptureStatusNone);
        SegJob js = j;
        if (ss) {
            for (int i = 0; i < 3; ++i) {
                js.side[i] = ss->s[i];
                js.join[i] = ss->join[i];
            }
            js.fork = ss->fork;
            // the direct set's streams and events are shared by every host thread
            std::unique_lock<std::mutex> lk(side_lock(dev), std::defer_lock);
            if (cap == hipStreamCaptureStatusNone) lk.lock();
            e2 = dispatch_seg(js, s);
        } else {
            e2 = dispatch_seg(js, s);
        }
        if (e2 != hipSuccess) return hip_fail(e2, "segmented sort launch");
        if (p.index_mode) {
            void* copy = base + p.off_vals_copy;
            e2 = launch_gather(copy, values, (const uint32_t*)j.vals, n, p.vb_user, cus, s);
            if (e2 == hipSuccess) e2 = launch_copy(values, copy, n * p.vb_user, cus, s);
            if (e2 != hipSuccess) return hip_fail(e2, "value gather");
        }
        return RSORT_SUCCESS;
    };
    if ((d->flags & RSORT_FLAG_GRAPH) && !events) {
        const GraphKey k = graph_key(1u, d, keys, values, workspace, *temp_bytes, n, offsets, offset_bytes,
                                     num_segments);
        return run_graphed(k, stream, enqueue);
    }
    return enqueue((hipStream_t)stream);
}

}  // namespace

hipError_t launch_iota(uint32_t* p, uint64_t n, int num_cus, hipStream_t s) {
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 1024, num_cus, 8)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_gather(void* dst, const void* src, const uint32_t* idx, uint64_t n, uint32_t rec_bytes,
                         int num_cus, hipStream_t s) {
    const uintptr_t al = (uintptr_t)dst | (uintptr_t)src | rec_bytes;
    const dim3 grid(grid_for(n * rec_bytes, 4096, num_cus, 8)), block(256);
    if ((al & 3u) == 0) {
        hipLaunchKernelGGL(k_gather<uint32_t>, grid, block, 0, s, (uint32_t*)dst, (const uint32_t*)src, idx, n,
                           rec_bytes / 4u);
    } else if ((al & 1u) == 0) {
        hipLaunchKernelGGL(k_gather<uint16_t>, grid, block, 0, s, (uint16_t*)dst, (const uint16_t*)src, idx, n,
                           rec_bytes / 2u);
    } else {
        hipLaunchKernelGGL(k_gather<uint8_t>, grid, block, 0, s, (uint8_t*)dst, (const uint8_t*)src, idx, n,
                           rec_bytes);
    }
    return hipGetLastError();
}

// Keys of n records (record_bytes apart, key at key_offset) into dst, in the
// widest unit that the record base, stride, offset and key width allow.
hipError_t launch_extract(void* dst, const void* rec, uint64_t n, uint32_t record_bytes, uint32_t key_offset,
                          uint32_t key_bytes, int num_cus, hipStream_t s) {
    const uintptr_t al = (uintptr_t)rec | record_bytes | key_offset | key_bytes;
    const dim3 grid(grid_for(n * key_bytes, 4096, num_cus, 8)), block(256);
    if ((al & 15u) == 0) {
        hipLaunchKernelGGL(k_extract_keys<uint4>, grid, block, 0, s, (uint4*)dst, (const uint4*)rec, n,
                           record_bytes / 16u, key_offset / 16u, key_bytes / 16u);
    } else if ((al & 7u) == 0) {
        hipLaunchKernelGGL(k_extract_keys<uint64_t>, grid, block, 0, s, (uint64_t*)dst, (const uint64_t*)rec, n,
                           record_bytes / 8u, key_offset / 8u, key_bytes / 8u);
    } else if ((al & 3u) == 0) {
        hipLaunchKernelGGL(k_extract_keys<uint32_t>, grid, block, 0, s, (uint32_t*)dst, (const uint32_t*)rec, n,
                           record_bytes / 4u, key_offset / 4u, key_bytes / 4u);
    } else if ((al & 1u) == 0) {
        hipLaunchKernelGGL(k_extract_keys<uint16_t>, grid, block, 0, s, (uint16_t*)dst, (const uint16_t*)rec, n,
                           record_bytes / 2u, key_offset / 2u, key_bytes / 2u);
    } else {
        hipLaunchKernelGGL(k_extract_keys<uint8_t>, grid, block, 0, s, (uint8_t*)dst, (const uint8_t*)rec, n,
                           record_bytes, key_offset, key_bytes);
    }
    return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, int num_cus, hipStream_t s) {
    const uintptr_t al = (uintptr_t)dst | (uintptr_t)src | (uintptr_t)bytes;
    const dim3 block(256);
    if ((al & 15u) == 0) {
        hipLaunchKernelGGL(k_copy<uint4>, dim3(grid_for(bytes / 16, 1024, num_cus, 8)), block, 0, s, (uint4*)dst,
                           (const uint4*)src, bytes / 16);
    } else if ((al & 3u) == 0) {
        hipLaunchKernelGGL(k_copy<uint32_t>, dim3(grid_for(bytes / 4, 1024, num_cus, 8)), block, 0, s,
                           (uint32_t*)dst, (const uint32_t*)src, bytes / 4);
    } else {
        hipLaunchKernelGGL(k_copy<uint8_t>, dim3(grid_for(bytes, 1024, num_cus, 8)), block, 0, s, (uint8_t*)dst,
                           (const uint8_t*)src, bytes);
    }
    return hipGetLastError();
}

}  // namespace rsa

extern "C" {

rsort_status rsort_sort(const rsort_desc* desc, void* keys, void* values, void* workspace, size_t* temp_bytes,
                        uint64_t n, void* stream) {
    return rsa::run_sort(desc, keys, values, workspace, temp_bytes, n, stream, nullptr, 0, nullptr);
}

rsort_status rsort_sort_noalloc(const rsort_desc* desc, void* keys, void* values, void* keys_alt, void* values_alt,
                                uint64_t n, void* stream) {
    return rsa::run_sort_noalloc(desc, keys, values, keys_alt, values_alt, n, stream);
}

// Records sorted by one key field: the keys are copied out of the records
// into the workspace, then sorted as keys with the whole record as the value
// (records of a power-of-two size up to 16 bytes move directly, others by a
// 32-bit index and a gather), so the record's own key bytes move with it.
rsort_status rsort_sort_records(const rsort_desc* desc, void* records, uint64_t n, uint32_t record_bytes,
                                uint32_t key_offset, void* workspace, size_t* temp_bytes, void* stream) {
    if (!desc) return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "descriptor is NULL");
    if (desc->value_bytes)
        return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "value_bytes must be 0: the record is the value");
    if (!temp_bytes) return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "temp_bytes is NULL");
    if (record_bytes == 0 || (uint64_t)key_offset + desc->key_bytes > record_bytes)
        return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "key of %u bytes at offset %u does not fit a %u-byte record",
                         desc->key_bytes, key_offset, record_bytes);
    rsort_desc d = *desc;
    d.value_bytes = record_bytes;
    size_t inner = 0;
    rsort_status st = rsa::run_sort(&d, nullptr, nullptr, nullptr, &inner, n, stream, nullptr, 0, nullptr);
    if (st != RSORT_SUCCESS) return st;
    const size_t keys_bytes = ((size_t)n * desc->key_bytes + 255u) & ~(size_t)255u;
    const size_t total = 256u + keys_bytes + inner;
    if (!workspace) {
        *temp_bytes = total;
        return RSORT_SUCCESS;
    }
    if (*temp_bytes < total)
        return rsa::fail(RSORT_ERROR_WORKSPACE_TOO_SMALL, "workspace of %zu bytes, %zu required", *temp_bytes, total);
    if (n <= 1) return RSORT_SUCCESS;
    if (!records) return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "records is NULL");
    if ((uintptr_t)workspace % 16) return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "workspace not 16-byte aligned");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return rsa::hip_fail(e, "hipGetDevice");
    uint8_t* keys = (uint8_t*)(((uintptr_t)workspace + 255u) & ~(uintptr_t)255u);
    e = rsa::launch_extract(keys, records, n, record_bytes, key_offset, desc->key_bytes, rsa::device_cus(dev),
                            (hipStream_t)stream);
    if (e != hipSuccess) return rsa::hip_fail(e, "key extraction");
    size_t inner_bytes = total - (size_t)(keys + keys_bytes - (uint8_t*)workspace);
    return rsa::run_sort(&d, keys, records, keys + keys_bytes, &inner_bytes, n, stream, nullptr, 0, nullptr);
}

rsort_status rsort_sort_profiled(const rsort_desc* desc, void* keys, void* values, void* workspace,
                                 size_t* temp_bytes, uint64_t n, void* stream, float* stage_ms, int max_stages,
                                 int* n_stages) {
    if (n_stages) *n_stages = 0;
    if (!workspace) return rsa::run_sort(desc, keys, values, workspace, temp_bytes, n, stream, nullptr, 0, nullptr);
    constexpr int kMaxEv = 40;
    hipEvent_t ev[kMaxEv];
    int made = 0;
    for (; made < kMaxEv; ++made) {
        if (hipEventCreate(&ev[made]) != hipSuccess) break;
    }
    int used = 0;
    rsort_status st = rsa::run_sort(desc, keys, values, workspace, temp_bytes, n, stream, ev, made, &used);
    if (st == RSORT_SUCCESS && used > 0) {
        hipError_t e = hipEventSynchronize(ev[used - 1]);
        if (e != hipSuccess) st = rsa::hip_fail(e, "hipEventSynchronize");
        int k = 0;
        for (int i = 1; st == RSORT_SUCCESS && i < used && k < max_stages; ++i, ++k) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[i - 1], ev[i]);
            if (stage_ms) stage_ms[k] = ms;
        }
        if (n_stages) *n_stages = k;
    }
    for (int i = 0; i < made; ++i) (void)hipEventDestroy(ev[i]);
    return st;
}

rsort_status rsort_segmented_sort(const rsort_desc* desc, void* keys, void* values, uint64_t n, const void* offsets,
                                  uint32_t offset_bytes, uint64_t num_segments, void* workspace, size_t* temp_bytes,
                                  void* stream) {
    return rsa::run_segmented(desc, keys, values, n, offsets, offset_bytes, num_segments, workspace, temp_bytes,
                              stream, nullptr, 0, nullptr);
}

rsort_status rsort_segmented_plan(const rsort_desc* desc, uint64_t n, uint32_t offset_bytes, uint64_t num_segments,
                                  rsort_segmented_plan_info* out) {
    if (!out) return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "plan output is NULL");
    rsa::SegPlan p;
    rsort_status st = rsa::make_seg_plan(desc, n, offset_bytes, num_segments, &p);
    if (st != RSORT_SUCCESS) return st;
    for (int b = 0; b < RSORT_SEGMENT_BINS; ++b) out->bin_capacity[b] = rsa::seg_cap((int)p.kb, (int)p.vb_core, b);
    out->tile = (uint32_t)p.tile;
    out->max_long_segments = p.max_large;
    out->max_tiles = p.max_tiles;
    out->wide = p.wide;
    out->index_mode = p.index_mode;
    out->workspace_bytes = p.total;
    return RSORT_SUCCESS;
}

rsort_status rsort_plan(const rsort_desc* desc, uint64_t n, rsort_plan_info* out) {
    if (!out) return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "plan output is NULL");
    rsa::Plan p;
    rsort_status st = rsa::make_plan(desc, n, &p);
    if (st != RSORT_SUCCESS) return st;
    out->passes = p.passes;
    out->threads = p.cfg.threads;
    out->items_per_thread = p.cfg.items;
    out->tile = (uint32_t)p.tile;
    out->num_tiles = p.num_tiles;
    out->wide = p.wide;
    out->index_mode = p.index_mode;
    out->workspace_bytes = p.total;
    out->lookback_bytes = p.lb_bytes;
    return RSORT_SUCCESS;
}

rsort_status rsort_tuning_get(uint32_t key_bytes, uint32_t value_bytes, rsort_tuning* out) {
    if (!out) return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "tuning output is NULL");
    if (rsa::kb_index(key_bytes) < 0 || rsa::vb_index(value_bytes) < 0)
        return rsa::fail(RSORT_ERROR_UNSUPPORTED, "no tuning entry for key %u / value %u bytes", key_bytes, value_bytes);
    const rsa::TuneEntry t = rsa::tune_get(key_bytes, value_bytes);
    out->tile_candidate = t.tile_cand;
    out->hist_blocks_per_cu = t.hist_bpc;
    out->lds_threshold = t.lds_threshold;
    out->small_tile_candidate = t.small_tile_cand;
    out->small_log2n = t.small_log2n;
    out->tiny_tile_candidate = t.tiny_tile_cand;
    out->tiny_log2n = t.tiny_log2n;
    out->reserved = 0;
    return RSORT_SUCCESS;
}

rsort_status rsort_tuning_set(uint32_t key_bytes, uint32_t value_bytes, const rsort_tuning* t) {
    if (!t || t->reserved) return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "tuning entry is NULL or reserved != 0");
    const int ki = rsa::kb_index(key_bytes), vi = rsa::vb_index(value_bytes);
    if (ki < 0 || vi < 0)
        return rsa::fail(RSORT_ERROR_UNSUPPORTED, "no tuning entry for key %u / value %u bytes", key_bytes, value_bytes);
    const rsa::TuneEntry e{t->tile_candidate, t->hist_blocks_per_cu, t->lds_threshold, t->small_tile_candidate,
                           t->small_log2n,    t->tiny_tile_candidate, t->tiny_log2n};
    if (!rsa::tune_valid(key_bytes, value_bytes, e))
        return rsa::fail(RSORT_ERROR_INVALID_ARGUMENT, "tuning out of range (tile candidates < %d, 1..8 histogram "
                         "workgroups per CU, LDS threshold <= %u, small_log2n and tiny_log2n <= 40)", rsa::kTileCands,
                         rsa::seg_cap((int)key_bytes, (int)value_bytes, rsa::kSegBins - 1));
    std::lock_guard<std::mutex> lk(rsa::g_tune_mu);
    rsa::tune_ensure_locked();
    rsa::g_tune[ki][vi] = e;
    return RSORT_SUCCESS;
}

int rsort_tuning_load(const char* path) {
    std::lock_guard<std::mutex> lk(rsa::g_tune_mu);
    if (!path) return rsa::tune_load_default_locked();
    rsa::tune_reset_locked();
    rsa::g_tune_ready = true;
    return *path ? rsa::tune_read_file(path) : 0;
}

const char* rsort_status_string(rsort_status s) {
    switch (s) {
        case RSORT_SUCCESS: return "success";
        case RSORT_ERROR_INVALID_ARGUMENT: return "invalid argument";
        case RSORT_ERROR_UNSUPPORTED: return "unsupported";
        case RSORT_ERROR_WORKSPACE_TOO_SMALL: return "workspace too small";
        case RSORT_ERROR_DEVICE: return "device error";
        case RSORT_ERROR_BUSY: return "busy";
    }
    return "unknown status";
}

const char* rsort_last_error(void) { return rsa::g_last_error; }

int rsort_debug_phase_slots(unsigned long long* out, int max_slots, int reset) {
    unsigned long long* d = rsa::debug_slots();
    if (!d || !out) return 0;
    unsigned long long h[64 * rsa::kDbgSlots];
    if (hipDeviceSynchronize() != hipSuccess) return 0;
    if (hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 0;
    const int k = max_slots < rsa::kDbgSlots ? max_slots : rsa::kDbgSlots;
    for (int i = 0; i < k; ++i) {
        unsigned long long v = 0;
        for (int c = 0; c < 64; ++c) v += h[c * rsa::kDbgSlots + i];
        out[i] = v;
    }
    if (reset) (void)hipMemset(d, 0, sizeof(h));
    return k;
}

int rsort_abi_version(void) { return RSORT_ABI_VERSION; }

}  // extern "C"
