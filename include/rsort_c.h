/*
 * rsort_c.h - C ABI of the MI355X (gfx950) onesweep LSD radix sort.
 *
 * Plain pointers and sizes only; no HIP or torch types.  The C++ drop-in
 * templates in <rsort/rsort.hpp> forward to these entry points, and a
 * ctypes / cgo / JNI binding can call them directly.
 *
 * Reference interface replaced (AwardOfSky/CUDA-Fast-Sort v0.4.0):
 *   rsort::sort / sort_descending / sort_pairs / sort_pairs_descending
 *     include/rsort/rsort.cuh:13-92
 *   size query convention (workspace == NULL -> write *temp_bytes, no device work)
 *     include/rsort/detail/radix_kernel.cuh:501-507
 *   in-place result, workspace holds the ping-pong buffer
 *     include/rsort/detail/radix_kernel.cuh:35-109, 468-473; README.md:100
 */
#ifndef RSORT_C_H_
#define RSORT_C_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSORT_ABI_VERSION 1

typedef enum rsort_status {
    RSORT_SUCCESS = 0,
    RSORT_ERROR_INVALID_ARGUMENT = 1, /* null pointer, bad descriptor, misaligned buffer */
    RSORT_ERROR_UNSUPPORTED = 2,      /* key width / kind / size outside what is built */
    RSORT_ERROR_WORKSPACE_TOO_SMALL = 3,
    RSORT_ERROR_DEVICE = 4,           /* a HIP runtime call failed */
    RSORT_ERROR_BUSY = 5              /* serving queue full / stopped */
} rsort_status;

/* Key kinds; the key is ordered by its value (radix_traits twiddle,
 * include/rsort/detail/tuning.cuh:297-397). */
typedef enum rsort_key_kind {
    RSORT_KEY_UNSIGNED = 0, /* 1,2,4,8,16 bytes (16 = unsigned __int128) */
    RSORT_KEY_SIGNED = 1,   /* 1,2,4,8,16 bytes, two's complement */
    RSORT_KEY_FLOAT = 2     /* 2 (IEEE half), 4, 8 bytes; 16 = long double bit
                               pattern ordered as a sign-magnitude 128-bit word
                               with the sign at bit 127, as the reference does
                               (radix_kernel.cuh:351-356, tuning.cuh:324) */
} rsort_key_kind;

/* Descriptor flags */
#define RSORT_FLAG_NO_PASS_SKIP 1u /* always run every pass (no uniform-digit skip) */
/* rsort_sort / rsort_segmented_sort: replay the call from a cached hipGraph.
 * The first call with a given set of arguments (descriptor, pointers, sizes,
 * workspace size, device) sorts directly and captures its launch sequence
 * into an executable graph; later calls with the same arguments launch that
 * graph on the caller's stream (one launch instead of one per kernel).  The
 * library keeps 16 such graphs (least recently used one dropped).  The
 * reference captures and instantiates a graph on every call
 * (include/rsort/detail/radix_kernel.cuh:583-600). */
#define RSORT_FLAG_GRAPH 2u

typedef struct rsort_desc {
    uint32_t key_kind;    /* rsort_key_kind */
    uint32_t key_bytes;   /* 1, 2, 4, 8 or 16 */
    uint32_t value_bytes; /* 0 = keys only; any width (1,2,4,8,16 native, others by index) */
    uint32_t descending;  /* 0 ascending, 1 descending (stable on the complemented key) */
    uint32_t flags;       /* RSORT_FLAG_* */
    uint32_t reserved;    /* must be 0 */
} rsort_desc;

/*
 * Stable LSD radix sort of n keys (and their values) in place.
 *
 * Replaces rsort::sort / sort_descending (rsort.cuh:13-50, value_bytes == 0)
 * and rsort::sort_pairs / sort_pairs_descending (rsort.cuh:53-92).
 *
 *   workspace == NULL : writes the required size to *temp_bytes, no device work.
 *   workspace != NULL : *temp_bytes must be >= that size; enqueues the sort on
 *                       `stream` (a hipStream_t, NULL = default stream) and
 *                       returns without synchronising.  keys/values are
 *                       overwritten with the sorted result.
 *
 * keys must be aligned to the key width (16 for 16-byte keys), values to
 * their width when it is a power of two.  n == 0 and n == 1 are no-ops.
 * One workspace must not be used by two sorts in flight at the same time.
 */
rsort_status rsort_sort(const rsort_desc* desc, void* keys, void* values, void* workspace, size_t* temp_bytes,
                        uint64_t n, void* stream);

/*
 * Allocation-free sort.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory or hand over a workspace at call time
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * allocation).  Same result as rsort_sort (stable, in place in keys /
 * values); keys_alt and values_alt are the caller's buffers of n keys and n
 * values (values_alt ignored without values), overwritten as the ping-pong
 * partners that rsort_sort keeps in its workspace.  The rest of the sort's
 * state (control header, digit counts, tile tickets, lookback words, digit
 * offsets) lives in a 32 MiB device array of the library, one per device,
 * held until the process ends; nothing is allocated per call.  Sorts whose
 * state exceeds it return RSORT_ERROR_UNSUPPORTED before any device work
 * (it holds u32 keys or u32/u32 pairs up to 2^28 keys, u64/u64 up to 2^26,
 * u128/u64 up to 2^24); so do value widths other than 0, 1, 2, 4, 8, 16.
 * Every call on one device shares that array: calls that overlap on a device
 * (two streams, two host threads) race and must be serialised by the caller.
 */
rsort_status rsort_sort_noalloc(const rsort_desc* desc, void* keys, void* values, void* keys_alt, void* values_alt,
                                uint64_t n, void* stream);

/*
 * Same as rsort_sort, then synchronises the stream and reports the device
 * time of each launched stage in milliseconds, in launch order:
 *   [0] control-block clear, [1] histogram, [2..2+P) one per pass,
 *   [2+P] copy-back (present when the pass count is odd or passes may be skipped).
 * Measurement helper for benchmarks; no reference counterpart (the reference
 * times the whole call with cudaEvents, benchmark/benchmark_sort.cuh:623-634).
 */
rsort_status rsort_sort_profiled(const rsort_desc* desc, void* keys, void* values, void* workspace,
                                 size_t* temp_bytes, uint64_t n, void* stream, float* stage_ms, int max_stages,
                                 int* n_stages);

/*
 * Bandwidth ceiling of the device (measurement helper for benchmarks; no
 * reference counterpart): one launch of a streaming kernel over `bytes`
 * (a multiple of 16) in 16-byte units per lane, blocks_per_cu (1..32)
 * 256-thread workgroups per CU.  mode bits 0-1: 0 reads src (dst: a 4 KiB
 * sink that is written only if the data hits an unlikely pattern), 1 writes
 * dst, 2 copies src to dst, 3 scatter-copies src to a separate dst (bytes a
 * multiple of 64 KiB: each 64 KiB tile of 4-byte keys is written as 256 runs
 * of 64 keys, run d at position tile * 64 of the d-th of 256 equal regions,
 * the write pattern of one radix pass); bit 2: non-temporal loads and
 * stores; bit 3 (modes 0-2): each workgroup takes contiguous 32 KiB chunks
 * (else every lane strides by the grid's width).  Enqueued on `stream`; time
 * it with events.
 */
rsort_status rsort_stream_bandwidth(uint32_t mode, void* dst, const void* src, uint64_t bytes,
                                    uint32_t blocks_per_cu, void* stream);

/*
 * Array-of-structs sort: n records of record_bytes each, sorted stably in
 * place by the key field of desc->key_bytes bytes at key_offset in every
 * record (desc->value_bytes must be 0: the rest of the record is the value).
 * Same query / workspace / stream convention as rsort_sort.  The keys are
 * copied out of the records into the workspace, then the records move as the
 * values of that key array (directly for record sizes 1/2/4/8/16, by a 32-bit
 * index and a gather otherwise, which needs n < 2^32).  Records of a
 * power-of-two size up to 16 must be aligned to that size.
 *
 * Replaces the reference's planned AoS interface: rsort.cuh:95-102 (AoS input
 * converted to SoA) and its C++26 reflection prototype, which splits a record
 * into its key member and the remaining members
 * (include/rsort/experimental/reflection.h:150-184).
 */
rsort_status rsort_sort_records(const rsort_desc* desc, void* records, uint64_t n, uint32_t record_bytes,
                                uint32_t key_offset, void* workspace, size_t* temp_bytes, void* stream);

/* Plan of a sort: what the host would launch for (desc, n).  Host only. */
typedef struct rsort_plan_info {
    uint32_t passes;          /* = key_bytes */
    uint32_t threads;         /* workgroup size of the pass kernel */
    uint32_t items_per_thread;
    uint32_t tile;            /* keys per workgroup = threads * items */
    uint64_t num_tiles;
    uint32_t wide;            /* 1: 64-bit offsets / lookback words */
    uint32_t index_mode;      /* 1: values moved by a 32-bit index + gather */
    uint64_t workspace_bytes;
    uint64_t lookback_bytes;
} rsort_plan_info;

rsort_status rsort_plan(const rsort_desc* desc, uint64_t n, rsort_plan_info* out);

/*
 * Segmented sort: sorts every segment [offsets[i], offsets[i+1]), i < num_segments,
 * of keys (and values) independently, stably and in place; keys outside every
 * segment are left alone.  No counterpart in the reference, which sorts one
 * array per call: this is its sort applied to a batch of arrays packed back
 * to back, in one enqueue (the same order inside each segment as rsort_sort).
 *
 *   offsets      device array of num_segments + 1 entries of offset_bytes (4 or 8)
 *                bytes; indices into keys/values, non-decreasing, <= n.
 *                Offsets are clamped to [0, n]; overlapping segments give an
 *                unspecified order (never an out-of-bounds access).
 *   n            extent of keys / values in elements
 *   workspace    as rsort_sort: NULL -> size query into *temp_bytes
 *
 * A segment of at most bin_capacity[8] keys (rsort_segmented_plan; 16384 for
 * 4-byte keys and values) is sorted by one workgroup in LDS, longer ones by
 * the onesweep passes.  Stream-ordered, no host synchronisation, so the whole
 * call can be captured into a hipGraph.
 */
rsort_status rsort_segmented_sort(const rsort_desc* desc, void* keys, void* values, uint64_t n, const void* offsets,
                                  uint32_t offset_bytes, uint64_t num_segments, void* workspace, size_t* temp_bytes,
                                  void* stream);

#define RSORT_SEGMENT_BINS 9

typedef struct rsort_segmented_plan_info {
    uint32_t bin_capacity[RSORT_SEGMENT_BINS]; /* keys per LDS workgroup of bin b (ascending) */
    uint32_t tile;                /* onesweep tile of the long segments */
    uint64_t max_long_segments;   /* workspace bound on segments longer than bin_capacity[8] */
    uint64_t max_tiles;
    uint32_t wide;
    uint32_t index_mode;
    uint64_t workspace_bytes;
} rsort_segmented_plan_info;

rsort_status rsort_segmented_plan(const rsort_desc* desc, uint64_t n, uint32_t offset_bytes, uint64_t num_segments,
                                  rsort_segmented_plan_info* out);

/*
 * Device-side generator and checks (enqueued on `stream`; no host sync).
 * Reference counterparts: init_keys / mix32 / set_seed_radix
 * (include/rsort/detail/utils.cuh:227-317, 435-437), check_sorted (:341-365),
 * check_pairings (:377-432), and the per-pass digit histogram
 * (include/rsort/detail/histogram.cuh:88-211).  The checks ADD the number of
 * violations to *d_violations (a device uint64 the caller zeroes), so a sort
 * of any n is checked with one device word read back.
 *
 *   rsort_init_keys      key bit patterns of the reference's generator:
 *                        mode 1 random (mix32 of index ^ seed), 2 byte_skip
 *                        (every other byte 0x55 / 0x00), 3 ascending index
 *   rsort_set_seed_radix the reference's seed of its i-th test array
 *   rsort_check_sorted   adjacent pairs out of order (float / double compare
 *                        as floating point, so NaNs never count; other kinds
 *                        on their order-preserving word)
 *   rsort_check_pairings values = original index (4 or 8 bytes): positions
 *                        whose value is out of range, whose key differs from
 *                        keys_orig[value], or whose equal-key successor has a
 *                        smaller value (stability)
 *   rsort_digit_counts   ADDS the per-pass digit counts of the (twiddled, per
 *                        desc) keys to d_counts[p*256 + d], key_bytes * 256
 *                        device uint64; equal before and after a sort
 */
rsort_status rsort_init_keys(uint32_t key_bytes, uint32_t mode, uint32_t seed, void* keys, uint64_t n, void* stream);
uint32_t rsort_set_seed_radix(int32_t index);
rsort_status rsort_check_sorted(const rsort_desc* desc, const void* keys, uint64_t n, unsigned long long* d_violations,
                                void* stream);
rsort_status rsort_check_pairings(uint32_t key_bytes, uint32_t value_bytes, const void* keys_orig,
                                  const void* keys_sorted, const void* vals_sorted, uint64_t n,
                                  unsigned long long* d_violations, void* stream);
rsort_status rsort_digit_counts(const rsort_desc* desc, const void* keys, uint64_t n, unsigned long long* d_counts,
                                void* stream);

/*
 * Tuning table (what the autotuner chooses), per key width x value width
 * (value widths other than 0,1,2,4,8,16 use the 4-byte entry):
 *   tile_candidate      pass-kernel tile: 0 = the built-in geometry, 1 = four
 *                       items fewer per thread, 2 = half the threads, 3 = half
 *                       the items per thread (1-3 exist for 4- and 8-byte
 *                       keys, keys only or with values of the key's width;
 *                       elsewhere they mean 0)
 *   hist_blocks_per_cu  histogram workgroups per CU, 1..8
 *   lds_threshold       segmented sort: longest segment sorted in LDS (LDS
 *                       bins larger than it are not used), <= bin_capacity[8]
 *   small_tile_candidate, small_log2n
 *                       the tile candidate of sorts below 2^small_log2n keys
 *                       (few tiles: a smaller tile fills more CUs)
 *   tiny_tile_candidate, tiny_log2n
 *                       the same for sorts below 2^tiny_log2n keys, checked
 *                       first (a second small-size tier)
 * Defaults are built in; at first use the library reads
 * $RSORT_AMD_TUNING_FILE, else tuning_gfx950.txt next to the library (lines
 * "key_bytes value_bytes tile_candidate hist_blocks_per_cu lds_threshold
 * [small_tile_candidate small_log2n [tiny_tile_candidate tiny_log2n]]").
 * Changing an entry changes workspace sizes: query again after a change.
 * The reference fixes these per build in a table for sm_86
 * (include/rsort/detail/tuning.cuh:138-229); there is no reference tuner.
 */
typedef struct rsort_tuning {
    uint32_t tile_candidate;       /* sorts above both small-size tiers */
    uint32_t hist_blocks_per_cu;
    uint32_t lds_threshold;
    uint32_t small_tile_candidate; /* sorts of n < 2^small_log2n keys */
    uint32_t small_log2n;          /* 0: no small-size tier */
    uint32_t tiny_tile_candidate;  /* sorts of n < 2^tiny_log2n keys (checked first) */
    uint32_t tiny_log2n;           /* 0: no tiny-size tier */
    uint32_t reserved;             /* must be 0 */
} rsort_tuning;

rsort_status rsort_tuning_get(uint32_t key_bytes, uint32_t value_bytes, rsort_tuning* out);
rsort_status rsort_tuning_set(uint32_t key_bytes, uint32_t value_bytes, const rsort_tuning* t);
/* Resets every entry to the built-in default, then reads `path`; NULL reads the
 * persisted table found at first use, "" keeps the built-in defaults only.
 * Returns the number of entries read, -1 if the file cannot be read. */
int rsort_tuning_load(const char* path);

/*
 * Serving loop: continuous batching of many small sort requests into
 * segmented sorts.  No counterpart in the reference (one sort per call); it
 * is for callers that sort many independent arrays as they arrive.
 *
 * A request is a device array of n keys (+ values) of the configured type,
 * sorted in place like rsort_sort.  A scheduler thread closes a batch when a
 * slot is full (batch_keys / batch_requests), when its oldest request has
 * waited window_us (0: as soon as a slot is free), or on flush.  Direct form
 * (default, power-of-two value widths): requests are binned by size on the
 * host and each LDS bin with requests is one launch sorting every request in
 * place; up to four longer requests get their own rsort_sort, more are sorted
 * together by one segmented sort in the slot's buffers.  Batched form
 * (RSORT_SERVE_BATCHED): gather -> segmented sort -> scatter through the
 * slot's buffers, captured once into a hipGraph unless RSORT_SERVE_NO_GRAPH.
 * Admission: a request must fit one slot (UNSUPPORTED otherwise) and the
 * queue holds at most max_pending requests (BUSY).  All device memory (depth
 * slots) is allocated at create.  The request's buffers must stay valid
 * until rsort_serve_wait returns.
 */
#define RSORT_SERVE_NO_GRAPH 1u
#define RSORT_SERVE_BATCHED 2u

typedef struct rsort_serve_config {
    rsort_desc desc;          /* element type and order of every request */
    uint64_t batch_keys;      /* keys per batch slot */
    uint32_t batch_requests;  /* requests per batch slot */
    uint32_t window_us;       /* batching window */
    uint32_t depth;           /* batch slots in flight, 1..8 */
    uint32_t max_pending;     /* queued requests before submit returns BUSY (0 = unbounded) */
    uint32_t flags;           /* RSORT_SERVE_* */
    uint32_t reserved;        /* must be 0 */
} rsort_serve_config;

typedef struct rsort_server rsort_server;

typedef struct rsort_request_info {
    uint64_t latency_us; /* submit -> completion observed */
    uint64_t queue_us;   /* submit -> batch launch */
    uint64_t batch;      /* batch sequence number */
    uint64_t n;
} rsort_request_info;

typedef struct rsort_serve_stats {
    uint64_t batches, completed, completed_keys, pending, in_flight;
    uint32_t graphs;     /* 1 when batches replay captured graphs */
    uint32_t direct;     /* 1 in the direct form */
    uint64_t workspace_bytes;
} rsort_serve_stats_info;

rsort_status rsort_serve_create(const rsort_serve_config* cfg, rsort_server** out);
rsort_status rsort_serve_submit(rsort_server* s, void* keys, void* values, uint64_t n, uint64_t* request_id);
/* Submits `count` requests at once (all or none; ids in request_ids). */
rsort_status rsort_serve_submit_many(rsort_server* s, void* const* keys, void* const* values, const uint64_t* n,
                                     uint64_t count, uint64_t* request_ids);
/* Blocks until the request is done (BUSY after timeout_us; 0 = poll;
 * RSORT_SERVE_WAIT_FOREVER or more = no limit) and releases its record. */
#define RSORT_SERVE_WAIT_FOREVER (1ull << 62)
rsort_status rsort_serve_wait(rsort_server* s, uint64_t request_id, uint64_t timeout_us, rsort_request_info* info);
/* Waits for `count` distinct requests under one time limit (as
 * rsort_serve_wait) and then releases them all, filling infos[i] when infos is
 * not NULL; on BUSY or an error none is released.  No reference counterpart:
 * the batch counterpart of rsort_serve_wait, as rsort_serve_submit_many is of
 * rsort_serve_submit. */
rsort_status rsort_serve_wait_many(rsort_server* s, const uint64_t* request_ids, uint64_t count, uint64_t timeout_us,
                                   rsort_request_info* infos);
/* Closes the open batch now and waits until nothing is queued or in flight. */
rsort_status rsort_serve_flush(rsort_server* s);
rsort_status rsort_serve_stats(rsort_server* s, rsort_serve_stats_info* out);
rsort_status rsort_serve_destroy(rsort_server* s);

/*
 * Debug builds compiled with -DRSA_PHASE_TIMING accumulate per-phase wall-clock
 * ticks (100 MHz) of the kernels; this copies the summed slots to `out` and
 * returns how many were written (0 in normal builds).  Slots of the pass
 * kernel: 0 ticket, 1 load, 2 rank, 3 scan + aggregate publish, 4 lookback,
 * 5 stage, 6 key scatter, 7 value stage + scatter, 8 lookback word loads,
 * 9 lookback not-ready polls, 10 total, 11 tiles; histogram: 16 count loop,
 * 17 reduction + global adds, 18 finalize, 20 total, 21 workgroups.
 * No reference counterpart (the reference tunes from SASS listings).
 */
int rsort_debug_phase_slots(unsigned long long* out, int max_slots, int reset);

/* Human-readable text of a status, and the last error message of the calling thread. */
const char* rsort_status_string(rsort_status s);
const char* rsort_last_error(void);
int rsort_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RSORT_C_H_ */
