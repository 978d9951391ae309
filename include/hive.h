/*
 * hive.h -- C-ABI of libhive.so, the MI355X (gfx950) Hive hash table.
 *
 * Plain C types only: device pointers are passed as typed pointers, HIP
 * streams as void* (a hipStream_t; NULL = the default stream), sizes as
 * size_t.  Every entry point returns hive_status_t and never exits the
 * process.  Citations are to the reference (SabbirPulok/HiveHashTable)
 * interfaces each entry point replaces.
 *
 * Table layouts (device memory owned by the handle), chosen per table:
 * HIVE_LAYOUT_BUCKET8 (the reference's bucket) and HIVE_LAYOUT_BUCKET16:
 *   slots[nb * S]   u64 packed kv = (uint64)value << 32 | key, 0 = empty;
 *                   S = 8 slots (one dense 64-byte bucket) or 16 (a bucket
 *                   owns its 128-byte line)
 *   stash[nb * 8]   u64, one 8-slot overflow bucket per canonical bucket
 *   stash_count[nb] u32 monotonic push pointer of each stash bucket
 *   meta[nb/16]     u32, 2 hint bits per bucket: overflow (a key whose H0 is
 *                   this bucket may sit in another candidate bucket) and
 *                   stash (stash_count > 0)
 *   locks[nb]       u32 per-canonical-bucket lock of the insert slow path
 * HIVE_LAYOUT_INLINE8: 8 slots per bucket, each bucket in its own 128-byte
 *   line with an in-line header after the slots: the hint word (31 overflow
 *   classes and the stash bit), the lock, the stash count and the index of
 *   the bucket's stash in a pool of nb / 4 + 64 8-slot stash buckets taken
 *   on first push (an insert that needs a stash bucket when the pool is
 *   exhausted fails).  Same protocol and results as HIVE_LAYOUT_BUCKET8.
 * Key 0 is the empty sentinel and is rejected by every operation
 * (include/utils.h:40-47 of the reference).
 */
#ifndef HIVE_H
#define HIVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  HIVE_OK = 0,
  HIVE_EINVAL = 1,     /* bad argument, unknown policy, bad size */
  HIVE_ENOMEM = 2,     /* device or host allocation failed */
  HIVE_EDEVICE = 3,    /* HIP runtime error, or no usable device */
  HIVE_ECAPACITY = 4,  /* resize target cannot hold the live entries */
  HIVE_ERANGE = 5      /* output buffer too small */
} hive_status_t;

/* Table layouts (the table layout comment above). */
typedef enum {
  HIVE_LAYOUT_BUCKET8 = 0,   /* 8-slot 64-byte buckets, separate hint / lock / stash arrays (the reference's bucket) */
  HIVE_LAYOUT_BUCKET16 = 1,  /* 16-slot 128-byte buckets, separate arrays */
  HIVE_LAYOUT_INLINE8 = 2    /* 8 slots + in-line header per 128-byte line, pooled stash */
} hive_layout_t;

/* Hash policies: include/HashPolicies.cuh:10-65.  The string names accepted
 * by hive_policy_from_name are the reference's (src/hive_dispatch.cu:829-864). */
typedef enum {
  HIVE_POLICY_DEFAULT2HASH = 0,   /* hash1, hash2 */
  HIVE_POLICY_MURMURCITY = 1,     /* murmur, city */
  HIVE_POLICY_TRIPLE = 2,         /* hash1, hash2, city */
  HIVE_POLICY_MURMURCITYBIT = 3,  /* murmur, city, hash1 */
  HIVE_POLICY_LOOKUP2 = 4,        /* crc64, crc32 */
  HIVE_POLICY_LOOKUP3 = 5         /* crc64, crc32, hash1 */
} hive_policy_t;

/* Operation types, numbered as the reference's OperationType (include/utils.h:10-16). */
enum { HIVE_OP_INSERT = 0, HIVE_OP_LOOKUP = 1, HIVE_OP_DELETE = 2, HIVE_OP_ATOMIC_INC = 3, HIVE_OP_NONE = 4 };

/* Per-operation result codes (uint8_t per op). */
enum {
  HIVE_R_FAILED = 0,      /* insert: canonical bucket, cuckoo and stash all full */
  HIVE_R_INSERTED = 1,    /* insert: new key placed in a bucket */
  HIVE_R_UPDATED = 2,     /* insert: key existed, value replaced */
  HIVE_R_STASHED = 3,     /* insert: new key placed in the overflow stash */
  HIVE_R_FOUND = 4,       /* lookup hit */
  HIVE_R_NOT_FOUND = 5,   /* lookup miss, or delete of an absent key */
  HIVE_R_ERASED = 6,      /* delete removed the key */
  HIVE_R_INVALID_KEY = 7, /* key 0 */
  HIVE_R_NOOP = 8         /* NONE / ATOMIC_INC: no effect, result 0 (include/hive_kernels.cuh:181-184) */
};

/* The reference's 16-byte Operation (include/utils.h:18-21): int type, 4 pad bytes, uint64 key. */
typedef struct {
  int32_t type;
  uint32_t pad;
  uint64_t key;
} hive_operation_t;

typedef struct hive_table hive_table_t;

typedef struct {
  uint64_t num_buckets;
  uint64_t slots_per_bucket; /* 8 or 16 */
  uint64_t capacity;         /* num_buckets * slots_per_bucket */
  uint64_t table_live;       /* entries stored in buckets */
  uint64_t stash_live;       /* entries stored in the stash */
  double load_factor;        /* table_live / capacity, as count_occupancy_kernel (include/hive_kernels.cuh:303-325) */
  uint64_t device_bytes;     /* device memory held by the handle */
  int policy;
  uint64_t slow_path_ops;    /* inserts that took the locked slow path since the last clear/import */
  int layout;                /* HIVE_LAYOUT_* */
} hive_stats_t;

/* Launch shape of the plain bulk kernels.  0 = library default / autotuned. */
typedef struct {
  uint32_t block_threads;    /* threads per block: 64..1024, multiple of 64 */
  uint32_t keys_per_lane;    /* lookups: 0 or 1 = one key per lane; 2 or 4 = each
                                lane probes that many keys at once, issuing a
                                round's bucket reads for all of them together
                                (same results); other ops: 1 */
  uint32_t inline_slow_path; /* inserts: 0 = stage slow-path ops in LDS behind one
                                block barrier (throughput), 1 = each lane runs its
                                own slow path at once (latency), 2 = two launches:
                                every fast path, then the slow paths from a device
                                list (hive_insert allocates the list, 12 bytes per
                                op, outside stream capture; other entry points and
                                batches larger than the list run as 0) */
  uint32_t combine_duplicates; /* inserts: 1 = ops of one block on the same key are
                                  combined into one table update before reaching
                                  memory (skewed batches; slow path inline) */
  uint32_t split_by_type;      /* mixed batches (read from the insert entry): 1 = in
                                  batches of >= 2^18 ops, inserts and deletes are
                                  compacted into dense lists before their phase
                                  kernels; 0 = every phase filters the whole batch */
  uint32_t lanes_per_key;      /* lookups: 0 or 1 = one lane per key reads whole
                                  buckets; 4 or 8 = that many lanes of a wavefront
                                  share a key, each reading part of every probed
                                  bucket (same results) */
} hive_launch_t;

/* ---- lifecycle: replaces hiveHashTableCreate / Destroy / Clear
 *      (include/hive_hash_table.cuh:92-144, src/hive_memory.cu:29-90,174-218).
 * num_slots is the reference's table_size: num_buckets = num_slots / 8
 * (src/hive_dispatch.cu:630-631).  The handle keeps the table resident on the
 * current device until hive_destroy. */
hive_status_t hive_create(uint64_t num_slots, int policy, hive_table_t** out);
/* hive_create with a layout (HIVE_LAYOUT_*; hive_create makes
 * HIVE_LAYOUT_BUCKET8).  num_buckets = num_slots / slots per bucket.  The
 * protocol is include/hive_hash_insert.cuh:898-1042 with SLOTS = the
 * layout's slots per bucket (the cuckoo stage tries 8 victims, slots 0..7)
 * and an 8-slot stash per bucket. */
hive_status_t hive_create_ex(uint64_t num_slots, int policy, int layout, hive_table_t** out);
hive_status_t hive_destroy(hive_table_t* t);
hive_status_t hive_clear(hive_table_t* t, void* stream);
hive_status_t hive_policy_from_name(const char* name, int* policy);

/* ---- bulk operations on device buffers, stream-ordered and asynchronous.
 * vals may be NULL for hive_insert (value = key + 1, the reference's
 * convention, include/hive_kernels.cuh:156,238).  Output pointers may be NULL.
 * hive_insert  replaces hive_insert_kernel (include/hive_kernels.cuh:205-266)
 * hive_lookup  replaces hive_lookup_kernel (include/hive_kernels.cuh:76-117),
 *              returning the value as well as the found flag
 * hive_erase   replaces the DELETE arm of hive_mixed_kernel (include/hive_kernels.cuh:175-180) */
hive_status_t hive_insert(hive_table_t* t, const uint32_t* keys, const uint32_t* vals, size_t n,
                          uint8_t* status, void* stream);
hive_status_t hive_lookup(hive_table_t* t, const uint32_t* keys, size_t n, uint32_t* vals_out,
                          uint8_t* status, void* stream);
hive_status_t hive_erase(hive_table_t* t, const uint32_t* keys, size_t n, uint8_t* status, void* stream);

/* Mixed batch: types[i] in HIVE_OP_*.  All operations of one call are
 * concurrent; the library executes them as lookups, then deletes, then
 * inserts (replaces hive_mixed_kernel, include/hive_kernels.cuh:119-192).
 * vals may be NULL (value = key + 1). out_vals receives lookup values. */
hive_status_t hive_mixed(hive_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                         size_t n, uint32_t* out_vals, uint8_t* status, void* stream);

/* hive_mixed without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory during the call -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments, semantics and results as hive_mixed; the type lists of a
 * batch of >= 2^18 ops (split_by_type) live in a static device array of the
 * library (36 MiB per GPU, held until the process ends) instead of a buffer
 * the handle allocates on first use.  At most 2^20 ops per call (HIVE_EINVAL
 * beyond).  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (other streams, other threads, other handles) race, so callers
 * order them. */
hive_status_t hive_mixed_noalloc(hive_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                                 size_t n, uint32_t* out_vals, uint8_t* status, void* stream);

/* The reference's kernel interface over its 16-byte Operation records with
 * uint64 0/1 results and value = key + 1:
 *   hive_insert_ops -> hive_insert_kernel (include/hive_kernels.cuh:205-266)
 *   hive_lookup_ops -> hive_lookup_kernel (include/hive_kernels.cuh:76-117)
 *   hive_mixed_ops  -> hive_mixed_kernel  (include/hive_kernels.cuh:119-192) */
hive_status_t hive_insert_ops(hive_table_t* t, const hive_operation_t* ops, size_t n, uint64_t* results, void* stream);
hive_status_t hive_lookup_ops(hive_table_t* t, const hive_operation_t* ops, size_t n, uint64_t* results, void* stream);
hive_status_t hive_mixed_ops(hive_table_t* t, const hive_operation_t* ops, size_t n, uint64_t* results, void* stream);

/* ---- streams of mixed batches (no reference counterpart: its harness
 * launches and synchronizes one hive_mixed_kernel per batch,
 * src/hive_dispatch.cu:119-164).  Batch b is ops [h_offsets[b],
 * h_offsets[b+1]) of the concatenated device arrays; every batch keeps the
 * hive_mixed order (lookups, deletes, inserts) and batch b+1 runs after batch
 * b, so all modes return the same results.
 *   HIVE_STREAM_LAUNCHES   three kernels per batch queued on the stream
 *   HIVE_STREAM_GRAPH      the same launches captured once into a hipGraph
 *                          and replayed (capture + instantiate in *setup_ms)
 *   HIVE_STREAM_PERSISTENT one resident kernel that claims tiles of
 *                          (batch, phase) from a device work queue
 *   HIVE_STREAM_WINDOWED   consecutive batches of up to window_ops ops in
 *                          all run as one window: each key's ops up to its
 *                          first written batch in the window run as one mixed
 *                          batch, its later ops in batch order after it (see
 *                          hive_set_stream_window); other batches run alone
 *                          as in LAUNCHES
 * h_offsets: host array of nbatches + 1 offsets.  batch_done_ns (host, may be
 * NULL) receives each batch's completion time in ns after the stream start
 * (device wall clock); in WINDOWED mode, every batch of a window completes
 * with the window.  Synchronizes the stream. */
enum { HIVE_STREAM_LAUNCHES = 0, HIVE_STREAM_GRAPH = 1, HIVE_STREAM_PERSISTENT = 2, HIVE_STREAM_WINDOWED = 3 };
hive_status_t hive_mixed_stream(hive_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                                const uint64_t* h_offsets, size_t nbatches, uint32_t* out_vals, uint8_t* status,
                                int mode, double* batch_done_ns, double* setup_ms, void* stream);
/* Shape of the persistent kernel: ops per tile (multiple of 256, 0 = about
 * one tile per resident block, 256..4096), resident blocks per CU (0 = 4),
 * and whether insert slow paths run inline (1) or staged per tile (0). */
hive_status_t hive_set_stream_shape(hive_table_t* t, uint32_t tile_ops, uint32_t blocks_per_cu,
                                    uint32_t inline_slow_path);
/* Windows of HIVE_STREAM_WINDOWED: runs of consecutive batches of fewer
 * than batch_ops ops each (0 = 2^19, at most 2^24) are coalesced into windows
 * of at most window_ops ops (2^10..2^24, 0 = 2^21); every other batch runs
 * alone.  The handle keeps 322-644 bytes of device scratch per window op
 * (161 bytes per slot of a power-of-two set of at least 2 slots per op) once
 * a windowed stream has run. */
hive_status_t hive_set_stream_window(hive_table_t* t, uint64_t window_ops, uint64_t batch_ops);

/* Autotuner for the stream forms (no reference counterpart): times the
 * sample stream on a scratch copy of the table -- so at the table's current
 * load factor, leaving the table unchanged -- for per-batch launches at
 * 128..1024 threads per block x {staged, inline} insert slow path (x split
 * by type or not, when the sample has batches of >= 2^18 ops), the
 * captured graph, windows of {2^15 .. 2^22} ops coalescing batches below
 * 1/32 .. 1/4 of the window (at most 2^19 ops) x {split by type or not, for
 * windows of >= 2^18 ops}, and, on
 * request (hive_autotune_stream_ex), the persistent kernel over tile sizes
 * {auto, 256, 1024, 4096} x {2, 4, 8} blocks per CU x {staged, inline}.
 * Every candidate's p99 batch service time is measured (a batch's time from
 * the previous completion to its own; the batches of a window complete
 * together).  The chosen form is the fastest whose p99 stays within the
 * latency bound -- by default the p99 of the per-batch launches plus 10 %, so
 * the choice does not trade latency for throughput -- and its launch shapes and
 * window are stored in the handle.  The fastest window regardless of latency
 * is reported too.  Synchronizes. */
typedef struct {
  uint32_t block_threads;  /* chosen threads per block of the plain kernels */
  uint32_t inline_slow_path; /* chosen insert slow-path placement of the plain kernels */
  uint32_t split_by_type;  /* chosen: big batches' inserts and deletes compacted (1) or not (0) */
  uint32_t tile_ops;       /* chosen persistent tile (0 = sized per batch) */
  uint32_t blocks_per_cu;  /* chosen persistent residency */
  uint32_t stream_inline;  /* chosen insert slow-path placement of the persistent kernel */
  int best_mode;           /* HIVE_STREAM_* fastest on the sample */
  double launches_ms;      /* best sample time per mode (device clock) */
  double graph_ms;
  double persistent_ms;
  uint32_t candidates;     /* configurations timed */
  uint64_t window_ops;     /* fastest window (regardless of latency): its size */
  uint64_t window_batch_ops; /* ... batches below this size are coalesced */
  double windowed_ms;      /* ... and its sample time */
  double windowed_p99_us;  /* ... and its p99 batch service time */
  double launches_p99_us;  /* p99 batch service time of the chosen per-batch launches */
  double graph_p99_us;
  double persistent_p99_us; /* (persistent fields: 0 when not tried) */
  double p99_bound_us;     /* the latency bound of the choice (0 = none) */
  uint64_t bounded_window_ops; /* fastest window within the bound (0 = none): its size */
  uint64_t bounded_window_batch_ops;
  double bounded_windowed_ms;
  double bounded_windowed_p99_us;
  uint32_t bounded_split_by_type;
  double windowed_ms_by_size[8];     /* per window size 2^15 .. 2^22: the best sample time */
  double windowed_p99_us_by_size[8]; /* ... and that candidate's p99 batch service time */
} hive_tune_t;
typedef struct {
  uint32_t try_persistent; /* 1 = also time the persistent kernel (0: it never won on MI355X) */
  double max_p99_us;       /* latency bound of the choice: < 0 = the per-batch launches' p99
                              times (1 + p99_slack), 0 = none (fastest form), > 0 = this many
                              microseconds */
  double p99_slack;        /* with max_p99_us < 0: tolerance over the launches' p99, >= 0 (the
                              p99 of a 100-batch sample is about its second-longest batch) */
} hive_tune_opts_t;
hive_status_t hive_autotune_stream(hive_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                                   const uint64_t* h_offsets, size_t nbatches, hive_tune_t* out, void* stream);
/* hive_autotune_stream with options (NULL = {0, -1, 0.1}: no persistent
 * kernel, bounded by the launches' p99 plus 10 %). */
hive_status_t hive_autotune_stream_ex(hive_table_t* t, const uint8_t* types, const uint32_t* keys,
                                      const uint32_t* vals, const uint64_t* h_offsets, size_t nbatches,
                                      const hive_tune_opts_t* opts, hive_tune_t* out, void* stream);

/* Autotuner for the bulk kernels (no reference counterpart), at the table's
 * current load factor: times the insert of the sample keys (device arrays,
 * vals may be NULL) on a scratch copy of the table at 256/512/1024 threads
 * per block x {staged, inline, second launch} slow path, and the lookup of the
 * sample lookup
 * keys on the table itself at 256/512/1024 threads x {1, 4, 8} lanes per key
 * and x {2, 4} keys per lane, then stores the fastest insert and lookup shapes
 * in the handle.  The table
 * is left unchanged.  Either sample may be empty (n = 0: that op is not
 * tuned).  Synchronizes. */
typedef struct {
  uint32_t insert_block_threads;
  uint32_t insert_inline_slow_path;
  uint32_t lookup_block_threads;
  uint32_t lookup_lanes_per_key;
  double insert_ms;               /* best sample times (device clock) */
  double lookup_ms;
  double lookup_ms_by_lanes[3];   /* best lookup time with 1, 4 and 8 lanes per key */
  double load_factor;             /* of the table while tuned */
  uint32_t candidates;
  uint32_t lookup_keys_per_lane;  /* chosen keys per lane (1 with lanes_per_key > 1) */
  double lookup_ms_by_keys[3];    /* best lookup time with 1, 2 and 4 keys per lane (one lane per key) */
  double insert_ms_by_slow_path[3]; /* best insert time with the slow path staged, inline, in a second
                                       launch (0 = not timed) */
} hive_bulk_tune_t;
hive_status_t hive_autotune_bulk(hive_table_t* t, const uint32_t* insert_keys, const uint32_t* insert_vals,
                                 size_t n_insert, const uint32_t* lookup_keys, size_t n_lookup, hive_bulk_tune_t* out,
                                 void* stream);

/* Bucket layout chosen by measurement (no reference counterpart: the
 * reference has one layout).  For each layout (HIVE_LAYOUT_*) the table's
 * content is rehashed into that layout, the insert sample is inserted into a
 * copy of it, and the lookup sample is looked up in the result (the faster
 * of two runs each); the layout with the least insert + lookup time is then
 * applied to the table, its content kept.  Times are indexed by layout id;
 * -1 marks a layout the content did not fit.  Synchronizes. */
typedef struct {
  int layout;            /* chosen and applied */
  double insert_ms[3];   /* by layout id (device clock) */
  double lookup_ms[3];
  uint32_t candidates;
} hive_layout_tune_t;
hive_status_t hive_autotune_layout(hive_table_t* t, const uint32_t* insert_keys, const uint32_t* insert_vals,
                                   size_t n_insert, const uint32_t* lookup_keys, size_t n_lookup,
                                   hive_layout_tune_t* out, void* stream);

/* ---- serving: continuous batching over one resident table
 * Replaces the reference's YCSB driver (src/hashTableBenchmark.cc:434-570,
 * src/hive_dispatch.cu:317-549), which runs a pre-built workload as one
 * batch per call.  hive_serve_trace replays a trace of client requests in
 * real time: request r owns ops [h_req_offsets[r], h_req_offsets[r+1]) of
 * the device arrays and arrives h_arrival_us[r] microseconds after the call
 * starts (non-decreasing).  Whenever fewer than max_inflight batches are
 * queued on the stream, every request that has arrived and is not yet
 * launched (up to max_batch_ops) is coalesced into one mixed batch with
 * hive_mixed semantics (lookups, then deletes, then inserts).  Requests
 * whose keys do not meet other requests of the same batch get exactly the
 * results they would get alone; ops of different requests on one key in one
 * batch follow the batch's phase order.  With ordered = 1, a batch of several
 * requests runs as a window in which each request is a batch of its own
 * (HIVE_STREAM_WINDOWED), so every request gets the results it would get
 * alone, run after the requests that arrived before it (at most 4096
 * requests are coalesced into one batch; the call fails with HIVE_EINVAL
 * when max_batch_ops or a request exceeds 2^25 ops in a trace of more than
 * 2^25 ops); a batch whose requests
 * neither insert nor delete runs as a plain batch, which gives the same
 * results.
 * Admission: before a batch whose inserts could lift the load factor above
 * max_load_factor, the table is resized in place by `growth` if free HBM
 * keeps hbm_reserve_bytes after the new table; if not, requests that insert
 * are admitted while stash use stays below stash_headroom of the stash
 * capacity, and rejected otherwise (their ops get HIVE_R_FAILED, unexecuted).
 * Shrink (min_load_factor > 0; off by default, as the reference's
 * DYNAMIC_RESIZE, src/hive_dispatch.cu:189-206): once the deletes launched
 * since the last exact count could have taken the load factor ((table +
 * stash entries) / capacity) below min_load_factor, and after the last
 * batch, the queue is drained and the exact count read; below
 * min_load_factor the table is rehashed in place to ceil(entries /
 * shrink_load_factor / slots per bucket) buckets (the reference's target of
 * 75 % occupancy with the default 0.75).
 * h_latency_us[r] (host, may be NULL) receives arrival-to-completion time
 * (completion = the host observing the batch's event).  Synchronizes. */
typedef struct {
  uint64_t max_batch_ops;     /* ops coalesced into one batch at most (0 = 2^22) */
  uint32_t max_inflight;      /* batches queued on the stream at once, 1..8 (0 = 2) */
  double max_load_factor;     /* resize before a batch that could exceed it (0 = 0.9) */
  double growth;              /* capacity multiplier of one resize, > 1 (0 = 2.0) */
  uint64_t hbm_reserve_bytes; /* free HBM a resize must leave (0 = 1 GiB) */
  double stash_headroom;      /* admit inserts without a resize while stash use < this (0 = 0.5) */
  uint32_t ordered;           /* 1 = each request sees every earlier request's writes, also inside a batch */
  double min_load_factor;     /* shrink below it (0 = never shrink; the reference's trigger is 0.25) */
  double shrink_load_factor;  /* load factor right after a shrink, in (min_load_factor, max_load_factor] (0 = 0.75) */
} hive_serve_config_t;
typedef struct {
  uint64_t requests;
  uint64_t batches;
  uint64_t ops;               /* ops executed (rejected requests excluded) */
  uint64_t rejected_requests;
  uint64_t resizes;
  uint64_t final_slots;
  double span_ms;             /* first arrival to last completion */
  double mops;                /* ops / span */
  double p50_us, p99_us, max_us; /* request latency */
  double mean_batch_ops;
  double resize_ms;           /* time spent growing (inside span) */
  uint64_t shrinks;           /* shrinks (resizes counts growths) */
  double shrink_ms;           /* time spent shrinking, drain and count included */
} hive_serve_report_t;
hive_status_t hive_serve_trace(hive_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                               const uint64_t* h_req_offsets, const double* h_arrival_us, size_t nreq,
                               uint32_t* out_vals, uint8_t* status, const hive_serve_config_t* cfg,
                               double* h_latency_us, hive_serve_report_t* report, void* stream);

/* ---- insert sequences with in-place growth (cfg3 shape).
 * Inserts keys [0, n) in batches of batch_keys; after batch b, if
 * lookups_per_batch > 0, looks up lookup_keys[b*lookups_per_batch ...]
 * (device, nbatches * lookups_per_batch keys; lookup_status may be NULL).
 * The table grows in place (x growth, repeated until it fits) before any
 * batch that could lift (table + stash entries) / capacity above
 * max_load_factor.
 *   HIVE_SEQUENCE_PLAIN: per batch, the host reads the exact count
 *     (hive_stats) and calls hive_resize -- the reference's DYNAMIC_RESIZE
 *     loop (src/hive_dispatch.cu:181-211) over this library.
 *   HIVE_SEQUENCE_GRAPH: the resize points are fixed up front from an upper
 *     bound of the count (every insert adds an entry), the target tables are
 *     allocated and cleared before capture, and the whole sequence -- rehash
 *     kernels included -- is captured once into a hipGraph and launched.
 * The report gives the wall time of the sequence (setup_ms: graph allocation,
 * capture and instantiation, outside elapsed_ms).  Synchronizes. */
enum { HIVE_SEQUENCE_PLAIN = 0, HIVE_SEQUENCE_GRAPH = 1 };
typedef struct {
  double elapsed_ms;
  double setup_ms;
  uint32_t resizes;
  uint64_t final_slots;
  uint64_t rehash_failed;   /* entries a rehash could not place (then HIVE_ECAPACITY) */
} hive_sequence_report_t;
hive_status_t hive_insert_sequence(hive_table_t* t, const uint32_t* keys, const uint32_t* vals, size_t n,
                                   size_t batch_keys, const uint32_t* lookup_keys, size_t lookups_per_batch,
                                   double max_load_factor, double growth, int mode, uint8_t* status,
                                   uint8_t* lookup_status, hive_sequence_report_t* report, void* stream);

/* ---- maintenance
 * hive_resize: rehash every live entry (buckets and stash) into
 *   new_num_slots / 8 buckets; replaces hive_hash_grow / hive_hash_shrink
 *   (include/hive_hash_resize.cuh:138-215,340-420) with a rehash that keeps
 *   lookups consistent.  Returns HIVE_ECAPACITY (table unchanged) if an entry
 *   could not be placed.  Synchronizes the stream.
 * hive_stash_drain: flatten live stash entries into out_kv (device, packed
 *   kv), reset the stash; replaces compact_stash_kernel +
 *   reinsert_stash_into_next_batch (include/hive_stash_table.cuh:33-155).
 *   *n_out (host) receives the entry count.  Synchronizes the stream.
 * hive_stash_reinsert: drain the stash and insert the drained entries back
 *   (with their values) through the normal insert path.  Synchronizes. */
hive_status_t hive_resize(hive_table_t* t, uint64_t new_num_slots, void* stream);
/* hive_set_layout: rehash the table into another layout (HIVE_LAYOUT_*) at
 * the same slot capacity, as hive_resize does (no reference counterpart: the
 * reference's bucket is fixed at build time, HIVE_BUCKET_SLOTS).
 * HIVE_ECAPACITY (table unchanged) if an entry could not be placed.
 * Synchronizes. */
hive_status_t hive_set_layout(hive_table_t* t, int layout, void* stream);
/* hive_dump: every live entry (bucket slots, then stash) as packed kv into
 *   out_kv (device, cap entries), in no particular order -- the content
 *   multiset, for checks and persistence (the reference has no export; its
 *   tables live for one dispatch call).  *n_out (host) receives the count;
 *   HIVE_ERANGE (nothing written) if it exceeds cap.  Synchronizes. */
hive_status_t hive_dump(hive_table_t* t, uint64_t* out_kv, size_t cap, size_t* n_out, void* stream);
hive_status_t hive_stash_drain(hive_table_t* t, uint64_t* out_kv, size_t cap, size_t* n_out, void* stream);
hive_status_t hive_stash_reinsert(hive_table_t* t, size_t* n_reinserted, void* stream);

/* ---- observation (synchronize the stream)
 * hive_stats replaces count_occupancy_kernel + the LF computation
 *   (include/hive_kernels.cuh:303-325, src/hive_dispatch.cu:166-177).
 * hive_count_exact scans the table instead of reading the counters.
 * hive_export_raw / hive_import_raw copy the raw layout to / from host
 *   memory (arrays sized num_buckets*slots_per_bucket, num_buckets*8,
 *   num_buckets: the separate layouts' slots, stash and stash counts; an
 *   in-line table is converted to and from them). */
hive_status_t hive_stats(hive_table_t* t, hive_stats_t* out, void* stream);
hive_status_t hive_count_exact(hive_table_t* t, uint64_t* table_live, uint64_t* stash_live, void* stream);
/* Probe statistics for a key set against the current table (no reference
 * counterpart; used for the algorithmic-bytes accounting): the distinct
 * candidate buckets an insert reads, the buckets a lookup reads before its
 * early exit, and the lookups that also read the stash. */
hive_status_t hive_probe_stats(hive_table_t* t, const uint32_t* keys, size_t n, uint64_t* candidate_buckets,
                               uint64_t* lookup_bucket_reads, uint64_t* lookup_stash_reads, void* stream);
hive_status_t hive_export_raw(hive_table_t* t, uint64_t* h_slots, uint64_t* h_stash, uint32_t* h_stash_count,
                              void* stream);
hive_status_t hive_import_raw(hive_table_t* t, const uint64_t* h_slots, const uint64_t* h_stash,
                              const uint32_t* h_stash_count, void* stream);
uint64_t hive_num_buckets(const hive_table_t* t);

/* ---- launch shape of the plain bulk kernels (autotuner output) */
hive_status_t hive_set_launch(hive_table_t* t, int op_type, const hive_launch_t* cfg);
hive_status_t hive_get_launch(const hive_table_t* t, int op_type, hive_launch_t* cfg);

/* ---- drop-in pair with the reference's host dispatch semantics.
 * Same arguments as hash_table_kernel_dispatch / hash_table_kernel_dispatch_YCSB
 * (include/hive_hash_table.cuh:50-77, src/hive_dispatch.cu:317-549,815-872),
 * with double& -> double*, std::string -> const char*, bool -> int,
 * HashTableDataLayout -> int.  Host buffers in and out; a table is created
 * and destroyed per call as in the reference.  Unknown policy names return
 * HIVE_EINVAL (the reference prints and falls back to Default2Hash).
 * elapsed_ms receives the mean kernel time per timed iteration. */
hive_status_t hive_hash_table_kernel_dispatch(const hive_operation_t* h_ops, size_t num_ops, size_t table_size,
                                              size_t threads_per_block, size_t num_blocks, size_t num_iterations,
                                              double* elapsed_ms, uint64_t* h_results, int verify, int layout,
                                              const char* hash_policy);
hive_status_t hive_hash_table_kernel_dispatch_ycsb(const hive_operation_t* h_prefill_ops,
                                                   const hive_operation_t* h_workload_ops, size_t num_prefill_ops,
                                                   size_t num_workload_ops, size_t table_size,
                                                   size_t threads_per_block, size_t num_iterations,
                                                   double* elapsed_ms, uint64_t* h_results, int verify, int layout);

/* ---- measurement kernels (same-footprint ceilings; no reference counterpart)
 * hive_ceiling_gather: n random 64-byte reads over a buffer of footprint
 *   bytes, one per lane (the bucket-read pattern of lookups).
 * hive_ceiling_cas: n random 64-bit CAS over the same footprint.
 * buf must hold footprint bytes of device memory; out gets one u32 per
 * block (anti-dead-code sink, may be NULL for cas). */
hive_status_t hive_ceiling_gather(const uint64_t* buf, uint64_t footprint_bytes, uint64_t n, uint32_t seed,
                                  uint32_t* sink, void* stream);
hive_status_t hive_ceiling_cas(uint64_t* buf, uint64_t footprint_bytes, uint64_t n, uint32_t seed, void* stream);
/* hive_ceiling_gather_lanes: the gather with lanes_per_bucket (1, 4, 8, 16)
 * lanes cooperating on each 64-byte bucket -- the lane-per-key against
 * wavefront-per-bucket layout measurement.  For lanes_per_bucket > 1 the
 * sink receives at most one u32 (sink[0]). */
/* hive_ceiling_gather_bytes: the gather with bucket_bytes-sized reads (64:
 *   the 8-slot bucket, 128: a 128-byte line). */
hive_status_t hive_ceiling_gather_bytes(const uint64_t* buf, uint64_t footprint_bytes, uint64_t n,
                                        uint32_t bucket_bytes, uint32_t seed, uint32_t* sink, void* stream);
hive_status_t hive_ceiling_gather_lanes(const uint64_t* buf, uint64_t footprint_bytes, uint64_t n,
                                        uint32_t lanes_per_bucket, uint32_t seed, uint32_t* sink, void* stream);

const char* hive_status_string(hive_status_t s);
const char* hive_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* HIVE_H */
