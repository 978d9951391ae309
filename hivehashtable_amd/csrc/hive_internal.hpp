// hive_internal.hpp -- host-side internals shared by the translation units of
// libhive.so: the table handle, error plumbing and the launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/hive.h"
#include "hive_kernels.hpp"

namespace hive {

extern thread_local std::string g_last_error;

inline hive_status_t fail(hive_status_t s, const std::string& what) {
  g_last_error = what;
  return s;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::hive::fail(HIVE_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define HIP_CHECK_LAUNCH() HIP_TRY(hipGetLastError())

constexpr uint32_t kDefaultBlock = 256;

}  // namespace hive

struct hive_table {
  int device = 0;
  int policy = 0;
  hive::TableView v{};
  size_t bytes = 0;
  hive_launch_t launch[3] = {{hive::kDefaultBlock, 1, 0, 0, 1, 1}, {hive::kDefaultBlock, 1, 0, 0, 1, 1},
                             {hive::kDefaultBlock, 1, 0, 0, 1, 1}};
  unsigned long long* scratch = nullptr;  // 8 u64 of device scratch
  uint32_t stream_tile = 0;              // persistent kernel: ops per tile (0 = sized per batch)
  uint32_t stream_blocks_per_cu = 0;     // persistent kernel: resident blocks per CU (0 = default)
  uint32_t stream_inline = 0;            // persistent kernel: insert slow path inline (1) or staged (0)
  uint32_t* part = nullptr;              // mixed-batch type lists: 3 x (keys, values, indices) + 3 counts
  size_t part_cap = 0;                   // ops the lists can hold
  uint64_t stream_window_ops = 0;        // windowed streams: most ops per window (0 = default)
  uint64_t stream_window_batch = 0;      // windowed streams: batches below this size are coalesced (0 = default)
  unsigned long long* win_set = nullptr; // windowed streams: key set of the current window, then chain records
  uint32_t* win_buf = nullptr;           // windowed streams: two buffers of per-op arrays, chain marks, control block
  size_t win_cap = 0;                    // key-set entries (a power of two; windows of up to win_cap / 2 ops)
  uint32_t win_epoch = 0;                // epoch of the last window (entries of other epochs are free)
  unsigned long long* pend = nullptr;    // split inserts: pending kv [cap], op index [cap] (u32), count
  size_t pend_cap = 0;                   // ops the pending list holds
};

namespace hive {

inline size_t meta_words(uint32_t nb) { return (static_cast<size_t>(nb) + kHintsPerWord - 1) / kHintsPerWord; }
inline size_t counter_words() { return static_cast<size_t>(3) * kCounterLanes * kCounterStride; }
inline unsigned grid_for(size_t n, uint32_t block) { return static_cast<unsigned>((n + block - 1) / block); }
// one launch covers at most 2^31 - 1 blocks
inline bool fits_one_launch(size_t n, uint32_t block) { return (n + block - 1) / block <= 0x7FFFFFFFull; }
inline bool valid_block(uint32_t b) { return b >= 64 && b <= static_cast<uint32_t>(kMaxBlock) && b % 64 == 0; }
inline unsigned fill_grid(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  return static_cast<unsigned>(g ? g : 1);
}

// A table's kind: its hash policy and layout (HIVE_LAYOUT_*).
// HIVE_POLICY_SWITCH runs BODY with P bound to the matching policy type.
inline int table_kind(int policy, uint32_t layout) { return policy + 8 * static_cast<int>(layout); }
inline int table_kind(const hive_table* t) {
  return table_kind(t->policy, t->v.layout == kLayoutInline ? HIVE_LAYOUT_INLINE8 : (t->v.S == 16 ? HIVE_LAYOUT_BUCKET16
                                                                                                : HIVE_LAYOUT_BUCKET8));
}

#define HIVE_POLICY_SWITCH(kind, BODY)                                                                            \
  switch (kind) {                                                                                                 \
    case 0 + HIVE_POLICY_DEFAULT2HASH: { using P = ::hive::PolicyDefault2; BODY; } break;                         \
    case 0 + HIVE_POLICY_MURMURCITY: { using P = ::hive::PolicyMurmurCity; BODY; } break;                         \
    case 0 + HIVE_POLICY_TRIPLE: { using P = ::hive::PolicyTriple; BODY; } break;                                 \
    case 0 + HIVE_POLICY_MURMURCITYBIT: { using P = ::hive::PolicyMurmurCityBit; BODY; } break;                   \
    case 0 + HIVE_POLICY_LOOKUP2: { using P = ::hive::PolicyLookup2; BODY; } break;                               \
    case 0 + HIVE_POLICY_LOOKUP3: { using P = ::hive::PolicyLookup3; BODY; } break;                               \
    case 8 + HIVE_POLICY_DEFAULT2HASH: { using P = ::hive::Wide16<::hive::PolicyDefault2>; BODY; } break;         \
    case 8 + HIVE_POLICY_MURMURCITY: { using P = ::hive::Wide16<::hive::PolicyMurmurCity>; BODY; } break;         \
    case 8 + HIVE_POLICY_TRIPLE: { using P = ::hive::Wide16<::hive::PolicyTriple>; BODY; } break;                 \
    case 8 + HIVE_POLICY_MURMURCITYBIT: { using P = ::hive::Wide16<::hive::PolicyMurmurCityBit>; BODY; } break;   \
    case 8 + HIVE_POLICY_LOOKUP2: { using P = ::hive::Wide16<::hive::PolicyLookup2>; BODY; } break;               \
    case 8 + HIVE_POLICY_LOOKUP3: { using P = ::hive::Wide16<::hive::PolicyLookup3>; BODY; } break;               \
    case 16 + HIVE_POLICY_DEFAULT2HASH: { using P = ::hive::Inline8<::hive::PolicyDefault2>; BODY; } break;       \
    case 16 + HIVE_POLICY_MURMURCITY: { using P = ::hive::Inline8<::hive::PolicyMurmurCity>; BODY; } break;       \
    case 16 + HIVE_POLICY_TRIPLE: { using P = ::hive::Inline8<::hive::PolicyTriple>; BODY; } break;               \
    case 16 + HIVE_POLICY_MURMURCITYBIT: { using P = ::hive::Inline8<::hive::PolicyMurmurCityBit>; BODY; } break; \
    case 16 + HIVE_POLICY_LOOKUP2: { using P = ::hive::Inline8<::hive::PolicyLookup2>; BODY; } break;             \
    case 16 + HIVE_POLICY_LOOKUP3: { using P = ::hive::Inline8<::hive::PolicyLookup3>; BODY; } break;             \
    default: return ::hive::fail(HIVE_EINVAL, "unknown policy or layout");                                        \
  }

// The pending list of split inserts (inline_slow_path = 2) for batches of up
// to n ops (at any block size), allocated by the bulk insert entry point and
// the bulk autotuner, never under stream capture; launches without a list
// big enough run staged.  Sub-list capacity: the ops of the blocks of one
// class (blockIdx mod kPendingLists), at most ceil(blocks / kPendingLists)
// blocks of at most kMaxBlock threads each, which n / kPendingLists +
// kMaxBlock bounds for any block size.
inline size_t pending_sub_cap(size_t n) { return (n + kPendingLists - 1) / kPendingLists + kMaxBlock; }
// (force: whatever the table's insert launch option is)
inline hive_status_t ensure_pending(hive_table* t, size_t n, hipStream_t s, bool force = false) {
  if ((!force && t->launch[OP_INSERT].inline_slow_path != 2) || n <= t->pend_cap || n > 0xFFFFFFFFull) return HIVE_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return HIVE_OK;
  }
  const size_t ent = pending_sub_cap(n) * kPendingLists;
  unsigned long long* p = nullptr;
  if (hipMalloc(&p, ent * 12 + static_cast<size_t>(kPendingLists) * kPendingCountStride * 4 + 256) != hipSuccess) {
    (void)hipGetLastError();
    return HIVE_OK;  // no list: split inserts run staged
  }
  if (t->pend) {
    (void)hipDeviceSynchronize();  // the old list may still be in use by queued work
    (void)hipFree(t->pend);
  }
  t->pend = p;
  t->pend_cap = n;
  return HIVE_OK;
}
inline PendingList pending_list(const hive_table* t) {
  const size_t sub = pending_sub_cap(t->pend_cap), ent = sub * kPendingLists;
  uint32_t* idx = reinterpret_cast<uint32_t*>(t->pend + ent);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(t->pend + ent + (ent + 1) / 2);
  return PendingList{t->pend, idx, cnt, sub};
}

template <class Src, class Out>
hive_status_t launch_insert(hive_table* t, Src src, size_t n, Out out, int phase, hipStream_t s,
                            BatchClock clk = BatchClock{nullptr, nullptr}) {
  if (n == 0) return HIVE_OK;
  const uint32_t blk = t->launch[OP_INSERT].block_threads;
  if (!fits_one_launch(n, blk)) return fail(HIVE_EINVAL, "batch too large for one launch");
  if (t->launch[OP_INSERT].inline_slow_path == 2 && !t->launch[OP_INSERT].combine_duplicates && t->pend &&
      n <= t->pend_cap) {
    const PendingList pl = pending_list(t);
    int ncu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, t->device));
    const unsigned slow_grid = static_cast<unsigned>(std::max(1, ncu) * 16);
    k_fill_u32<<<kPendingLists * kPendingCountStride / 64, 64, 0, s>>>(pl.count, kPendingLists * kPendingCountStride,
                                                                       0u);
    HIVE_POLICY_SWITCH(table_kind(t), (k_insert_fast<P, Src, Out><<<grid_for(n, blk), blk, 0, s>>>(t->v, src, n, out,
                                                                                                phase, pl)));
    HIVE_POLICY_SWITCH(table_kind(t), (k_insert_slow<P, Out><<<slow_grid, 256, 0, s>>>(t->v, pl, out, clk)));
    HIP_CHECK_LAUNCH();
    return HIVE_OK;
  }
  if (t->launch[OP_INSERT].combine_duplicates) {
    HIVE_POLICY_SWITCH(table_kind(t), (k_insert_combined<P, Src, Out><<<grid_for(n, blk), blk, 0, s>>>(t->v, src, n, out,
                                                                                                    phase, clk)));
  } else if (t->launch[OP_INSERT].inline_slow_path == 1) {
    HIVE_POLICY_SWITCH(table_kind(t), (k_insert<P, Src, Out, true><<<grid_for(n, blk), blk, 0, s>>>(t->v, src, n, out,
                                                                                                 phase, clk)));
  } else {
    HIVE_POLICY_SWITCH(table_kind(t), (k_insert<P, Src, Out, false><<<grid_for(n, blk), blk, 0, s>>>(t->v, src, n, out,
                                                                                                  phase, clk)));
  }
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}

template <class Src, class Out>
hive_status_t launch_lookup(hive_table* t, Src src, size_t n, Out out, int phase, hipStream_t s) {
  if (n == 0) return HIVE_OK;
  const uint32_t blk = t->launch[OP_LOOKUP].block_threads;
  const uint32_t lanes = t->launch[OP_LOOKUP].lanes_per_key > 1 ? t->launch[OP_LOOKUP].lanes_per_key : 1;
  const uint32_t kpl = t->launch[OP_LOOKUP].keys_per_lane > 1 ? t->launch[OP_LOOKUP].keys_per_lane : 1;
  if (!fits_one_launch(n * lanes, blk)) return fail(HIVE_EINVAL, "batch too large for one launch");
  if (kpl == 4) {
    HIVE_POLICY_SWITCH(table_kind(t), (k_lookup_multi<P, Src, Out, 4><<<grid_for(n, blk * 4), blk, 0, s>>>(t->v, src, n, out,
                                                                                                    phase)));
  } else if (kpl == 2) {
    HIVE_POLICY_SWITCH(table_kind(t), (k_lookup_multi<P, Src, Out, 2><<<grid_for(n, blk * 2), blk, 0, s>>>(t->v, src, n, out,
                                                                                                    phase)));
  } else if (lanes == 8) {
    HIVE_POLICY_SWITCH(table_kind(t), (k_lookup_coop<P, Src, Out, 8><<<grid_for(n * 8, blk), blk, 0, s>>>(t->v, src, n, out,
                                                                                                   phase)));
  } else if (lanes == 4) {
    HIVE_POLICY_SWITCH(table_kind(t), (k_lookup_coop<P, Src, Out, 4><<<grid_for(n * 4, blk), blk, 0, s>>>(t->v, src, n, out,
                                                                                                   phase)));
  } else {
    HIVE_POLICY_SWITCH(table_kind(t), (k_lookup<P, Src, Out><<<grid_for(n, blk), blk, 0, s>>>(t->v, src, n, out, phase)));
  }
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}

template <class Src, class Out>
hive_status_t launch_erase(hive_table* t, Src src, size_t n, Out out, int phase, hipStream_t s) {
  if (n == 0) return HIVE_OK;
  const uint32_t blk = t->launch[OP_DELETE].block_threads;
  if (!fits_one_launch(n, blk)) return fail(HIVE_EINVAL, "batch too large for one launch");
  HIVE_POLICY_SWITCH(table_kind(t), (k_erase<P, Src, Out><<<grid_for(n, blk), blk, 0, s>>>(t->v, src, n, out, phase)));
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}

// a mixed batch = three stream-ordered phases: lookups, deletes, inserts
// Phases of a mixed batch that the caller knows to have ops (the lookup phase
// always runs: it also answers NONE / ATOMIC_INC ops).  A phase without ops
// still costs a launch whose blocks find nothing to do.
constexpr uint32_t kPhaseDelete = 1, kPhaseInsert = 2, kAllPhases = 3;

template <class Src, class Out>
hive_status_t launch_mixed(hive_table* t, Src src, size_t n, Out out, hipStream_t s,
                           BatchClock clk = BatchClock{nullptr, nullptr}, uint32_t phases = kAllPhases) {
  hive_status_t r;
  if ((r = launch_lookup(t, src, n, out, OP_LOOKUP, s)) != HIVE_OK) return r;
  if ((phases & kPhaseDelete) && (r = launch_erase(t, src, n, out, OP_DELETE, s)) != HIVE_OK) return r;
  if (!(phases & kPhaseInsert) && clk.ctr == nullptr) return HIVE_OK;
  return launch_insert(t, src, n, out, OP_INSERT, s, clk);
}

// In mixed batches of at least kCompactMinOps ops the inserts and deletes are
// compacted first (k_partition), so that their phase kernels run on dense
// lists instead of every wavefront of the batch carrying a few insert lanes
// through the whole insert chain; lookups keep the type-filtered kernel.
constexpr size_t kCompactMinOps = size_t(1) << 18;

// type-list scratch for batches of up to n ops (allocated outside any capture)
inline hive_status_t ensure_partition(hive_table* t, size_t n) {
  if (n < kCompactMinOps || n > 0xFFFFFFFFull || n <= t->part_cap) return HIVE_OK;
  uint32_t* p = nullptr;
  if (hipMalloc(&p, (9 * n + 4) * sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    return HIVE_OK;  // no scratch: batches run unsplit
  }
  if (t->part) {
    (void)hipDeviceSynchronize();  // the old lists may still be in use by queued work
    (void)hipFree(t->part);
  }
  t->part = p;
  t->part_cap = n;
  return HIVE_OK;
}

// a native mixed batch: split by type when it is large and scratch is there
// Type lists of a mixed batch of n ops in `part` (9 n + 3 words): keys, values
// and op indices [3 n each, by type], then the three counts.
constexpr size_t partition_words(size_t n) { return 9 * n + 3; }
inline bool use_partition(const hive_table* t, size_t n, const uint8_t* types) {
  return t->launch[OP_INSERT].split_by_type && n >= kCompactMinOps && n <= 0xFFFFFFFFull && types;
}
inline hive_status_t partition_batch(SrcNative src, size_t n, uint32_t* part, hipStream_t s) {
  // about 512 blocks, each owning a whole number of 1024-op tiles
  constexpr size_t kBlocks = 512;
  const size_t tiles = (n + kPartitionBlock - 1) / kPartitionBlock;
  const size_t span = ((tiles + kBlocks - 1) / kBlocks) * kPartitionBlock;
  k_fill_u32<<<1, 64, 0, s>>>(part + 9 * n, 3, 0u);
  k_partition<<<grid_for(n, static_cast<uint32_t>(span)), kPartitionBlock, 0, s>>>(
      src.types, src.keys, src.vals, n, span, part, part + 3 * n, part + 6 * n, part + 9 * n);
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}
// the phases of a mixed batch whose inserts and deletes partition_batch listed
inline hive_status_t launch_mixed_listed(hive_table* t, SrcNative src, size_t n, OutNative out, const uint32_t* part,
                                         hipStream_t s, BatchClock clk = BatchClock{nullptr, nullptr},
                                         uint32_t phases = kAllPhases) {
  const uint32_t* lkeys = part;
  const uint32_t* lvals = part + 3 * n;
  const uint32_t* lidx = part + 6 * n;
  const uint32_t* cnt = part + 9 * n;
  hive_status_t r;
  auto list = [&](int ty) {
    const size_t o = static_cast<size_t>(ty) * n;
    return SrcList{lkeys + o, lvals + o, cnt + ty, ty};
  };
  auto outs = [&](int ty) { return OutIdx{out.status, out.vals, lidx + static_cast<size_t>(ty) * n}; };
  if ((r = launch_lookup(t, src, n, out, OP_LOOKUP, s)) != HIVE_OK) return r;  // also answers NONE ops
  if ((phases & kPhaseDelete) &&
      (r = launch_erase(t, list(OP_DELETE), n, outs(OP_DELETE), OP_DELETE, s)) != HIVE_OK)
    return r;
  if (!(phases & kPhaseInsert) && clk.ctr == nullptr) return HIVE_OK;
  return launch_insert(t, list(OP_INSERT), n, outs(OP_INSERT), OP_INSERT, s, clk);
}

inline hive_status_t launch_mixed_native(hive_table* t, SrcNative src, size_t n, OutNative out, hipStream_t s,
                                         BatchClock clk = BatchClock{nullptr, nullptr}, uint32_t phases = kAllPhases) {
  if (!(phases & (kPhaseDelete | kPhaseInsert)) || !use_partition(t, n, src.types) || n > t->part_cap)
    return launch_mixed(t, src, n, out, s, clk, phases);
  hive_status_t r;
  if ((r = partition_batch(src, n, t->part, s)) != HIVE_OK) return r;
  return launch_mixed_listed(t, src, n, out, t->part, s, clk, phases);
}

// ---- windowed streams (hive_window.hip)
constexpr uint64_t kDefaultWindowOps = uint64_t(1) << 21;
constexpr uint64_t kDefaultWindowBatchOps = uint64_t(1) << 19;
constexpr size_t kMaxWindowBatches = 4096;
struct WindowPlan {
  size_t b0, b1;   // batches [b0, b1)
  uint32_t t0, t1; // tile descriptors [t0, t1) (multi-batch windows only)
};
hive_status_t ensure_window(hive_table* t, size_t ops);
hive_status_t free_window(hive_table* t);
hive_status_t plan_windows(const uint64_t* off, size_t nb, uint64_t window_ops, uint64_t batch_ops,
                           std::vector<WindowPlan>& plan, std::vector<uint4>& desc);
hive_status_t enqueue_windows(hive_table* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                              const uint64_t* off, uint32_t* out_vals, uint8_t* status, unsigned int* ctrs,
                              unsigned long long* ts, const std::vector<WindowPlan>& plan, const uint4* d_desc,
                              int ncu, hipStream_t s);
hive_status_t enqueue_window(hive_table* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                             uint64_t lo, uint64_t n, uint32_t* out_vals, uint8_t* status, const uint4* desc,
                             uint32_t ntiles, uint32_t base, unsigned long long* ts, uint32_t nts, int p, int ncu,
                             hipStream_t ks, hipStream_t s, hipEvent_t ready, uint32_t phases = kAllPhases);

// device buffers freed together when a call returns
struct DevArena {
  std::vector<void*> ptrs;
  ~DevArena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(T));
    if (e == hipSuccess) {
      ptrs.push_back(q);
      *p = static_cast<T*>(q);
    }
    return e;
  }
};

// every array of a view lives in one allocation, starting at v.slots (one
// hipMalloc / hipFree per view: each costs 0.1-0.5 ms, a resize's stall)
inline void free_view(TableView& v) {
  if (v.slots) (void)hipFree(v.slots);
  v = TableView{};
}

// ---- table views by layout (HIVE_LAYOUT_*)
inline bool valid_layout(int layout) {
  return layout == HIVE_LAYOUT_BUCKET8 || layout == HIVE_LAYOUT_BUCKET16 || layout == HIVE_LAYOUT_INLINE8;
}
inline uint32_t layout_slots(int layout) { return layout == HIVE_LAYOUT_BUCKET16 ? 16u : 8u; }
inline int layout_of(const TableView& v) {
  return v.layout == kLayoutInline ? HIVE_LAYOUT_INLINE8 : (v.S == 16 ? HIVE_LAYOUT_BUCKET16 : HIVE_LAYOUT_BUCKET8);
}
// stash buckets of an in-line table's pool: a quarter of the buckets
inline uint32_t pool_buckets(uint32_t nb) { return nb / 4 + 64; }
// counters, then (in-line layout) the pool cursors, each on a line of its own,
// then the holes flag on a line of its own
inline size_t pool_cursor_bytes() { return static_cast<size_t>(kPoolCursors) * kPoolCursorStride * 4; }
inline size_t counter_bytes() { return counter_words() * 8 + pool_cursor_bytes() + 128; }

inline size_t view_bytes(uint32_t nb, int layout) {
  if (layout == HIVE_LAYOUT_INLINE8)
    return static_cast<size_t>(nb) * kInlineStride * 8 + static_cast<size_t>(pool_buckets(nb)) * kStashSlots * 8 +
           counter_bytes();
  return static_cast<size_t>(nb) * layout_slots(layout) * 8 + static_cast<size_t>(nb) * kStashSlots * 8 +
         static_cast<size_t>(nb) * 4 * 2 + meta_words(nb) * 4 + counter_bytes();
}

inline hive_status_t alloc_view(TableView& v, uint32_t nb, int layout) {
  v = TableView{};
  v.nb = nb;
  v.S = layout_slots(layout);
  v.layout = layout == HIVE_LAYOUT_INLINE8 ? kLayoutInline : kLayoutSeparate;
  const bool inl = layout == HIVE_LAYOUT_INLINE8;
  if (inl) v.pool_cap = pool_buckets(nb);
  // carve: slots (or in-line lines), stash (or pool), then the separate
  // layout's per-bucket arrays, counters last; every part 256-byte aligned
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t b_slots = static_cast<size_t>(nb) * (inl ? kInlineStride : v.S) * 8;
  const size_t b_stash = static_cast<size_t>(inl ? v.pool_cap : nb) * kStashSlots * 8;
  const size_t b_u32 = inl ? 0 : static_cast<size_t>(nb) * 4;
  const size_t b_meta = inl ? 0 : meta_words(nb) * 4;
  const size_t total = up(b_slots) + up(b_stash) + 2 * up(b_u32) + up(b_meta) + up(counter_bytes());
  char* base = nullptr;
  const hipError_t e = hipMalloc(&base, total);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    v = TableView{};
    return fail(e == hipErrorOutOfMemory ? HIVE_ENOMEM : HIVE_EDEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base + off;
    off += up(bytes);
    return p;
  };
  v.slots = reinterpret_cast<unsigned long long*>(take(b_slots));
  v.stash = reinterpret_cast<unsigned long long*>(take(b_stash));
  if (!inl) {
    v.stash_count = reinterpret_cast<uint32_t*>(take(b_u32));
    v.locks = reinterpret_cast<uint32_t*>(take(b_u32));
    v.meta = reinterpret_cast<uint32_t*>(take(b_meta));
  }
  v.counters = reinterpret_cast<unsigned long long*>(take(counter_bytes()));
  if (inl) v.pool_next = reinterpret_cast<uint32_t*>(v.counters + counter_words());
  v.holes = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(v.counters + counter_words()) + pool_cursor_bytes());
  return HIVE_OK;
}

// zero buckets (with their in-line headers), stash pointers, hint bits, locks
// and counters (stash slots at or above a bucket's count are never read, so
// the stash body is left as is).
inline hive_status_t clear_view(const TableView& v, hipStream_t s) {
  const size_t cw = counter_bytes() / 8;
  k_fill_u64<<<fill_grid(cw), 256, 0, s>>>(v.counters, cw, 0ull);
  if (v.layout == kLayoutInline) {
    const size_t ns = static_cast<size_t>(v.nb) * kInlineStride;
    k_fill_u64<<<fill_grid(ns), 256, 0, s>>>(v.slots, ns, 0ull);
    HIP_CHECK_LAUNCH();
    return HIVE_OK;
  }
  const size_t ns = static_cast<size_t>(v.nb) * v.S;
  k_fill_u64<<<fill_grid(ns), 256, 0, s>>>(v.slots, ns, 0ull);
  k_fill_u32<<<fill_grid(v.nb), 256, 0, s>>>(v.stash_count, v.nb, 0u);
  k_fill_u32<<<fill_grid(meta_words(v.nb)), 256, 0, s>>>(v.meta, meta_words(v.nb), 0u);
  k_fill_u32<<<fill_grid(v.nb), 256, 0, s>>>(v.locks, v.nb, 0u);
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}


// rehash every live entry of src into dst (k_rehash, instantiated for dst's
// policy P and src's layout); failures counted in *failed
template <class P>
hive_status_t launch_rehash(const TableView& dst, const TableView& src, unsigned long long* failed, hipStream_t s) {
  const size_t n = rehash_threads(src.nb, src.S, src.layout == kLayoutInline);
  const unsigned g = grid_for(n, kDefaultBlock);
  switch (layout_of(src)) {
    case HIVE_LAYOUT_BUCKET16: k_rehash<P, LayoutOf<16, false>><<<g, kDefaultBlock, 0, s>>>(dst, src, failed); break;
    case HIVE_LAYOUT_INLINE8: k_rehash<P, LayoutOf<8, true>><<<g, kDefaultBlock, 0, s>>>(dst, src, failed); break;
    default: k_rehash<P, LayoutOf<8, false>><<<g, kDefaultBlock, 0, s>>>(dst, src, failed); break;
  }
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}

inline hive_status_t check_table(const hive_table* t) {
  if (!t) return fail(HIVE_EINVAL, "null table");
  return HIVE_OK;
}

inline hive_status_t set_device(const hive_table* t) {
  HIP_TRY(hipSetDevice(t->device));
  return HIVE_OK;
}

}  // namespace hive
