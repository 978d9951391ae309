// hive_kernels.hpp -- __global__ kernels of the MI355X Hive hash table.
//
// Work mapping: one operation per lane (lane-per-key).  A lane reads each
// 64-byte candidate bucket with four 16-byte loads, so a wavefront keeps 64
// independent random bucket reads in flight.  Divergent insert slow paths
// (cuckoo eviction, stash push) are staged through LDS: lanes whose fast path
// cannot finish append their op to a per-block LDS queue, and after a block
// barrier the queue is processed by consecutive lanes, so the slow path runs
// on densely packed wavefronts instead of one lane per wave.
#pragma once

#include <type_traits>

#include "hive_device.hpp"

namespace hive {

// ---------------------------------------------------------------- sources
// deferred_type(): the one op type a phase kernel skips without writing a
// result -- OP_DEFERRED for the internal types of a window (an op its chain
// kernels resolve), -1 for every caller-supplied source, whose ops of any
// type other than insert / lookup / delete get R_NOOP
// (include/hive_kernels.cuh:181-184).
// Batch inputs are read once and results written once: nontemporal loads and
// stores keep them from displacing table lines in L2 (2^26 lookups at LF 0.9:
// 1.763 -> 1.741 ms).
struct SrcNative {
  const uint32_t* keys;
  const uint32_t* vals;   // NULL: value = key + 1
  const uint8_t* types;   // NULL: every op has the kernel's own type
  int deferred = -1;      // OP_DEFERRED only for a window's internal types
  __device__ __forceinline__ uint32_t key(size_t i) const { return __builtin_nontemporal_load(keys + i); }
  __device__ __forceinline__ uint32_t val(size_t i, uint32_t k) const {
    return vals ? __builtin_nontemporal_load(vals + i) : k + 1u;
  }
  __device__ __forceinline__ int type(size_t i, int dflt) const { return types ? static_cast<int>(types[i]) : dflt; }
  __device__ __forceinline__ int deferred_type() const { return deferred; }
};

struct SrcOps {  // value = key + 1 (include/hive_kernels.cuh:156,238)
  const Op16* ops;
  __device__ __forceinline__ uint32_t key(size_t i) const { return static_cast<uint32_t>(ops[i].key); }
  __device__ __forceinline__ uint32_t val(size_t, uint32_t k) const { return k + 1u; }
  __device__ __forceinline__ int type(size_t i, int) const { return ops[i].type; }
  __device__ __forceinline__ int deferred_type() const { return -1; }
};

struct SrcKV {  // packed kv records (stash re-insert)
  const unsigned long long* kv;
  __device__ __forceinline__ uint32_t key(size_t i) const { return kv_key(kv[i]); }
  __device__ __forceinline__ uint32_t val(size_t i, uint32_t) const { return kv_val(kv[i]); }
  __device__ __forceinline__ int type(size_t, int dflt) const { return dflt; }
  __device__ __forceinline__ int deferred_type() const { return -1; }
};

template <class Src>
struct requires_bound : std::false_type {};

// One op type of a mixed batch, compacted by k_partition: list entry j holds
// the key and value of batch op idx[j].  Entries at or past *cnt do not exist
// and report another valid type, so every phase kernel skips them without
// touching the table or the outputs.  Keys and values are stored in the list
// (not fetched through idx), so a lane's first load is its own key, as in an
// uncompacted batch.  own = the list's type.
struct SrcList {
  const uint32_t* keys;
  const uint32_t* vals;  // NULL: value = key + 1
  const uint32_t* cnt;
  int own;
  __device__ __forceinline__ size_t bound() const { return *cnt; }  // blocks at or past it have nothing to do
  __device__ __forceinline__ uint32_t key(size_t j) const { return keys[j]; }
  __device__ __forceinline__ uint32_t val(size_t j, uint32_t k) const { return vals ? vals[j] : k + 1u; }
  __device__ __forceinline__ int type(size_t j, int) const { return j < *cnt ? own : (own + 1) % 3; }
  __device__ __forceinline__ int deferred_type() const { return -1; }
};

template <>
struct requires_bound<SrcList> : std::true_type {};

// Whole blocks past a compacted list's length return at once (after their
// batch-clock arrival); other sources have no such bound.
template <class Src>
__device__ __forceinline__ bool block_past_list(const Src& src) {
  if constexpr (requires_bound<Src>::value) return static_cast<size_t>(blockIdx.x) * blockDim.x >= src.bound();
  else return false;
}

// ---------------------------------------------------------------- outputs
struct OutNative {
  uint8_t* status;
  uint32_t* vals;
  __device__ __forceinline__ void put(size_t i, uint8_t st, uint32_t v) const {
    if (status) __builtin_nontemporal_store(st, status + i);
    if (vals) __builtin_nontemporal_store(v, vals + i);
  }
};

struct OutIdx {  // results of a compacted list, written at the batch op's index
  uint8_t* status;
  uint32_t* vals;
  const uint32_t* idx;
  __device__ __forceinline__ void put(size_t j, uint8_t st, uint32_t v) const {
    const uint32_t i = idx[j];
    if (status) status[i] = st;
    if (vals) vals[i] = v;
  }
};

struct OutU64 {  // 0/1 per op (include/hive_kernels.cuh:110-113,187-190)
  unsigned long long* res;
  __device__ __forceinline__ void put(size_t i, uint8_t st, uint32_t) const {
    if (res)
      res[i] = (st == R_INSERTED || st == R_UPDATED || st == R_STASHED || st == R_FOUND || st == R_ERASED) ? 1ull : 0ull;
  }
};

struct OutCount {  // discards per-op results, counts failures
  unsigned long long* failed;
  __device__ __forceinline__ void put(size_t, uint8_t st, uint32_t) const {
    if (st == R_FAILED) atomicAdd(failed, 1ull);
  }
};

// ---------------------------------------------------------------- counters
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Per-wavefront counter update: every lane of the wave must call it (no
// barrier, so waves whose work is done can retire independently).  Group 0
// counts bucket entries, 1 stash entries, 2 ops that took the insert slow path.
__device__ __forceinline__ void wave_count(const TableView& t, int dtab, int dstash, int dslow = 0) {
  const int wt = wave_sum(dtab);
  const int ws = wave_sum(dstash);
  const int wl = wave_sum(dslow);
  if ((threadIdx.x & 63) == 0) {
    const int lane = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kCounterLanes - 1);
    if (wt) atomicAdd(t.counters + (0 * kCounterLanes + lane) * kCounterStride,
                      static_cast<unsigned long long>(static_cast<long long>(wt)));
    if (ws) atomicAdd(t.counters + (1 * kCounterLanes + lane) * kCounterStride,
                      static_cast<unsigned long long>(static_cast<long long>(ws)));
    if (wl) atomicAdd(t.counters + (2 * kCounterLanes + lane) * kCounterStride,
                      static_cast<unsigned long long>(static_cast<long long>(wl)));
  }
}

// Per-block counter update, for kernels with a barrier after which every
// lane calls it once (the block counters zeroed before that barrier): each
// wavefront adds its sums to the block's LDS counters, and the last
// wavefront to arrive adds the block's totals to the device counters -- one
// device-scope atomic per counter per block instead of per wavefront.
struct BlockCounts {
  int tab, stash, slow;
  uint32_t arrived;
};
__device__ __forceinline__ void block_count(const TableView& t, BlockCounts& bc, int dtab, int dstash, int dslow) {
  const int wt = wave_sum(dtab);
  const int ws = wave_sum(dstash);
  const int wl = wave_sum(dslow);
  if ((threadIdx.x & 63) == 0) {
    if (wt) __hip_atomic_fetch_add(&bc.tab, wt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (ws) __hip_atomic_fetch_add(&bc.stash, ws, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (wl) __hip_atomic_fetch_add(&bc.slow, wl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint32_t waves = blockDim.x >> 6;
    if (__hip_atomic_fetch_add(&bc.arrived, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == waves - 1u) {
      const int a = __hip_atomic_load(&bc.tab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      const int b = __hip_atomic_load(&bc.stash, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      const int c = __hip_atomic_load(&bc.slow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      const int lane = static_cast<int>(blockIdx.x & (kCounterLanes - 1));
      if (a) atomicAdd(t.counters + (0 * kCounterLanes + lane) * kCounterStride,
                       static_cast<unsigned long long>(static_cast<long long>(a)));
      if (b) atomicAdd(t.counters + (1 * kCounterLanes + lane) * kCounterStride,
                       static_cast<unsigned long long>(static_cast<long long>(b)));
      if (c) atomicAdd(t.counters + (2 * kCounterLanes + lane) * kCounterStride,
                       static_cast<unsigned long long>(static_cast<long long>(c)));
    }
  }
}

constexpr int kMaxBlock = 1024;

// Optional completion clock of a launch: the last block to finish stamps the
// device wall clock (100 MHz on gfx950) into *ts.  Used to time batch
// sequences without extra launches.
struct BatchClock {
  unsigned int* ctr;
  unsigned long long* ts;
  __device__ __forceinline__ void finish() const {
    if (ctr == nullptr) return;
    wait_vmem();
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(ctr, 1u) == gridDim.x - 1)
      __hip_atomic_store(ts, static_cast<unsigned long long>(wall_clock64()), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
};


// ---------------------------------------------------------------- insert
// phase < 0: every op is an insert.  phase == OP_INSERT: only ops whose type
// is INSERT (the insert phase of a mixed batch).
// Inline = false: ops whose fast path cannot finish are staged in an LDS queue
//   and, after one block barrier, run their slow path on packed lanes
//   (fewer partially idle wavefronts: throughput).
// Inline = true: each lane continues straight into its slow path, no barrier
//   (shortest critical path: latency of small batches).
template <class P, class Src, class Out, bool Inline>
__global__ __launch_bounds__(kMaxBlock) void k_insert(TableView t, Src src, size_t n, Out out, int phase,
                                                      BatchClock clk = BatchClock{nullptr, nullptr}) {
  if (block_past_list(src)) {
    clk.finish();
    return;
  }
  const size_t base = static_cast<size_t>(blockIdx.x) * blockDim.x;
  const size_t i = base + threadIdx.x;
  const bool full_hints = *t.holes == 0u;
  int dtab = 0, dstash = 0, dslow = 0;
  // key and value are requested together with the op type (one round trip)
  const uint32_t key = i < n ? src.key(i) : 0u;
  const uint32_t val = i < n ? src.val(i, key) : 0u;
  if constexpr (Inline) {
    if (i < n && (phase < 0 || src.type(i, OP_INSERT) == OP_INSERT)) {
      uint8_t r = fast_insert<P>(t, key, val, full_hints);
      if (r == R_PENDING) {
        r = slow_insert<P>(t, key, val);
        dslow = 1;
      }
      out.put(i, r, 0);
      dtab += (r == R_INSERTED);
      dstash += (r == R_STASHED);
    }
  } else {
    __shared__ uint32_t queue[kMaxBlock];
    __shared__ uint32_t qn;
    __shared__ BlockCounts bc;
    if (threadIdx.x == 0) {
      qn = 0;
      bc = BlockCounts{0, 0, 0, 0u};
    }
    __syncthreads();
    if (i < n && (phase < 0 || src.type(i, OP_INSERT) == OP_INSERT)) {
      const uint8_t r = fast_insert<P>(t, key, val, full_hints);
      if (r == R_PENDING) {
        queue[atomicAdd(&qn, 1u)] = threadIdx.x;
      } else {
        out.put(i, r, 0);
        dtab += (r == R_INSERTED);
      }
    }
    __syncthreads();
    if (threadIdx.x < qn) {  // LDS-staged slow path on packed lanes
      const size_t idx = base + queue[threadIdx.x];
      const uint32_t key = src.key(idx);
      const uint8_t r = slow_insert<P>(t, key, src.val(idx, key));
      out.put(idx, r, 0);
      dtab += (r == R_INSERTED);
      dstash += (r == R_STASHED);
      dslow = 1;
    }
    block_count(t, bc, dtab, dstash, dslow);
    clk.finish();
    return;
  }
  wave_count(t, dtab, dstash, dslow);
  clk.finish();
}

// Split insert (launch option inline_slow_path = 2): two launches.  The first
// runs every lane's fast path and appends the ops that must take the slow
// path to a device list (key and value packed, and the op's index); the
// second runs the slow path over that list.  The fast-path kernel then
// carries only the fast path's registers (51 VGPRs, 8 waves per SIMD,
// against 78 and 6 for k_insert with the in-line 8-slot layout), so more of
// its single CASes are in flight.  The list is kPendingLists sub-lists, one
// per class of blocks (blockIdx mod kPendingLists), each with its own
// counter on its own 128-byte line: appends from every wavefront to one
// counter serialized at one address (the fast kernel took 7.7 ms instead of
// the 3 its atomics need).  Sub-list l holds the ops of the blocks of its
// class, at most sub_cap; its counter is zero on entry to the first launch.
constexpr uint32_t kPendingLists = 64;
constexpr uint32_t kPendingCountStride = 32;  // u32 between sub-list counters (128 B)
struct PendingList {
  unsigned long long* kv;
  uint32_t* idx;
  uint32_t* count;   // [kPendingLists * kPendingCountStride]
  size_t sub_cap;    // ops per sub-list
};

template <class P, class Src, class Out>
__global__ __launch_bounds__(kMaxBlock) void k_insert_fast(TableView t, Src src, size_t n, Out out, int phase,
                                                           PendingList pl) {
  if (block_past_list(src)) return;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int dtab = 0;
  const bool full_hints = *t.holes == 0u;
  const uint32_t key = i < n ? src.key(i) : 0u;  // requested together with the op type
  const uint32_t val = i < n ? src.val(i, key) : 0u;
  bool pend = false;
  if (i < n && (phase < 0 || src.type(i, OP_INSERT) == OP_INSERT)) {
    const uint8_t r = fast_insert<P>(t, key, val, full_hints);
    if (r == R_PENDING) {
      pend = true;
    } else {
      out.put(i, r, 0);
      dtab += (r == R_INSERTED);
    }
  }
  const unsigned long long b = __ballot(pend);
  if (b) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t l = blockIdx.x % kPendingLists;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(pl.count + l * kPendingCountStride, static_cast<uint32_t>(__popcll(b)));
    base = static_cast<uint32_t>(__shfl(static_cast<int>(base), 0));
    if (pend) {
      const size_t j = l * pl.sub_cap + base + static_cast<uint32_t>(__popcll(b & ((1ull << lane) - 1ull)));
      pl.kv[j] = kv_pack(key, val);
      pl.idx[j] = static_cast<uint32_t>(i);
    }
  }
  wave_count(t, dtab, 0, 0);
}

// The pending ops of all sub-lists form one index space (their counts'
// prefix sums, computed by each block in LDS), so the whole grid works on
// all of them at once.
template <class P, class Out>
__global__ __launch_bounds__(256) void k_insert_slow(TableView t, PendingList pl, Out out,
                                                     BatchClock clk = BatchClock{nullptr, nullptr}) {
  __shared__ uint32_t start[kPendingLists + 1];  // start[l]: first global index of sub-list l
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t l = 0; l < kPendingLists; ++l) {
      start[l] = acc;
      acc += pl.count[l * kPendingCountStride];  // written by the previous launch
    }
    start[kPendingLists] = acc;
  }
  __syncthreads();
  const uint32_t total = start[kPendingLists];
  int dtab = 0, dstash = 0, dslow = 0;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = kPendingLists;  // the sub-list l with start[l] <= g < start[l + 1]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (start[mid] <= g) lo = mid;
      else hi = mid;
    }
    const size_t e = lo * pl.sub_cap + (g - start[lo]);
    const unsigned long long kv = pl.kv[e];
    const uint8_t r = slow_insert<P>(t, kv_key(kv), kv_val(kv));
    out.put(pl.idx[e], r, 0);
    dtab += (r == R_INSERTED);
    dstash += (r == R_STASHED);
    ++dslow;
  }
  wave_count(t, dtab, dstash, dslow);
  clk.finish();
}

// Inserts with the duplicate keys of one block combined before they reach
// memory (launch option for skewed batches, where thousands of ops update a
// few hot keys).  Each key's first lane in a block-local LDS hash table is its
// leader and performs the insert with its own value (inline slow path); the
// other lanes of the key are linearized before it: if the leader placed the
// key, the first follower reports that placement and every other member
// UPDATED; if it updated or failed, every member reports the same.  The value
// left in the table is the leader's, the last of the group.
template <class P, class Src, class Out>
__global__ __launch_bounds__(kMaxBlock) void k_insert_combined(TableView t, Src src, size_t n, Out out, int phase,
                                                               BatchClock clk = BatchClock{nullptr, nullptr}) {
  if (block_past_list(src)) {
    clk.finish();
    return;
  }
  constexpr uint32_t kTab = 2 * kMaxBlock;  // at most half full
  __shared__ unsigned long long tab[kTab];  // key << 32 | (lane + 1), 0 = empty
  __shared__ uint32_t followers[kMaxBlock];
  __shared__ uint8_t res[kMaxBlock];
  for (uint32_t j = threadIdx.x; j < kTab; j += blockDim.x) tab[j] = 0ull;
  followers[threadIdx.x] = 0;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint32_t key = i < n ? src.key(i) : 0u;  // requested together with the op type
  const uint32_t val = i < n ? src.val(i, key) : 0u;
  const bool mine = i < n && (phase < 0 || src.type(i, OP_INSERT) == OP_INSERT);
  __syncthreads();
  uint32_t leader = threadIdx.x;
  bool first_follower = false;
  if (mine && key != 0) {
    const unsigned long long word = (static_cast<unsigned long long>(key) << 32) | (threadIdx.x + 1u);
    uint32_t h = (key * 0x9E3779B1u) >> 21;  // 11 bits: [0, kTab)
    for (;;) {
      const unsigned long long prev = atomicCAS(&tab[h], 0ull, word);
      if (prev == 0ull) break;
      if (static_cast<uint32_t>(prev >> 32) == key) {
        leader = static_cast<uint32_t>(prev & 0xFFFFFFFFull) - 1u;
        break;
      }
      h = (h + 1) & (kTab - 1);
    }
    if (leader != threadIdx.x) first_follower = atomicAdd(&followers[leader], 1u) == 0u;
  }
  __syncthreads();
  int dtab = 0, dstash = 0, dslow = 0;
  uint8_t r = R_FAILED;
  if (mine && leader == threadIdx.x) {
    r = fast_insert<P>(t, key, val, *t.holes == 0u);
    if (r == R_PENDING) {
      r = slow_insert<P>(t, key, val);
      dslow = 1;
    }
    dtab += (r == R_INSERTED);
    dstash += (r == R_STASHED);
    res[threadIdx.x] = r;
  }
  __syncthreads();
  if (mine) {
    uint8_t s;
    if (leader == threadIdx.x) {
      s = (followers[threadIdx.x] && (r == R_INSERTED || r == R_STASHED)) ? R_UPDATED : r;
    } else {
      const uint8_t rl = res[leader];
      s = (rl == R_INSERTED || rl == R_STASHED) ? (first_follower ? rl : R_UPDATED) : rl;
    }
    out.put(i, s, 0);
  }
  wave_count(t, dtab, dstash, dslow);
  clk.finish();
}

// ---------------------------------------------------------------- partition
// Compact the inserts and deletes of a mixed batch (n < 2^32): list t
// (OP_INSERT, OP_DELETE) occupies entries [t * n, t * n + cnt[t]) of lkeys /
// lvals / lidx (the op's key, value and batch index); cnt is zero on entry.
// Lookups stay in the batch (their phase kernel filters by type; they are the
// bulk of read-heavy batches, where a copy would only add traffic).  Order
// within a list is arbitrary: a phase's ops are concurrent.
// A block owns `span` consecutive ops (a multiple of its 1024 threads): it
// counts them by type first (wavefront ballots, LDS totals), takes its
// positions in each list with one device-scope atomic per type, then ranks
// and writes them.  Device-scope atomics on one address serialize (about
// 0.2 G/s), so one atomic per 1024 ops -- 8 k on two counters for a 4 M-op
// batch -- held the kernel near that rate; a span of many tiles keeps it to
// a few hundred.
constexpr uint32_t kPartitionBlock = 1024;
static __global__ __launch_bounds__(kPartitionBlock) void k_partition(const uint8_t* types, const uint32_t* keys,
                                                                      const uint32_t* vals, size_t n, size_t span,
                                                                      uint32_t* lkeys, uint32_t* lvals, uint32_t* lidx,
                                                                      uint32_t* cnt) {
  __shared__ uint32_t c[3], run[3], base[3];
  if (threadIdx.x < 3) {
    c[threadIdx.x] = 0;
    run[threadIdx.x] = 0;
  }
  __syncthreads();
  const size_t lo = static_cast<size_t>(blockIdx.x) * span;
  const size_t hi = lo + span < n ? lo + span : n;
  const uint32_t lane = threadIdx.x & 63u;
  const unsigned long long below = (1ull << lane) - 1ull;
  auto type_at = [&](size_t i) {
    int ty = -1;
    if (i < hi) {
      ty = static_cast<int>(types[i]);
      if (ty != OP_INSERT && ty != OP_DELETE) ty = -1;
    }
    return ty;
  };
  // pass 1: the block's counts per type
  for (size_t t0 = lo; t0 < hi; t0 += blockDim.x) {
    const int ty = type_at(t0 + threadIdx.x);
    const unsigned long long mi = __ballot(ty == OP_INSERT), md = __ballot(ty == OP_DELETE);
    if (lane == 0) {
      if (mi) atomicAdd(&c[OP_INSERT], static_cast<uint32_t>(__popcll(mi)));
      if (md) atomicAdd(&c[OP_DELETE], static_cast<uint32_t>(__popcll(md)));
    }
  }
  __syncthreads();
  if ((threadIdx.x == OP_INSERT || threadIdx.x == OP_DELETE) && c[threadIdx.x])
    base[threadIdx.x] = atomicAdd(cnt + threadIdx.x, c[threadIdx.x]);
  __syncthreads();
  // pass 2: ranks within the block (wavefront ballot, one LDS atomic per
  // wavefront and list) and the writes
  for (size_t t0 = lo; t0 < hi; t0 += blockDim.x) {
    const size_t i = t0 + threadIdx.x;
    const int ty = type_at(i);
    const unsigned long long mi = __ballot(ty == OP_INSERT), md = __ballot(ty == OP_DELETE);
    uint32_t wi = 0, wd = 0;
    if (mi) {
      const uint32_t lead = static_cast<uint32_t>(__ffsll(static_cast<long long>(mi)) - 1);
      uint32_t x = 0;
      if (lane == lead) x = atomicAdd(&run[OP_INSERT], static_cast<uint32_t>(__popcll(mi)));
      wi = __shfl(x, static_cast<int>(lead));
    }
    if (md) {
      const uint32_t lead = static_cast<uint32_t>(__ffsll(static_cast<long long>(md)) - 1);
      uint32_t x = 0;
      if (lane == lead) x = atomicAdd(&run[OP_DELETE], static_cast<uint32_t>(__popcll(md)));
      wd = __shfl(x, static_cast<int>(lead));
    }
    if (ty >= 0) {
      const bool ins = ty == OP_INSERT;
      const size_t pos = static_cast<size_t>(ty) * n + base[ty] + (ins ? wi : wd) +
                         static_cast<uint32_t>(__popcll((ins ? mi : md) & below));
      const uint32_t key = keys[i];
      lkeys[pos] = key;
      lvals[pos] = vals ? vals[i] : key + 1u;
      lidx[pos] = static_cast<uint32_t>(i);
    }
  }
}

// ---------------------------------------------------------------- lookup
template <class P, class Src, class Out>
__global__ __launch_bounds__(kMaxBlock) void k_lookup(TableView t, Src src, size_t n, Out out, int phase) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t key = src.key(i);  // requested together with the op type
  if (phase >= 0) {
    const int ty = src.type(i, OP_LOOKUP);
    if (ty != OP_LOOKUP) {
      if (ty != OP_INSERT && ty != OP_DELETE && ty != src.deferred_type()) out.put(i, R_NOOP, 0);
      return;
    }
  }
  uint32_t v;
  const uint8_t r = lookup_one<P>(t, key, v, *t.holes == 0u);
  out.put(i, r, v);
}

// the value of a bucket's lowest slot in match mask m (m != 0), by selects
// over the unrolled slots: a dynamic index into a register array would put
// the bucket in scratch memory
template <int S>
__device__ __forceinline__ uint32_t lowest_match_val(const unsigned long long (&s)[S], uint32_t m) {
  const int j = __builtin_ctz(m);
  uint32_t v = 0;
#pragma unroll
  for (int q = 0; q < S; ++q)
    if (q == j) v = kv_val(s[q]);
  return v;
}

// K keys per lane (K = 2 or 4): each lane runs lookup_one's probe sequence
// for K keys at once, issuing a round's bucket reads for all of them before
// consuming any, so a wavefront keeps K times as many random reads in flight
// and pays the key load, launch and output once per K keys.  Key k of a
// lane is op blockIdx.x * blockDim.x * K + k * blockDim.x + threadIdx.x,
// so key loads and result stores stay coalesced.
template <class P, class Src, class Out, int K>
__global__ __launch_bounds__(kMaxBlock) void k_lookup_multi(TableView t, Src src, size_t n, Out out, int phase) {
  const size_t base = static_cast<size_t>(blockIdx.x) * blockDim.x * K + threadIdx.x;
  uint32_t key[K], val[K], b0[K];
  bool open[K], act[K];
  uint8_t res[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const size_t i = base + static_cast<size_t>(k) * blockDim.x;
    act[k] = i < n;
    key[k] = act[k] ? src.key(i) : 0u;  // requested together with the op type
    if (act[k] && phase >= 0) {
      const int ty = src.type(i, OP_LOOKUP);
      if (ty != OP_LOOKUP) {
        if (ty != OP_INSERT && ty != OP_DELETE && ty != src.deferred_type()) out.put(i, R_NOOP, 0);
        act[k] = false;
      }
    }
    val[k] = 0;
    res[k] = R_NOT_FOUND;
    open[k] = act[k] && key[k] != 0u;
    b0[k] = bucket_of<P>(0, key[k], t.nb);
  }
  {  // round 1: every key's H0
    const bool full_hints = *t.holes == 0u;  // a miss in an H0 with an empty slot is final (note_hole)
    unsigned long long s[K][P::kS];
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (open[k]) read_bucket<P>(bucket_ptr<P>(t, b0[k]), s[k]);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!open[k]) continue;
      const uint32_t m = match_mask(s[k], key[k]);
      if (m) {
        val[k] = lowest_match_val(s[k], m);
        res[k] = R_FOUND;
        open[k] = false;
      } else if (full_hints && empty_mask(s[k]) != 0u) {
        open[k] = false;
      }
    }
  }
  // round 2: the hint words of the keys H0 missed (read only after the miss,
  // as in lookup_one)
  Hdr h[K];
#pragma unroll
  for (int k = 0; k < K; ++k) h[k] = open[k] ? read_hdr<P>(t, b0[k]) : Hdr{0u, 0u, 0u};
  // rounds 3..: the other candidates of the keys whose overflow class is set
#pragma unroll
  for (int c = 1; c < P::NH; ++c) {
    unsigned long long s[K][P::kS];
    bool probe[K];
    uint32_t bc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      bc[k] = bucket_of<P>(c, key[k], t.nb);
      bool dup = false;
#pragma unroll
      for (int j = 0; j < c; ++j) dup |= bucket_of<P>(j, key[k], t.nb) == bc[k];
      probe[k] = open[k] && (h[k].hint & ovf_bit_of<P>(key[k])) && !dup;
      if (probe[k]) read_bucket<P>(bucket_ptr<P>(t, bc[k]), s[k]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!probe[k]) continue;
      const uint32_t m = match_mask(s[k], key[k]);
      if (m) {
        val[k] = lowest_match_val(s[k], m);
        res[k] = R_FOUND;
        open[k] = false;
      }
    }
  }
  // the canonical stash (rare: only buckets whose stash bit is set)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!open[k] || !(h[k].hint & stash_bit<P>())) continue;
    const uint32_t cnt = stash_count_of<P>(t, b0[k], h[k]);
    const unsigned long long* sb = stash_of<P>(t, b0[k], h[k]);
    for (uint32_t j = 0; j < cnt; ++j) {
      const unsigned long long cur = read_u64<P>(sb + j);
      if (cur != 0ull && kv_key(cur) == key[k]) {
        val[k] = kv_val(cur);
        res[k] = R_FOUND;
        break;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (act[k]) out.put(base + static_cast<size_t>(k) * blockDim.x, res[k], val[k]);
}

// ------------------------------------------------------ cooperative lookup
// L lanes per key (L = 4 or 8, a group of consecutive lanes of one
// wavefront): lane r of a group loads slots [r*S/L, (r+1)*S/L) of each probed
// bucket of S slots, and one wavefront ballot gives every group the lowest
// matching slot.  The probe sequence and result are those of lookup_one: H0,
// the other candidates if H0's overflow bit is set, then stash[H0] if its
// stash bit is set (stash entries below its count, one per lane of an
// 8-lane group, two per lane of a 4-lane group).
template <int L>
__device__ __forceinline__ uint32_t group_bits(bool pred) {
  const unsigned long long b = __ballot(pred);
  const uint32_t base = (threadIdx.x & 63u) & ~static_cast<uint32_t>(L - 1);
  return static_cast<uint32_t>(b >> base) & ((1u << L) - 1u);
}

// this lane's part of a probed region: W consecutive slots from p
template <int W>
__device__ __forceinline__ void load_part(const unsigned long long* p, unsigned long long (&v)[W]) {
  if constexpr (W == 1) {
    v[0] = p[0];
  } else {
#pragma unroll
    for (int h = 0; h < W / 2; ++h) {
      const uint4 q = reinterpret_cast<const uint4*>(p)[h];
      v[2 * h] = (static_cast<unsigned long long>(q.y) << 32) | q.x;
      v[2 * h + 1] = (static_cast<unsigned long long>(q.w) << 32) | q.z;
    }
  }
}

// lowest slot of the group's region holding key: its value to every lane
template <int L, int W>
__device__ __forceinline__ bool group_match(const unsigned long long (&v)[W], uint32_t key, uint32_t& val) {
  int mine = -1;
#pragma unroll
  for (int j = W - 1; j >= 0; --j)
    if (v[j] != 0ull && kv_key(v[j]) == key) mine = j;
  const uint32_t bits = group_bits<L>(mine >= 0);
  if (!bits) return false;
  const uint32_t src = __builtin_ctz(bits);
  uint32_t my_val = 0;
#pragma unroll
  for (int j = 0; j < W; ++j)
    if (j == mine) my_val = kv_val(v[j]);
  const int lane0 = static_cast<int>((threadIdx.x & 63u) & ~static_cast<uint32_t>(L - 1));
  val = static_cast<uint32_t>(__shfl(static_cast<int>(my_val), lane0 + static_cast<int>(src)));
  return true;
}

template <class P, class Src, class Out, int L>
__global__ __launch_bounds__(kMaxBlock) void k_lookup_coop(TableView t, Src src, size_t n, Out out, int phase) {
  constexpr int W = P::kS / L;         // bucket slots per lane
  constexpr int WS = kStashSlots / L;  // stash slots per lane
  const size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) / L;
  const uint32_t r = threadIdx.x & (L - 1);
  if (i >= n) return;  // whole groups: n ops cover whole groups of lanes
  const uint32_t key = src.key(i);
  if (phase >= 0) {
    const int ty = src.type(i, OP_LOOKUP);
    if (ty != OP_LOOKUP) {
      if (r == 0 && ty != OP_INSERT && ty != OP_DELETE && ty != src.deferred_type()) out.put(i, R_NOOP, 0);
      return;
    }
  }
  uint32_t val = 0;
  uint8_t res = R_NOT_FOUND;
  if (key != 0) {
    uint32_t b[P::NH];
#pragma unroll
    for (int h = 0; h < P::NH; ++h) b[h] = bucket_of<P>(h, key, t.nb);
    unsigned long long v[W];
    load_part<W>(bucket_ptr<P>(t, b[0]) + r * W, v);
    if (group_match<L, W>(v, key, val)) {
      res = R_FOUND;
    } else {
      const Hdr hd = read_hdr<P>(t, b[0]);
      const uint32_t hint = hd.hint;
      if (hint & ovf_bit_of<P>(key)) {
#pragma unroll
        for (int h = 1; h < P::NH; ++h) {
          if (res == R_FOUND || is_dup<P>(b, h)) continue;
          load_part<W>(bucket_ptr<P>(t, b[h]) + r * W, v);
          if (group_match<L, W>(v, key, val)) res = R_FOUND;
        }
      }
      if (res != R_FOUND && (hint & stash_bit<P>())) {
        const uint32_t cnt = stash_count_of<P>(t, b[0], hd);
        unsigned long long vs[WS];
        load_part<WS>(stash_of<P>(t, b[0], hd) + r * WS, vs);
#pragma unroll
        for (int j = 0; j < WS; ++j)
          if (r * WS + j >= cnt) vs[j] = 0ull;  // past the count: not an entry
        if (group_match<L, WS>(vs, key, val)) res = R_FOUND;
      }
    }
  }
  if (r == 0) out.put(i, res, res == R_FOUND ? val : 0u);
}

// ---------------------------------------------------------------- erase
template <class P, class Src, class Out>
__global__ __launch_bounds__(kMaxBlock) void k_erase(TableView t, Src src, size_t n, Out out, int phase) {
  if (block_past_list(src)) return;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int rt = 0, rs = 0;
  const uint32_t key = i < n ? src.key(i) : 0u;  // requested together with the op type
  if (i < n && (phase < 0 || src.type(i, OP_DELETE) == OP_DELETE)) {
    const uint8_t r = erase_one<P>(t, key, rt, rs);
    out.put(i, r, 0);
  }
  wave_count(t, -rt, -rs);
}

// ---------------------------------------------------------------- rehash
// One thread per source entry, inserted into dst with the same fast/slow
// protocol.  A wavefront takes one slot of a row of 64 consecutive source
// buckets; region 0 holds the bucket slots of the source (PS: the source's
// layout, which may differ from dst's), region 1 its 8 stash slots.
// The keys of one source bucket share H0 in the source and land in the same
// two destination buckets, where lanes running at the same time would race
// for the same empty slots.  So a row's slots go to wavefronts kRehashChunk
// apart (rows are taken in chunks of kRehashChunk, slot by slot: about one
// chip's worth of resident wavefronts between them), and a chunk's source
// lines stay cached for its passes.
constexpr size_t kRehashRows = 64;      // source buckets per wavefront
constexpr size_t kRehashChunk = 8192;   // rows per chunk
__host__ __device__ inline size_t rehash_rows(uint32_t nb) { return (static_cast<size_t>(nb) + kRehashRows - 1) / kRehashRows; }
// threads of a k_rehash launch over a source of nb buckets of ss slots: one
// per bucket slot, then one per stash slot (separate layout) or one per
// bucket (in-line layout: its header is read once, its stash entries walked)
inline size_t rehash_threads(uint32_t nb, uint32_t ss, bool inl) {
  return rehash_rows(nb) * (ss + (inl ? 1 : kStashSlots)) * kRehashRows;
}
// the layout part of a policy (no hashes): a rehash's source
template <int S, bool Inl>
struct LayoutOf : PolicyBase {
  static constexpr int kS = S;
  static constexpr bool kInline = Inl;
};
template <class P, class PS>
__global__ __launch_bounds__(kMaxBlock) void k_rehash(TableView dst, TableView src, unsigned long long* failed) {
  constexpr int SS = PS::kS;
  __shared__ unsigned long long queue[kMaxBlock];
  __shared__ uint32_t qn;
  if (threadIdx.x == 0) qn = 0;
  __syncthreads();
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t rows = rehash_rows(src.nb), region0 = rows * SS;
  const size_t w = i / kRehashRows;
  unsigned long long kv = 0;
  int dtab = 0, dstash = 0;
  if constexpr (PS::kInline) {
    // in-line source: one thread per bucket walks its pooled stash (rare:
    // entries go straight through the inline slow path)
    if (w >= region0) {
      const size_t b = (w - region0) * kRehashRows + i % kRehashRows;
      if (b < src.nb) {
        const Hdr h = read_hdr<PS>(src, static_cast<uint32_t>(b));
        const uint32_t cnt = stash_count_of<PS>(src, static_cast<uint32_t>(b), h);
        const unsigned long long* sb = cnt ? stash_of<PS>(src, static_cast<uint32_t>(b), h) : nullptr;
        for (uint32_t j = 0; j < cnt; ++j) {
          const unsigned long long e = sb[j];
          if (e == 0ull) continue;
          uint8_t r = fast_insert<P>(dst, kv_key(e), kv_val(e), *dst.holes == 0u);
          if (r == R_PENDING) r = slow_insert<P>(dst, kv_key(e), kv_val(e));
          dtab += (r == R_INSERTED);
          dstash += (r == R_STASHED);
          if (r == R_FAILED) atomicAdd(failed, 1ull);
        }
      }
    }
  }
  if (w < region0 + (PS::kInline ? 0 : rows * kStashSlots)) {
    const bool in_stash = w >= region0;
    const size_t wr = in_stash ? w - region0 : w;
    const size_t ns = in_stash ? kStashSlots : SS;  // slots per row of this region
    const size_t chunk = wr / (kRehashChunk * ns), within = wr % (kRehashChunk * ns);
    const size_t crow = chunk * kRehashChunk, nrow = rows - crow < kRehashChunk ? rows - crow : kRehashChunk;
    const uint32_t slot = static_cast<uint32_t>(within / nrow);
    const size_t b = (crow + within % nrow) * kRehashRows + i % kRehashRows;
    if (b < src.nb) {
      if (!in_stash) {
        kv = bucket_ptr<PS>(src, static_cast<uint32_t>(b))[slot];
      } else {
        const Hdr h = read_hdr<PS>(src, static_cast<uint32_t>(b));
        if (slot < stash_count_of<PS>(src, static_cast<uint32_t>(b), h))
          kv = stash_of<PS>(src, static_cast<uint32_t>(b), h)[slot];
      }
    }
  }
  if (kv != 0ull) {
    const uint8_t r = fast_insert<P>(dst, kv_key(kv), kv_val(kv), *dst.holes == 0u);
    if (r == R_PENDING) queue[atomicAdd(&qn, 1u)] = kv;
    else dtab += (r == R_INSERTED);
  }
  __syncthreads();
  if (threadIdx.x < qn) {
    const unsigned long long q = queue[threadIdx.x];
    const uint8_t r = slow_insert<P>(dst, kv_key(q), kv_val(q));
    dtab += (r == R_INSERTED);
    dstash += (r == R_STASHED);
    if (r == R_FAILED) atomicAdd(failed, 1ull);
  }
  wave_count(dst, dtab, dstash);
}

// ---------------------------------------------------------------- utility
static __global__ void k_fill_u64(unsigned long long* p, size_t n, unsigned long long v) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

static __global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

// exact occupancy (include/hive_kernels.cuh:303-325, counting kv != 0): one
// thread per bucket slot; the first of each bucket also counts its stash
template <class P>
__global__ void k_count_exact(TableView t, unsigned long long* out) {
  const size_t ntab = static_cast<size_t>(t.nb) * P::kS;
  unsigned long long ct = 0, cs = 0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < ntab;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const uint32_t b = static_cast<uint32_t>(i / P::kS), j = static_cast<uint32_t>(i % P::kS);
    ct += (bucket_ptr<P>(t, b)[j] != 0ull);
    if (j == 0) {
      const Hdr h = read_hdr<P>(t, b);
      const uint32_t n = stash_count_of<P>(t, b, h);
      if (n) {
        const unsigned long long* sb = stash_of<P>(t, b, h);
        for (uint32_t q = 0; q < n; ++q) cs += (sb[q] != 0ull);
      }
    }
  }
  int a = static_cast<int>(ct), c = static_cast<int>(cs);
  a = wave_sum(a);
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) {
    if (a) atomicAdd(out, static_cast<unsigned long long>(a));
    if (c) atomicAdd(out + 1, static_cast<unsigned long long>(c));
  }
}

// Rebuild the hint bits from a table's content (after import): stash bit from
// each stash count, overflow bit of H0(key) for every entry stored outside
// its H0.  The hint bits must be zero first.  One thread per slot.
template <class P>
__global__ void k_rebuild_meta(TableView t) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(t.nb) * P::kS) return;
  const uint32_t b = static_cast<uint32_t>(i / P::kS);
  const unsigned long long kv = bucket_ptr<P>(t, b)[i % P::kS];
  if (kv != 0ull) {
    const uint32_t h0 = bucket_of<P>(0, kv_key(kv), t.nb);
    if (h0 != b) hint_set<P>(t, h0, ovf_bit_of<P>(kv_key(kv)));
  }
  if (i % P::kS == 0 && stash_count_of<P>(t, b, read_hdr<P>(t, b)) > 0) hint_set<P>(t, b, stash_bit<P>());
}

// drop every stash bit, keep the overflow bits, reset every stash (stash
// drain): one thread per hint word (separate layout) or per bucket (in-line)
template <class P>
__global__ void k_reset_stash(TableView t) {
  const size_t w = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if constexpr (P::kInline) {
    if (w < kPoolCursors) t.pool_next[w * kPoolCursorStride] = 0u;  // launched with at least 256 threads
    if (w >= t.nb) return;
    uint32_t* h = header_ptr<P>(t, static_cast<uint32_t>(w));
    h[0] &= ~kInlineStashBit;
    h[2] = 0u;
    h[3] = 0u;
  } else {
    uint32_t keep = 0;
    for (uint32_t j = 0; j < kHintsPerWord; ++j) keep |= kOverflowMask << (j * kHintBits);
    if (w < (static_cast<size_t>(t.nb) + kHintsPerWord - 1) / kHintsPerWord) t.meta[w] &= keep;
    for (size_t b = w * kHintsPerWord; b < (w + 1) * kHintsPerWord && b < t.nb; ++b) t.stash_count[b] = 0u;
  }
}

// stash drain (include/hive_stash_table.cuh:33-64; holes are skipped)
template <class P>
__global__ void k_stash_count_live(TableView t, unsigned long long* total) {
  const size_t b = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int c = 0;
  if (b < t.nb) {
    const Hdr h = read_hdr<P>(t, static_cast<uint32_t>(b));
    const uint32_t cnt = stash_count_of<P>(t, static_cast<uint32_t>(b), h);
    if (cnt) {
      const unsigned long long* sb = stash_of<P>(t, static_cast<uint32_t>(b), h);
      for (uint32_t j = 0; j < cnt; ++j) c += (sb[j] != 0ull);
    }
  }
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(total, static_cast<unsigned long long>(c));
}

template <class P>
__global__ void k_stash_compact(TableView t, unsigned long long* out, unsigned long long* pos) {
  const size_t b = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= t.nb) return;
  const Hdr h = read_hdr<P>(t, static_cast<uint32_t>(b));
  const uint32_t cnt = stash_count_of<P>(t, static_cast<uint32_t>(b), h);
  if (!cnt) return;
  const unsigned long long* sb = stash_of<P>(t, static_cast<uint32_t>(b), h);
  for (uint32_t j = 0; j < cnt; ++j) {
    const unsigned long long kv = sb[j];
    if (kv != 0ull) out[atomicAdd(pos, 1ull)] = kv;
  }
}

// The in-line layout in the separate layout's raw arrays (export) and back
// (import): slots[nb * 8], stash[nb * 8], stash_count[nb]; one thread per
// bucket.  Import takes a pool stash bucket for every bucket with a count
// and leaves the hint words zero (k_rebuild_meta follows).
template <class P>
__global__ void k_inline_to_raw(TableView t, unsigned long long* slots, unsigned long long* stash, uint32_t* count) {
  const size_t b = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= t.nb) return;
  const unsigned long long* bk = bucket_ptr<P>(t, static_cast<uint32_t>(b));
  for (int j = 0; j < P::kS; ++j) slots[b * P::kS + j] = bk[j];
  const Hdr h = read_hdr<P>(t, static_cast<uint32_t>(b));
  const uint32_t n = stash_count_of<P>(t, static_cast<uint32_t>(b), h);
  count[b] = h.count;
  const unsigned long long* sb = n ? stash_of<P>(t, static_cast<uint32_t>(b), h) : nullptr;
  for (int j = 0; j < kStashSlots; ++j) stash[b * kStashSlots + j] = (static_cast<uint32_t>(j) < n) ? sb[j] : 0ull;
}
template <class P>
__global__ void k_raw_to_inline(TableView t, const unsigned long long* slots, const unsigned long long* stash,
                                const uint32_t* count, unsigned long long* failed) {
  const size_t b = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= t.nb) return;
  unsigned long long* bk = bucket_ptr<P>(t, static_cast<uint32_t>(b));
  for (int j = 0; j < P::kS; ++j) bk[j] = slots[b * P::kS + j];
  uint32_t* h = header_ptr<P>(t, static_cast<uint32_t>(b));
  h[0] = 0u;
  h[1] = 0u;
  h[2] = 0u;
  h[3] = 0u;
  const uint32_t n = count[b];
  if (n == 0u) return;
  const uint32_t ref = pool_take(t, static_cast<uint32_t>(b));
  if (ref == 0u) {
    atomicAdd(failed, 1ull);
    return;
  }
  h[2] = n;
  h[3] = ref;
  unsigned long long* sb = t.stash + static_cast<size_t>(ref - 1u) * kStashSlots;
  for (int j = 0; j < kStashSlots; ++j) sb[j] = stash[b * kStashSlots + j];
}

// live bucket entries of the nb * S bucket slots, appended at *pos with one
// atomic per wavefront
template <class P>
__global__ void k_dump_slots(TableView t, unsigned long long* out, unsigned long long* pos) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t n = static_cast<size_t>(t.nb) * P::kS;
  const unsigned long long kv = i < n ? bucket_ptr<P>(t, static_cast<uint32_t>(i / P::kS))[i % P::kS] : 0ull;
  const unsigned long long live = __ballot(kv != 0ull);
  if (live == 0ull) return;  // uniform across the wavefront
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int first = __ffsll(static_cast<long long>(live)) - 1;
  unsigned long long base = 0;
  if (lane == first) base = atomicAdd(pos, static_cast<unsigned long long>(__popcll(live)));
  base = __shfl(base, first);
  if (kv != 0ull) out[base + __popcll(live & ((1ull << lane) - 1ull))] = kv;
}

// ---------------------------------------------------------------- probe stats
// Per key, following the probe order of fast_insert / lookup_one exactly:
//   out[0] buckets an insert's fast path reads (H0, plus the other distinct
//          candidates when H0's overflow bit is set)
//   out[1] buckets a lookup reads before its early exit
//   out[2] lookups that also read the stash
//   out[3] lookups that read the hint word (H0 misses)
// Block totals go to out[(blockIdx % 64) * 8 + field] (spread counters).
template <class P>
__global__ __launch_bounds__(256) void k_probe_stats(TableView t, const uint32_t* keys, size_t n,
                                                     unsigned long long* out) {
  __shared__ int blk[4];
  if (threadIdx.x < 4) blk[threadIdx.x] = 0;
  __syncthreads();
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int v[4] = {0, 0, 0, 0};
  if (i < n && keys[i] != 0) {
    const uint32_t key = keys[i];
    uint32_t b[P::NH];
#pragma unroll
    for (int j = 0; j < P::NH; ++j) b[j] = bucket_of<P>(j, key, t.nb);
    const uint32_t hint = read_hdr<P>(t, b[0]).hint;
    v[0] = 1;
    v[1] = 1;
    unsigned long long s[P::kS];
    load_bucket(bucket_ptr<P>(t, b[0]), s);
    bool found = match_mask(s, key) != 0;
    if (!found) v[3] = 1;
#pragma unroll
    for (int j = 1; j < P::NH; ++j) {
      if (!(hint & ovf_bit_of<P>(key)) || is_dup<P>(b, j)) continue;
      ++v[0];
      if (found) continue;
      ++v[1];
      load_bucket(bucket_ptr<P>(t, b[j]), s);
      found = match_mask(s, key) != 0;
    }
    if (!found && (hint & stash_bit<P>())) v[2] = 1;
  }
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const int w = wave_sum(v[f]);
    if ((threadIdx.x & 63) == 0 && w) atomicAdd(&blk[f], w);
  }
  __syncthreads();
  if (threadIdx.x < 4 && blk[threadIdx.x])
    atomicAdd(out + (blockIdx.x & 63) * 8 + threadIdx.x, static_cast<unsigned long long>(blk[threadIdx.x]));
}

// ---------------------------------------------------------------- ceilings
__device__ __forceinline__ uint32_t ceiling_index(size_t i, uint32_t seed) {
  return mix_murmur(static_cast<uint32_t>(i) ^ seed) ^ mix_hash2(static_cast<uint32_t>(i >> 32) + seed);
}

// n random S * 8-byte reads (one per lane), same load shape as a bucket
// probe (S = 8: 64 bytes, S = 16: a 128-byte line)
template <int S>
static __global__ __launch_bounds__(256) void k_ceiling_gather(const unsigned long long* buf, uint32_t nbuckets, size_t n,
                                                        uint32_t seed, uint32_t* sink) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = ceiling_index(i, seed) % nbuckets;
  unsigned long long s[S];
  load_bucket(buf + static_cast<size_t>(b) * S, s);
  unsigned long long acc = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) acc ^= s[j];
  if (static_cast<uint32_t>(acc) == 0x9e3779b9u && sink) sink[blockIdx.x] = static_cast<uint32_t>(acc >> 32);
}

// n random 64-byte bucket reads with L lanes cooperating on each bucket (one
// 64/L-byte piece per lane, L = 4, 8 or 16): the wavefront-per-bucket layout
// against which the lane-per-key kernels are measured
template <int L>
static __global__ __launch_bounds__(256) void k_ceiling_gather_coop(const unsigned long long* buf, uint32_t nbuckets,
                                                                    size_t n, uint32_t seed, uint32_t* sink) {
  const size_t tid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t i = tid / L;
  if (i >= n) return;
  const uint32_t part = static_cast<uint32_t>(tid % L);
  const uint32_t b = ceiling_index(i, seed) % nbuckets;
  uint32_t acc;
  if constexpr (L == 4) {
    const uint4 v = reinterpret_cast<const uint4*>(buf)[static_cast<size_t>(b) * 4 + part];
    acc = v.x ^ v.y ^ v.z ^ v.w;
  } else if constexpr (L == 8) {
    const uint2 v = reinterpret_cast<const uint2*>(buf)[static_cast<size_t>(b) * 8 + part];
    acc = v.x ^ v.y;
  } else {
    acc = reinterpret_cast<const uint32_t*>(buf)[static_cast<size_t>(b) * 16 + part];
  }
  if (acc == 0x9e3779b9u && sink) sink[0] = acc;  // anti-dead-code sink: one u32
}

// n random 64-bit CAS (one per lane)
static __global__ __launch_bounds__(256) void k_ceiling_cas(unsigned long long* buf, size_t nwords, size_t n, uint32_t seed) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t w = (static_cast<size_t>(ceiling_index(i, seed)) * 8u + (i & 7)) % nwords;
  cas64(buf + w, 0ull, static_cast<unsigned long long>(i) + 1ull);
}

}  // namespace hive
