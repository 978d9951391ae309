// hive_window.hip -- windowed streams: consecutive small batches of a stream
// run as one window, with the per-batch results of running them one by one.
//
// No reference counterpart: the reference's harness launches and synchronizes
// one hive_mixed_kernel per batch (src/hive_dispatch.cu:119-164), and so do
// this library's per-batch launches.  A small batch costs the latency of its
// three dependent phase kernels whatever its size; a window pays that latency
// once for many batches.
//
// Why the results are those of batch-by-batch execution: the outcome of an op
// depends only on the ops on the same key before it (the table's capacity
// aside, see below), and the phases of a batch are ordered lookups, deletes,
// inserts.  So it suffices that, for every key, its ops run in (batch, phase)
// order -- except that lookups with no write of the key between them may run
// in any order.  Per window:
//   k_win_mark   every op records its key in a window key set, keeping the
//                first batch of the window that writes (inserts or deletes)
//                the key;
//   k_win_level  ops up to that batch keep their type (before it the key is
//                only looked up); later ones are deferred (OP_DEFERRED) and
//                appended to a pending list;
//   phase kernels the window's lookups, deletes, inserts, as one mixed batch:
//                for each key, its lookups before its first write, then the
//                ops of that batch in phase order;
//   k_win_rounds round r runs, for every key with pending ops, the ops of its
//                earliest remaining batch, again lookups, deletes, inserts,
//                separated by grid barriers, until no op is pending.
// Ops on different keys run concurrently, as inside one batch.  What the
// table's capacity decides -- whether an insert lands in a bucket or the stash,
// and an insert failing on a full stash -- can differ, as it can between two
// inserts of one batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "hive_internal.hpp"

using namespace hive;

namespace {

constexpr uint32_t kWinBlock = 256;       // ops per tile of the mark/level kernels
constexpr uint32_t kRoundBlock = 1024;    // threads per block of the rounds kernel
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;
constexpr uint32_t kDoneBatch = 0xFFFFFFFFu;
constexpr uint32_t kNoWrite = 0xFFFFu;    // first-write field of a key only looked up
constexpr uint32_t kLine = 32;            // u32 words per control counter (own 128-byte line)

// Window key-set entry: epoch (16) | first batch of the window that writes
// the key, kNoWrite if none (16) | key (32).  Entries of other epochs are
// free, so the set is never cleared between windows.
__device__ __forceinline__ unsigned long long went(uint32_t epoch, uint32_t fw, uint32_t key) {
  return (static_cast<unsigned long long>((epoch << 16) | fw) << 32) | key;
}
__device__ __forceinline__ uint32_t went_epoch(unsigned long long e) { return static_cast<uint32_t>(e >> 48); }
__device__ __forceinline__ uint32_t went_fw(unsigned long long e) { return static_cast<uint32_t>(e >> 32) & 0xFFFFu; }
__device__ __forceinline__ uint32_t win_hash(uint32_t key) { return fmix32(key * 0x9E3779B1u + 0x7F4A7C15u); }

// Round word of a key: the earliest batch among its pending writes, tagged
// with the window and round it belongs to: (window seq << 12 | round) << 16 |
// (0xFFFF - batch).  atomicMax keeps the newest tag and, within it, the
// earliest batch, so the words never need clearing.  A word without the
// round's tag means the key has no pending write.
__device__ __forceinline__ unsigned long long round_word(unsigned long long seq, uint32_t r, uint32_t b) {
  return (((seq << 12) | r) << 16) | (0xFFFFu - b);
}
// the earliest pending written batch of a key in round r (kNoWrite: none)
__device__ __forceinline__ uint32_t round_write_batch(unsigned long long w, unsigned long long seq, uint32_t r) {
  return (w >> 16) == ((seq << 12) | r) ? 0xFFFFu - static_cast<uint32_t>(w & 0xFFFFu) : kNoWrite;
}

// control block of one window buffer, one 128-byte line per counter, all
// reset by k_win_mark: [0..2] ops pending for a round, [3] barrier arrivals,
// [4] barrier generation, [5..7] deletes and [8..10] inserts of a round,
// [11] blocks finished.  Per-round counters rotate over 3 slots (round r uses
// slot r%3 and clears slot (r+1)%3 or (r+2)%3 for a later round), so a slot
// is cleared only after every block has passed a barrier since reading it.
// The error word is shared by both buffers and sticky for the whole stream.
struct WinCtl {
  unsigned int* c;
  unsigned int* err;
  __device__ unsigned int* pending(uint32_t slot) const { return c + slot * kLine; }
  __device__ unsigned int* arrive() const { return c + 3 * kLine; }
  __device__ unsigned int* gen() const { return c + 4 * kLine; }
  __device__ unsigned int* count(int op, uint32_t slot) const {
    return c + ((op == OP_DELETE ? 5u : 8u) + slot) * kLine;
  }
  __device__ unsigned int* finished() const { return c + 11 * kLine; }
  __device__ unsigned int* error() const { return err; }
};
constexpr uint32_t kCtlCounters = 12;
constexpr uint32_t kCtlWords = (kCtlCounters + 1) * kLine;  // + the shared error word's line

// One block per tile of 256 ops of one batch: desc = {first op, end op,
// batch number, -}; the batch's index in the window is its number - base.
// Block 0 also resets the control block.
// Only writes (inserts, deletes) enter the set: a key's first written batch
// is all that decides deferral, and lookups find their key's entry, if any,
// by a read-only probe in k_win_level.  Slot reads are plain (a stale view
// only makes a CAS fail and return the current entry) except right after the
// set was cleared (coherent = 1).
__global__ __launch_bounds__(kWinBlock) void k_win_mark(const uint4* desc, uint32_t base, const uint8_t* types,
                                                        const uint32_t* keys, uint32_t wlo, unsigned long long* set,
                                                        uint32_t mask, uint32_t epoch, uint32_t coherent, uint32_t* eidx,
                                                        WinCtl ctl) {
  if (blockIdx.x == 0 && threadIdx.x < kCtlCounters)
    __hip_atomic_store(ctl.c + threadIdx.x * kLine, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint4 d = desc[blockIdx.x];
  const uint32_t i = d.x + threadIdx.x;
  if (i >= d.y) return;
  const int ty = types[i];
  const uint32_t key = keys[i];
  uint32_t e = kNoEntry;
  if (key != 0u && (ty == OP_INSERT || ty == OP_DELETE)) {
    const uint32_t fw = d.z - base, cap = mask + 1u;
    uint32_t h = win_hash(key) & mask;
    for (uint32_t probe = 0; probe < cap && e == kNoEntry; ++probe, h = (h + 1u) & mask) {
      unsigned long long cur = coherent ? ld_sync(set + h) : set[h];
      for (;;) {
        if (went_epoch(cur) != epoch) {  // free in this window: claim it
          const unsigned long long seen = cas64(set + h, cur, went(epoch, fw, key));
          if (seen == cur) { e = h; break; }
          cur = seen;
          continue;
        }
        if (static_cast<uint32_t>(cur) != key) break;  // another key: probe on
        if (went_fw(cur) <= fw) { e = h; break; }
        const unsigned long long seen = cas64(set + h, cur, went(epoch, fw, key));
        if (seen == cur) { e = h; break; }
        cur = seen;
      }
    }
  }
  eidx[i - wlo] = e;
}

// Ops in a batch after their key's first write in the window are deferred:
// typed OP_DEFERRED for the phase kernels and appended to the pending list
// (wavefront-aggregated); deferred writes fold their batch into round 0's
// word.  Ops up to and including that batch keep their type: before it the
// key is only looked up, so they see the table as it was before the window.
__global__ __launch_bounds__(kWinBlock) void k_win_level(const uint4* desc, uint32_t base, const uint8_t* types,
                                                         const uint32_t* keys, uint32_t wlo,
                                                         const unsigned long long* set, uint32_t mask, uint32_t epoch,
                                                         const uint32_t* eidx, uint8_t* wtype, unsigned long long* rw,
                                                         unsigned long long seq, uint32_t* pidx, uint32_t* pbid,
                                                         uint32_t* pe, WinCtl ctl) {
  const uint4 d = desc[blockIdx.x];
  const uint32_t i = d.x + threadIdx.x;
  const uint32_t bw = d.z - base;  // batch index in the window
  bool pend = false;
  uint32_t e = kNoEntry;
  if (i < d.y) {
    int ty = types[i];
    if (ty == OP_LOOKUP) {
      const uint32_t key = keys[i];
      // the key's entry exists only if the window writes the key
      for (uint32_t h = win_hash(key) & mask, probe = 0; key != 0u && probe <= mask; ++probe, h = (h + 1u) & mask) {
        const unsigned long long en = set[h];
        if (went_epoch(en) != epoch) break;  // a free slot ends the probe: no write of the key
        if (static_cast<uint32_t>(en) == key) {
          e = h;
          pend = bw > went_fw(en);
          break;
        }
      }
    } else {
      e = eidx[i - wlo];
      if (e != kNoEntry) pend = bw > went_fw(set[e]);
    }
    if (ty == OP_DEFERRED) ty = HIVE_OP_NONE;  // not an input type
    wtype[i - wlo] = pend ? static_cast<uint8_t>(OP_DEFERRED) : static_cast<uint8_t>(ty);
  }
  const unsigned long long m = __ballot(pend);
  if (m == 0ull) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t leader = static_cast<uint32_t>(__builtin_ctzll(m));
  uint32_t first = 0;
  if (lane == leader) first = atomicAdd(ctl.pending(0), static_cast<uint32_t>(__popcll(m)));
  first = __shfl(first, static_cast<int>(leader));
  if (pend) {
    const uint32_t pos = first + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
    pidx[pos] = i;
    pbid[pos] = bw;
    pe[pos] = e;
    if (types[i] != OP_LOOKUP) atomicMax(rw + e, round_word(seq, 0, bw));
  }
}

// Barrier over the nblk active blocks of the launch (all resident: the grid
// is at most one block per CU); a lone block needs only its own barrier.
// Each thread's memory operations are performed before its block arrives.  A
// wait longer than the limit, or an error raised elsewhere, sets the error
// flag and returns false in every block that sees it.
__device__ bool grid_sync(const WinCtl& ctl, uint32_t nblk, unsigned long long wait_ticks) {
  __shared__ uint32_t s_ok;
  wait_vmem();
  __syncthreads();
  if (nblk == 1u) return true;
  if (threadIdx.x == 0) {
    uint32_t ok = 1;
    const uint32_t g = ld_sync32(ctl.gen());
    const uint32_t arrived = __hip_atomic_fetch_add(ctl.arrive(), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nblk - 1u) {
      __hip_atomic_store(ctl.arrive(), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      wait_vmem();  // the reset is performed before the generation moves on
      __hip_atomic_fetch_add(ctl.gen(), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const unsigned long long t0 = wall_clock64();
      uint32_t spins = 0;
      while (ld_sync32(ctl.gen()) == g) {
        __builtin_amdgcn_s_sleep(2);
        if ((++spins & 63u) == 0 && (ld_sync32(ctl.error()) != 0u || wall_clock64() - t0 > wait_ticks)) {
          atomicOr(ctl.error(), 1u);
          ok = 0;
          break;
        }
      }
    }
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0u;
}

struct RoundArgs {
  const uint8_t* types;
  const uint32_t* keys;
  const uint32_t* vals;
  uint32_t* out_vals;
  uint8_t* status;
  const uint32_t* pidx;
  uint32_t* pbid;             // kDoneBatch once the op has run
  const uint32_t* pe;
  unsigned long long* rw;     // 2 slots x cap round words
  uint32_t cap;
  unsigned long long seq;     // this window's sequence number
  WinCtl ctl;
  unsigned long long wait_ticks;
  unsigned long long* ts;     // completion stamps of the window's batches
  uint32_t nts;
};

// one clock reading for every batch of the window (called by a whole block)
__device__ void stamp_window(const RoundArgs& a) {
  __shared__ unsigned long long s_now;
  if (threadIdx.x == 0) s_now = wall_clock64();
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < a.nts; j += blockDim.x)
    __hip_atomic_store(a.ts + j, s_now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Deferred ops, round by round.  With m the key's earliest pending written
// batch (round word of slot r%2, complete since round r-1 or k_win_level; no
// word: no pending write), round r runs the key's pending lookups of batches
// up to m -- lookups between two writes commute, so they all see the writes
// before them -- then the deletes and inserts of batch m.  Writes that stay
// fold their batch into slot (r+1)%2 under round r+1's tag.  The rounds a key
// needs are its written batches, not every batch that touches it.  The lookup
// pass also counts the round's deletes and inserts and the ops that stay, so
// the barriers after it are only taken when a later pass has work.  Table
// reads are device-coherent: passes are separated by barriers, not kernel
// boundaries.  The last block stamps the window's batches complete.
// Only the blocks the pending list needs (one op per thread where it fits)
// take part; the others return at once.
template <class P0>
__global__ __launch_bounds__(kRoundBlock) void k_win_rounds(TableView t, RoundArgs a) {
  using P = CoherentReads<P0>;
  __shared__ uint32_t s_cnt[3];
  const uint32_t npend = ld_sync32(a.ctl.pending(0));
  if (npend == 0u) {  // the same answer in every block; the phase kernels are done
    if (blockIdx.x == 0) stamp_window(a);
    return;
  }
  const uint32_t nblk = min(gridDim.x, (npend + blockDim.x - 1u) / blockDim.x);
  if (blockIdx.x >= nblk) return;
  int dtab = 0, dstash = 0, dslow = 0;
  const uint32_t stride = nblk * blockDim.x;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t r = 0;; ++r) {
    const unsigned long long* cur = a.rw + (r & 1u) * a.cap;
    unsigned long long* nxt = a.rw + ((r + 1u) & 1u) * a.cap;
    if (threadIdx.x < 3) s_cnt[threadIdx.x] = 0;
    const uint32_t cs = r % 3u;
    if (tid == 0) {
      __hip_atomic_store(a.ctl.pending((r + 2u) % 3u), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(a.ctl.count(OP_DELETE, (r + 1u) % 3u), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(a.ctl.count(OP_INSERT, (r + 1u) % 3u), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    // lookups of this round; ops of later rounds fold their batch into the next word
    uint32_t stay = 0, ndel = 0, nins = 0;
    for (uint32_t j = tid; j < npend; j += stride) {
      const uint32_t b = a.pbid[j];
      if (b == kDoneBatch) continue;
      const uint32_t e = a.pe[j];
      const uint32_t m = round_write_batch(ld_sync(cur + e), a.seq, r);
      const uint32_t i = a.pidx[j];
      const int ty = a.types[i];
      if (ty != OP_LOOKUP) {
        if (b != m) {  // a later written batch: a later round
          atomicMax(nxt + e, round_word(a.seq, r + 1u, b));
          ++stay;
          continue;
        }
        ndel += ty == OP_DELETE;
        nins += ty == OP_INSERT;
        continue;
      }
      if (b > m) {  // after this round's write of the key
        ++stay;
        continue;
      }
      uint32_t v;
      const uint8_t res = lookup_one<P>(t, a.keys[i], v);
      if (a.status) a.status[i] = res;
      if (a.out_vals) a.out_vals[i] = v;
      a.pbid[j] = kDoneBatch;
    }
    if (stay) atomicAdd(&s_cnt[0], stay);
    if (ndel) atomicAdd(&s_cnt[1], ndel);
    if (nins) atomicAdd(&s_cnt[2], nins);
    __syncthreads();
    if (threadIdx.x == 0) {
      if (s_cnt[0]) atomicAdd(a.ctl.pending((r + 1u) % 3u), s_cnt[0]);
      if (s_cnt[1]) atomicAdd(a.ctl.count(OP_DELETE, cs), s_cnt[1]);
      if (s_cnt[2]) atomicAdd(a.ctl.count(OP_INSERT, cs), s_cnt[2]);
    }
    if (!grid_sync(a.ctl, nblk, a.wait_ticks)) break;
    // final after the barrier, and the same in every block
    const bool more = ld_sync32(a.ctl.pending((r + 1u) % 3u)) != 0u;
    const bool any_ins = ld_sync32(a.ctl.count(OP_INSERT, cs)) != 0u;
    if (ld_sync32(a.ctl.count(OP_DELETE, cs)) != 0u) {
      for (uint32_t j = tid; j < npend; j += stride) {  // deletes of this round
        const uint32_t b = a.pbid[j];
        if (b == kDoneBatch || round_write_batch(ld_sync(cur + a.pe[j]), a.seq, r) != b) continue;
        const uint32_t i = a.pidx[j];
        if (a.types[i] != OP_DELETE) continue;
        int rt, rs;
        const uint8_t res = erase_one<P>(t, a.keys[i], rt, rs);
        if (a.status) a.status[i] = res;
        if (a.out_vals) a.out_vals[i] = 0;
        dtab -= rt;
        dstash -= rs;
        a.pbid[j] = kDoneBatch;
      }
      if ((any_ins || more) && !grid_sync(a.ctl, nblk, a.wait_ticks)) break;
    }
    if (any_ins) {
      for (uint32_t j = tid; j < npend; j += stride) {  // inserts of this round
        const uint32_t b = a.pbid[j];
        if (b == kDoneBatch || round_write_batch(ld_sync(cur + a.pe[j]), a.seq, r) != b) continue;
        const uint32_t i = a.pidx[j];
        const uint32_t key = a.keys[i];
        const uint32_t val = a.vals ? a.vals[i] : key + 1u;
        uint8_t res = fast_insert<P>(t, key, val);
        if (res == R_PENDING) {
          res = slow_insert<P>(t, key, val);
          ++dslow;
          dstash += (res == R_STASHED);
        }
        dtab += (res == R_INSERTED);
        if (a.status) a.status[i] = res;
        if (a.out_vals) a.out_vals[i] = 0;
        a.pbid[j] = kDoneBatch;
      }
      if (more && !grid_sync(a.ctl, nblk, a.wait_ticks)) break;
    }
    if (!more) break;
  }
  // the last block to finish stamps the window's batches complete
  __shared__ uint32_t s_last;
  wait_vmem();
  __syncthreads();
  if (threadIdx.x == 0)
    s_last = __hip_atomic_fetch_add(a.ctl.finished(), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nblk - 1u;
  __syncthreads();
  if (s_last) stamp_window(a);
  wave_count(t, dtab, dstash, dslow);
}

__global__ void k_stamp_range(unsigned long long* ts, uint32_t n) {
  __shared__ unsigned long long s_now;
  if (threadIdx.x == 0) s_now = wall_clock64();
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < n; j += blockDim.x)
    __hip_atomic_store(ts + j, s_now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

namespace hive {

namespace {
// win_set (u64): the key set [cap], then per buffer the round words [2 cap].
// win_buf (u32): per buffer eidx, pidx, pbid, pe [cap / 2 each], wtype
// [cap / 2 bytes], its control block on a 128-byte line; then the shared
// error word's line.
size_t ctl_offset(size_t cap) {
  const size_t half = cap / 2;
  return (4 * half + (half + 3) / 4 + kLine - 1) / kLine * kLine;
}
size_t buffer_words(size_t cap) { return ctl_offset(cap) + kCtlCounters * kLine; }

struct WindowBuffer {
  unsigned long long* rw;
  uint32_t *eidx, *pidx, *pbid, *pe;
  uint8_t* wtype;
  WinCtl ctl;
};

WindowBuffer window_buffer(const hive_table* t, int p) {
  const size_t cap = t->win_cap, half = cap / 2;
  uint32_t* base = t->win_buf + static_cast<size_t>(p) * buffer_words(cap);
  WindowBuffer w;
  w.rw = t->win_set + cap + static_cast<size_t>(p) * 2 * cap;
  w.eidx = base;
  w.pidx = base + half;
  w.pbid = base + 2 * half;
  w.pe = base + 3 * half;
  w.wtype = reinterpret_cast<uint8_t*>(base + 4 * half);
  w.ctl = WinCtl{base + ctl_offset(cap), t->win_buf + 2 * buffer_words(cap)};
  return w;
}
}  // namespace

unsigned int* window_error_word(const hive_table* t) {
  return t->win_buf ? t->win_buf + 2 * buffer_words(t->win_cap) : nullptr;
}

// key set (2 entries per op, a power of two) and two window buffers (round
// words, per-op and pending arrays) for windows of up to `ops` ops
hive_status_t ensure_window(hive_table* t, size_t ops) {
  size_t cap = 1024;
  while (cap < 2 * ops) cap <<= 1;
  if (t->win_set && cap <= t->win_cap) return HIVE_OK;
  if (cap > (size_t(1) << 31)) return fail(HIVE_EINVAL, "window too large");
  const size_t words = 2 * buffer_words(cap) + kLine;
  unsigned long long* set = nullptr;
  uint32_t* buf = nullptr;
  if (hipMalloc(&set, 5 * cap * 8) != hipSuccess) {
    (void)hipGetLastError();
    return fail(HIVE_ENOMEM, "window key set");
  }
  if (hipMalloc(&buf, words * 4) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(set);
    return fail(HIVE_ENOMEM, "window buffers");
  }
  if (hipMemset(set, 0, 5 * cap * 8) != hipSuccess || hipMemset(buf, 0, words * 4) != hipSuccess) {
    (void)hipFree(set);
    (void)hipFree(buf);
    return fail(HIVE_EDEVICE, "window buffers: clear");
  }
  if (t->win_set) {
    (void)hipDeviceSynchronize();  // queued work may still use the old buffers
    (void)hipFree(t->win_set);
    (void)hipFree(t->win_buf);
  }
  t->win_set = set;
  t->win_buf = buf;
  t->win_cap = cap;
  t->win_epoch = 0;  // every entry and round word is zero: older than any window
  t->win_seq = 0;
  return HIVE_OK;
}

namespace {
// side stream and events of one windowed stream, released when it returns
struct Pipeline {
  hipStream_t ms = nullptr;
  hipEvent_t start = nullptr, done = nullptr, ready[2] = {nullptr, nullptr}, free[2] = {nullptr, nullptr};
  ~Pipeline() {
    if (ms) (void)hipStreamSynchronize(ms);
    for (hipEvent_t e : {start, done, ready[0], ready[1], free[0], free[1]})
      if (e) (void)hipEventDestroy(e);
    if (ms) (void)hipStreamDestroy(ms);
  }
};
}  // namespace

// One multi-batch window: ops [lo, lo + n), described by ntiles tiles from
// desc (batch numbers base .. base + nbatches - 1), in window buffer p.  The
// key-set passes run on ks, the table passes on s; when they differ, s waits
// for the key-set passes through `ready`, and the caller orders ks after the
// buffer's last use.  ts/nts: completion stamps of the window's batches (nts
// may be 0).
hive_status_t enqueue_window(hive_table* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                             uint64_t lo, uint64_t n, uint32_t* out_vals, uint8_t* status, const uint4* desc,
                             uint32_t ntiles, uint32_t base, unsigned long long* ts, uint32_t nts, int p, int ncu,
                             unsigned long long wait_ticks, hipStream_t ks, hipStream_t s, hipEvent_t ready) {
  const size_t cap = t->win_cap;
  unsigned long long* set = t->win_set;
  const uint32_t mask = static_cast<uint32_t>(cap - 1);
  const WindowBuffer wb = window_buffer(t, p);
  uint32_t coherent = 0;
  if (++t->win_epoch > 0xFFFFu) {  // epochs wrapped: free every entry
    k_fill_u64<<<fill_grid(cap), 256, 0, ks>>>(set, cap, 0ull);
    HIP_CHECK_LAUNCH();
    t->win_epoch = 1;
    coherent = 1;
  }
  const unsigned long long seq = ++t->win_seq;
  if (seq >= (1ull << 36)) return fail(HIVE_EINVAL, "windowed batches: window sequence exhausted");
  k_win_mark<<<ntiles, kWinBlock, 0, ks>>>(desc, base, types, keys, static_cast<uint32_t>(lo), set, mask,
                                           t->win_epoch, coherent, wb.eidx, wb.ctl);
  HIP_CHECK_LAUNCH();
  k_win_level<<<ntiles, kWinBlock, 0, ks>>>(desc, base, types, keys, static_cast<uint32_t>(lo), set, mask,
                                            t->win_epoch, wb.eidx, wb.wtype, wb.rw, seq, wb.pidx, wb.pbid, wb.pe,
                                            wb.ctl);
  HIP_CHECK_LAUNCH();
  if (ks != s) {
    HIP_TRY(hipEventRecord(ready, ks));
    HIP_TRY(hipStreamWaitEvent(s, ready, 0));
  }
  hive_status_t r = launch_mixed_native(t, SrcNative{keys + lo, vals ? vals + lo : nullptr, wb.wtype}, n,
                                        OutNative{status ? status + lo : nullptr, out_vals ? out_vals + lo : nullptr}, s);
  if (r != HIVE_OK) return r;
  RoundArgs ra{types, keys, vals, out_vals, status, wb.pidx, wb.pbid, wb.pe, wb.rw, static_cast<uint32_t>(cap), seq,
               wb.ctl, wait_ticks, ts, nts};
  // a small grid, every block resident: the rounds are separated by grid barriers
  const unsigned round_grid = static_cast<unsigned>(std::max(1, std::min(ncu, 64)));
  HIVE_POLICY_SWITCH(t->policy, (k_win_rounds<P><<<round_grid, kRoundBlock, 0, s>>>(t->v, ra)));
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}

// Multi-batch windows alternate between two buffers.  The key-set passes
// (k_win_mark, k_win_level) touch only the inputs and the window buffer, so
// they run on a side stream one window ahead, overlapping the table work of
// the previous window; a window's phase kernels wait for its key-set passes,
// and a buffer's next key-set passes wait for its last window's rounds.
hive_status_t enqueue_windows(hive_table* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                              const uint64_t* off, uint32_t* out_vals, uint8_t* status, unsigned int* ctrs,
                              unsigned long long* ts, const std::vector<WindowPlan>& plan, const uint4* d_desc,
                              int ncu, unsigned long long wait_ticks, hipStream_t s) {
  Pipeline pl;
  HIP_TRY(hipStreamCreateWithFlags(&pl.ms, hipStreamNonBlocking));
  for (hipEvent_t* e : {&pl.start, &pl.done, &pl.ready[0], &pl.ready[1], &pl.free[0], &pl.free[1]})
    HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(pl.start, s));  // inputs written on s before the call
  HIP_TRY(hipStreamWaitEvent(pl.ms, pl.start, 0));
  bool used[2] = {false, false};
  uint32_t k = 0;
  for (const WindowPlan& w : plan) {
    const uint64_t lo = off[w.b0], n = off[w.b1] - lo;
    const uint32_t ntiles = w.t1 - w.t0;
    if (w.b1 - w.b0 == 1 && n) {  // one batch: its plain phase kernels
      hive_status_t r = launch_mixed_native(t, SrcNative{keys + lo, vals ? vals + lo : nullptr, types + lo}, n,
                                            OutNative{status ? status + lo : nullptr, out_vals ? out_vals + lo : nullptr},
                                            s, BatchClock{ctrs + w.b0, ts + w.b0 + 1});
      if (r != HIVE_OK) return r;
      continue;
    }
    if (ntiles == 0) {  // only empty batches
      k_stamp_range<<<1, 64, 0, s>>>(ts + w.b0 + 1, static_cast<uint32_t>(w.b1 - w.b0));
      HIP_CHECK_LAUNCH();
      continue;
    }
    const int p = static_cast<int>(k++ & 1u);
    if (used[p]) HIP_TRY(hipStreamWaitEvent(pl.ms, pl.free[p], 0));
    hive_status_t r = enqueue_window(t, types, keys, vals, lo, n, out_vals, status, d_desc + w.t0, ntiles,
                                     static_cast<uint32_t>(w.b0), ts + w.b0 + 1, static_cast<uint32_t>(w.b1 - w.b0),
                                     p, ncu, wait_ticks, pl.ms, s, pl.ready[p]);
    if (r != HIVE_OK) return r;
    HIP_TRY(hipEventRecord(pl.free[p], s));
    used[p] = true;
  }
  HIP_TRY(hipEventRecord(pl.done, pl.ms));  // the caller's stream covers the side stream's work
  HIP_TRY(hipStreamWaitEvent(s, pl.done, 0));
  return HIVE_OK;
}

// Windows: runs of consecutive batches of fewer than batch_ops ops each, while
// their ops fit in window_ops (at most kMaxWindowBatches batches); any other
// batch is a window alone and runs as plain phase kernels.  Tiles of 256 ops
// (one batch each, with its batch number) describe the ops of multi-batch
// windows.
hive_status_t plan_windows(const uint64_t* off, size_t nb, uint64_t window_ops, uint64_t batch_ops,
                           std::vector<WindowPlan>& plan, std::vector<uint4>& desc) {
  plan.clear();
  desc.clear();
  auto size = [&](size_t b) { return off[b + 1] - off[b]; };
  for (size_t b = 0; b < nb;) {
    size_t b1 = b + 1;
    uint64_t ops = size(b);
    if (ops < batch_ops)
      while (b1 < nb && b1 - b < kMaxWindowBatches && size(b1) < batch_ops && ops + size(b1) <= window_ops) {
        ops += size(b1);
        ++b1;
      }
    WindowPlan w{b, b1, static_cast<uint32_t>(desc.size()), 0};
    if (b1 - b > 1)
      for (size_t c = b; c < b1; ++c)
        for (uint64_t i = off[c]; i < off[c + 1]; i += kWinBlock) {
          if (desc.size() >= 0x7FFFFFFFu) return fail(HIVE_EINVAL, "windowed stream: too many tiles");
          desc.push_back(make_uint4(static_cast<uint32_t>(i),
                                    static_cast<uint32_t>(std::min<uint64_t>(i + kWinBlock, off[c + 1])),
                                    static_cast<uint32_t>(c), 0u));
        }
    w.t1 = static_cast<uint32_t>(desc.size());
    plan.push_back(w);
    b = b1;
  }
  return HIVE_OK;
}

hive_status_t free_window(hive_table* t) {
  if (t->win_set) (void)hipFree(t->win_set);
  if (t->win_buf) (void)hipFree(t->win_buf);
  t->win_set = nullptr;
  t->win_buf = nullptr;
  t->win_cap = 0;
  return HIVE_OK;
}

}  // namespace hive
