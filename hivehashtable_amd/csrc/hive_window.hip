// hive_window.hip -- windowed batches: consecutive small batches of a stream
// (or the requests of one serving batch) run as one window, with the results
// of running them one by one.
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
// inserts.  So it suffices that every op sees, for its key, exactly the
// writes of the batches before its own (and, for an insert, its own batch's
// deletes).  Per window:
//   k_win_mark      every write records its key in a window key set, keeping
//                   the first batch of the window that writes the key (fw);
//   k_win_level     ops of batches up to fw keep their type (before fw the
//                   key is only looked up); later ones are deferred
//                   (OP_DEFERRED, marked in place);
//   phase kernels   the window's lookups, deletes, inserts, as one mixed
//                   batch: for each key, its lookups before fw, then fw's ops
//                   in phase order;
//   chain kernels   the deferred ops of a key form its chain.  Its state
//                   after fw is looked up once; every deferred op's result
//                   follows from that state and the latest deferred write
//                   batch before its own; only the chain's final state is
//                   applied to the table.  No op waits for another.
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
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;
constexpr uint32_t kClaimer = 0x80000000u;  // eidx flag: this op claimed its key's entry (entries < 2^26)
constexpr uint32_t kNoWrite = 0xFFFFu;    // no written batch
constexpr uint32_t kListed = 24;          // written batches kept per chain for the chain lookups
constexpr uint32_t kLine = 32;            // u32 words per control counter (own 128-byte line)

// Window key-set entry: epoch (16) | first batch of the window that writes
// the key (16) | key (32).  Entries of other epochs are free, so the set is
// never cleared between windows.
__device__ __forceinline__ unsigned long long went(uint32_t epoch, uint32_t fw, uint32_t key) {
  return (static_cast<unsigned long long>((epoch << 16) | fw) << 32) | key;
}
__device__ __forceinline__ uint32_t went_epoch(unsigned long long e) { return static_cast<uint32_t>(e >> 48); }
__device__ __forceinline__ uint32_t went_fw(unsigned long long e) { return static_cast<uint32_t>(e >> 32) & 0xFFFFu; }
__device__ __forceinline__ uint32_t win_hash(uint32_t key) { return fmix32(key * 0x9E3779B1u + 0x7F4A7C15u); }

// control block of one window buffer, one 128-byte line per word, reset by
// k_win_mark: [0] nonzero if any op is deferred, [1] blocks of the last chain
// kernel finished
struct WinCtl {
  unsigned int* c;
  __device__ unsigned int* pending() const { return c; }
  __device__ unsigned int* finished() const { return c + 1 * kLine; }
};
constexpr uint32_t kCtlCounters = 2;
// internal types of deferred ops whose write carries their chain's final
// state (set by k_chain_resolve, read by the apply passes)
constexpr uint8_t kApplyDelete = 0xFC, kApplyInsert = 0xFD;

// One block per tile of 256 ops of one batch: desc = {first op, end op,
// batch number, -}; the batch's index in the window is its number - base.
// Block 0 also resets the control block.
// Only writes (inserts, deletes) enter the set: a key's first written batch
// is all that decides deferral, and lookups find their key's entry, if any,
// by a read-only probe in k_win_level.  The writes of one tile (one batch)
// meet in LDS first, so a hot key costs the set one probe per tile; slot
// reads are device-coherent, so the probe finds a claimed entry without a
// CAS.  The one op that claims a key's entry flags its eidx with kClaimer.
constexpr uint32_t kMarkSlots = 2 * kWinBlock;
__global__ __launch_bounds__(kWinBlock) void k_win_mark(const uint4* desc, uint32_t base, const uint8_t* types,
                                                        const uint32_t* keys, uint32_t wlo, unsigned long long* set,
                                                        uint32_t mask, uint32_t epoch, uint32_t* eidx, WinCtl ctl) {
  __shared__ uint32_t s_key[kMarkSlots], s_e[kMarkSlots];  // tile's written keys (0 = free), their entries
  if (blockIdx.x == 0 && threadIdx.x < kCtlCounters)
    __hip_atomic_store(ctl.c + threadIdx.x * kLine, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (uint32_t q = threadIdx.x; q < kMarkSlots; q += kWinBlock) s_key[q] = 0u;
  __syncthreads();
  const uint4 d = desc[blockIdx.x];
  const uint32_t i = d.x + threadIdx.x;
  const int ty = i < d.y ? types[i] : HIVE_OP_NONE;
  const uint32_t key = i < d.y ? keys[i] : 0u;
  const bool write = key != 0u && (ty == OP_INSERT || ty == OP_DELETE);
  // the first write of each key in the tile probes the set for all of them
  uint32_t q = win_hash(key) & (kMarkSlots - 1u);
  bool lead = false;
  if (write)
    for (;;) {  // at most kWinBlock keys in kMarkSlots slots: a free slot always exists
      const uint32_t cur = atomicCAS(s_key + q, 0u, key);
      if (cur == 0u) { lead = true; break; }
      if (cur == key) break;
      q = (q + 1u) & (kMarkSlots - 1u);
    }
  uint32_t e = kNoEntry;
  if (lead) {
    const uint32_t fw = d.z - base, cap = mask + 1u;
    uint32_t h = win_hash(key) & mask;
    for (uint32_t probe = 0; probe < cap && e == kNoEntry; ++probe, h = (h + 1u) & mask) {
      unsigned long long cur = ld_sync(set + h);
      for (;;) {
        if (went_epoch(cur) != epoch) {  // free in this window: claim it
          const unsigned long long seen = cas64(set + h, cur, went(epoch, fw, key));
          if (seen == cur) { e = h | kClaimer; break; }
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
    s_e[q] = e == kNoEntry ? kNoEntry : (e & ~kClaimer);
  }
  __syncthreads();
  if (write && !lead) e = s_e[q];
  if (i < d.y) eidx[i - wlo] = e;
}

// Ops in a batch after their key's first write in the window are deferred:
// typed OP_DEFERRED in place, with their key entry and fw << 16 | batch at
// their window position (no list, so no shared counter); the key is marked
// chained and the window's pending flag raised.  Ops up to and including that
// batch keep their type: before it the key is only looked up, so they see the
// table as it was before the window.
__global__ __launch_bounds__(kWinBlock) void k_win_level(const uint4* desc, uint32_t base, const uint8_t* types,
                                                         const uint32_t* keys, uint32_t wlo,
                                                         const unsigned long long* set, uint32_t mask, uint32_t epoch,
                                                         const uint32_t* eidx, uint8_t* wtype, uint32_t* pbid,
                                                         uint32_t* pe, uint32_t* chained, WinCtl ctl) {
  const uint4 d = desc[blockIdx.x];
  const uint32_t i = d.x + threadIdx.x;
  const uint32_t bw = d.z - base;  // batch index in the window
  bool pend = false;
  if (i < d.y) {
    int ty = types[i];
    uint32_t e = kNoEntry, fw = 0;
    if (ty == OP_LOOKUP) {
      const uint32_t key = keys[i];
      // the key's entry exists only if the window writes the key
      for (uint32_t h = win_hash(key) & mask, probe = 0; key != 0u && probe <= mask; ++probe, h = (h + 1u) & mask) {
        const unsigned long long en = set[h];
        if (went_epoch(en) != epoch) break;  // a free slot ends the probe: no write of the key
        if (static_cast<uint32_t>(en) == key) {
          e = h;
          fw = went_fw(en);
          pend = bw > fw;
          break;
        }
      }
    } else {
      e = eidx[i - wlo];
      if (e != kNoEntry) {
        e &= ~kClaimer;
        fw = went_fw(set[e]);
        pend = bw > fw;
      }
    }
    if (ty == OP_DEFERRED || ty == kApplyDelete || ty == kApplyInsert) ty = HIVE_OP_NONE;  // not input types
    wtype[i - wlo] = pend ? static_cast<uint8_t>(OP_DEFERRED) : static_cast<uint8_t>(ty);
    if (pend) {
      pbid[i - wlo] = (fw << 16) | bw;
      pe[i - wlo] = e;
      chained[e] = epoch;  // every op of the chain stores the same word
    }
  }
  const unsigned long long m = __ballot(pend);
  if (m != 0ull && (threadIdx.x & 63u) == static_cast<uint32_t>(__builtin_ctzll(m)) && ctl.pending()[0] == 0u)
    atomicOr(ctl.pending(), 1u);
}

// ---------------------------------------------------------------- chains
// Chain state.  Every word is tagged with the window's epoch (the listed
// batches are read only below the count of a chain word of this epoch), so
// nothing is cleared between windows.  A record is one access: 32 bytes for
// a pair, a 64-byte line for a key.
// A (key entry, written batch) pair of the window, in a write-batch hash
// table (linear probing, slot h):
struct PairRec {
  unsigned long long key;     // epoch << 48 | e << 12 | batch
  unsigned long long insmax;  // epoch << 32 | (op + 1) of the pair's last insert
  unsigned long long insmin;  // epoch << 32 | ~op of the pair's first insert
  unsigned long long delmin;  // epoch << 32 | ~op of the pair's first delete
};
// A key's chain, by key entry e:
struct KeyRec {
  unsigned long long init;   // epoch << 33 | found << 32 | value: the key after fw
  unsigned long long chain;  // epoch << 48 | written batches << 16 | the last of them
  uint16_t wl[kListed];      // the first kListed written batches, in noting order
};

static_assert(sizeof(PairRec) == 32 && sizeof(KeyRec) == 64, "chain records are 32 and 64 bytes");

struct ChainArgs {
  const uint8_t* types;
  const uint32_t* keys;
  const uint32_t* vals;
  uint32_t* out_vals;
  uint8_t* status;
  uint8_t* wtype;        // the window's internal op types (OP_DEFERRED, kApply*)
  const uint32_t* pbid;  // per window position: fw << 16 | batch of a deferred op
  const uint32_t* pe;    // per window position: key entry of a deferred op
  const uint32_t* eidx;  // per window position: key entry of a write, kClaimer-flagged for the entry's claimer
  const uint32_t* chained;  // per key entry: the window's epoch if the key has deferred ops
  uint32_t wlo, n;       // the window's first op and op count
  PairRec* pr;
  KeyRec* kr;
  uint32_t mask;   // write-batch slots - 1
  uint32_t epoch;
  WinCtl ctl;
  unsigned long long* ts;  // completion stamps of the window's batches
  uint32_t nts;
};

__device__ __forceinline__ uint32_t tag32(unsigned long long w) { return static_cast<uint32_t>(w >> 32); }
__device__ __forceinline__ uint32_t wb_hash(uint32_t e, uint32_t b) {
  return fmix32(e * 0x9E3779B1u + b * 0x85EBCA77u + 0x165667B1u);
}
__device__ __forceinline__ unsigned long long wb_word(uint32_t epoch, uint32_t e, uint32_t b) {
  return (static_cast<unsigned long long>(epoch) << 48) | (static_cast<unsigned long long>(e) << 12) | b;
}
// slot of a (key entry, batch) pair noted in this window, or kNoEntry
__device__ __forceinline__ uint32_t wb_find(const ChainArgs& a, uint32_t e, uint32_t b) {
  const unsigned long long want = wb_word(a.epoch, e, b);
  for (uint32_t h = wb_hash(e, b) & a.mask, probe = 0; probe <= a.mask; ++probe, h = (h + 1u) & a.mask) {
    const unsigned long long cur = a.pr[h].key;
    if (cur == want) return h;
    if (static_cast<uint32_t>(cur >> 48) != a.epoch) return kNoEntry;
  }
  return kNoEntry;
}
__device__ __forceinline__ uint32_t op_value(const ChainArgs& a, uint32_t i) {
  return a.vals ? a.vals[i] : a.keys[i] + 1u;
}

// Note one (key entry e, batch b) pair with its last insert (op + 1, or 0),
// first insert and first delete (op, or ~0u).  The op that adds the pair to
// the write-batch table adds b to the key's chain.  A hot pair is noted by
// many blocks: its words are read coherently first and written only by an
// atomic that can still change them.
__device__ void note_pair(const ChainArgs& a, uint32_t e, uint32_t b, uint32_t imax1, uint32_t imin, uint32_t dmin) {
  const unsigned long long want = wb_word(a.epoch, e, b);
  uint32_t h = wb_hash(e, b) & a.mask;
  bool claimed = false;
  for (uint32_t probe = 0; probe <= a.mask; ++probe, h = (h + 1u) & a.mask) {
    unsigned long long cur = ld_sync(&a.pr[h].key);
    if (cur != want && static_cast<uint32_t>(cur >> 48) != a.epoch) {
      const unsigned long long seen = cas64(&a.pr[h].key, cur, want);
      if (seen == cur) { claimed = true; break; }
      cur = seen;
    }
    if (cur == want) break;
  }
  if (claimed) {  // a new written batch of the chain: count it, keep the latest and the first kListed
    KeyRec& k = a.kr[e];
    unsigned long long cur = ld_sync(&k.chain);
    for (;;) {
      const bool fresh = static_cast<uint32_t>(cur >> 48) == a.epoch;
      const uint32_t n = fresh ? static_cast<uint32_t>(cur >> 16) & 0xFFFFu : 0u;
      const uint32_t last = fresh ? std::max(static_cast<uint32_t>(cur) & 0xFFFFu, b) : b;
      const unsigned long long next = (static_cast<unsigned long long>(a.epoch) << 48) |
                                      (static_cast<unsigned long long>(n + 1u) << 16) | last;
      const unsigned long long seen = cas64(&k.chain, cur, next);
      if (seen == cur) {
        if (n < kListed) k.wl[n] = static_cast<uint16_t>(b);
        break;
      }
      cur = seen;
    }
  }
  // the claimer knows the pair's words are older than the window; anyone
  // else reads them first
  PairRec& r = a.pr[h];
  const unsigned long long tag = static_cast<unsigned long long>(a.epoch) << 32;
  if (imax1 != 0u && (claimed || ld_sync(&r.insmax) < (tag | imax1))) atomicMax(&r.insmax, tag | imax1);
  if (imin != ~0u && (claimed || ld_sync(&r.insmin) < (tag | ~imin))) atomicMax(&r.insmin, tag | ~imin);
  if (dmin != ~0u && (claimed || ld_sync(&r.delmin) < (tag | ~dmin))) atomicMax(&r.delmin, tag | ~dmin);
}

// k_chain_note aggregates the deferred writes of a block's kWinBlock
// consecutive window positions per (key, batch) pair in LDS; the pair's first
// op there notes the aggregate, so a hot pair costs one note per block
// instead of one per op, and notes of different pairs still run in parallel.
constexpr uint32_t kNoteSlots = 2 * kWinBlock;  // at most kWinBlock pairs: a free slot always exists

// Chain pass 1 (reads the table, writes none of it): the claimer of every
// chained key's entry looks the key up (its state after fw); every deferred
// write's (key, batch) pair is noted with the pair's first/last insert and
// first delete.
template <class P>
__global__ __launch_bounds__(kWinBlock) void k_chain_note(TableView t, ChainArgs a) {
  __shared__ unsigned long long s_pair[kNoteSlots];  // (e << 12 | batch) + 1, 0 = free
  __shared__ uint32_t s_imax[kNoteSlots], s_imin[kNoteSlots], s_dmin[kNoteSlots];
  if (a.ctl.pending()[0] == 0u) return;  // the same for every block
  for (uint32_t q = threadIdx.x; q < kNoteSlots; q += kWinBlock) {
    s_pair[q] = 0ull;
    s_imax[q] = 0u;
    s_imin[q] = ~0u;
    s_dmin[q] = ~0u;
  }
  __syncthreads();
  // whole blocks step together (the LDS barriers): every thread runs every step
  for (uint32_t j0 = blockIdx.x * kWinBlock; j0 < a.n; j0 += gridDim.x * kWinBlock) {
    const uint32_t j = j0 + threadIdx.x;
    bool lead = false, any = false;
    uint32_t q = 0, e = 0, b = 0;
    if (j < a.n) {
      const uint32_t ce = a.eidx[j];
      if (ce != kNoEntry && (ce & kClaimer) != 0u && a.chained[ce & ~kClaimer] == a.epoch) {
        uint32_t v;
        const bool found = lookup_one<P>(t, a.keys[a.wlo + j], v) == R_FOUND;
        a.kr[ce & ~kClaimer].init =
            (static_cast<unsigned long long>(a.epoch) << 33) | (static_cast<unsigned long long>(found) << 32) | v;
      }
      const uint32_t i = a.wlo + j;
      const int ty = a.wtype[j] == OP_DEFERRED ? a.types[i] : HIVE_OP_NONE;
      if (ty == OP_INSERT || ty == OP_DELETE) {
        e = a.pe[j];
        b = a.pbid[j] & 0xFFFu;
        const unsigned long long pk = ((static_cast<unsigned long long>(e) << 12) | b) + 1ull;
        q = wb_hash(e, b) & (kNoteSlots - 1u);
        for (;;) {
          const unsigned long long cur = atomicCAS(s_pair + q, 0ull, pk);
          if (cur == 0ull) { lead = true; break; }
          if (cur == pk) break;
          q = (q + 1u) & (kNoteSlots - 1u);
        }
        if (ty == OP_INSERT) {
          atomicMax(s_imax + q, i + 1u);
          atomicMin(s_imin + q, i);
        } else {
          atomicMin(s_dmin + q, i);
        }
        any = true;
      }
    }
    if (!__syncthreads_or(any)) continue;
    uint32_t imax1 = 0u, imin = ~0u, dmin = ~0u;
    if (lead) {  // take the aggregate and free the slot for the next step
      imax1 = s_imax[q];
      imin = s_imin[q];
      dmin = s_dmin[q];
      s_pair[q] = 0ull;
      s_imax[q] = 0u;
      s_imin[q] = ~0u;
      s_dmin[q] = ~0u;
    }
    __syncthreads();
    if (lead) note_pair(a, e, b, imax1, imin, dmin);
  }
}

// Chain pass 2 (no table access): every deferred op's result from the key's
// state after fw and the latest written batch w of its chain before its own
// batch b (state after w: present with w's last insert's value if w inserts,
// absent if it only deletes).  Within batch b, an insert sees b's deletes
// first; the first delete / insert of b (lowest op) is the one that finds
// the key as it was, the later ones find it erased / present.  The op that
// carries the chain's final state (the last insert of the last written
// batch, or its first delete if it only deletes) is re-typed to be applied.
__global__ __launch_bounds__(kWinBlock) void k_chain_resolve(ChainArgs a) {
  if (a.ctl.pending()[0] == 0u) return;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < a.n; j += gridDim.x * blockDim.x) {
    if (a.wtype[j] != OP_DEFERRED) continue;
    const uint32_t i = a.wlo + j, pb = a.pbid[j], b = pb & 0xFFFFu, fw = pb >> 16, e = a.pe[j];
    const int ty = a.types[i];
    const KeyRec kr = a.kr[e];
    bool present = ((kr.init >> 32) & 1ull) != 0ull;
    uint32_t val = static_cast<uint32_t>(kr.init);
    // the chain's latest written batch before b
    const bool fresh = static_cast<uint32_t>(kr.chain >> 48) == a.epoch;
    const uint32_t nw = fresh ? static_cast<uint32_t>(kr.chain >> 16) & 0xFFFFu : 0u;
    uint32_t hw = kNoEntry;
    if (nw <= kListed) {
      uint32_t w = kNoWrite;
      for (uint32_t s = 0; s < kListed; ++s) {
        const uint32_t bb = kr.wl[s];
        if (s < nw && bb < b && (w == kNoWrite || bb > w)) w = bb;
      }
      if (w != kNoWrite) hw = wb_find(a, e, w);
    } else {  // a chain with many written batches: they are dense, walk back
      for (uint32_t bb = b; bb > fw + 1u && hw == kNoEntry;) hw = wb_find(a, e, --bb);
    }
    if (hw != kNoEntry) {
      const unsigned long long im = a.pr[hw].insmax;
      present = tag32(im) == a.epoch;
      val = present ? op_value(a, static_cast<uint32_t>(im) - 1u) : 0u;
    }
    uint8_t res;
    uint32_t out = 0;
    if (ty == OP_LOOKUP) {
      res = present ? R_FOUND : R_NOT_FOUND;
      out = present ? val : 0u;
    } else if (const uint32_t hb = wb_find(a, e, b); hb == kNoEntry) {
      res = R_FAILED;  // unreachable: pass 1 noted every deferred write's pair
    } else {
      const PairRec pr = a.pr[hb];
      const unsigned long long dm = pr.delmin, im = pr.insmax;
      const bool has_del = tag32(dm) == a.epoch, has_ins = tag32(im) == a.epoch;
      const bool last_batch = fresh && (static_cast<uint32_t>(kr.chain) & 0xFFFFu) == b;
      if (ty == OP_DELETE) {
        const bool first = 0xFFFFFFFFu - static_cast<uint32_t>(dm) == i;
        res = first && present ? R_ERASED : R_NOT_FOUND;
        if (last_batch && !has_ins && first) a.wtype[j] = kApplyDelete;
      } else {
        const bool first = 0xFFFFFFFFu - static_cast<uint32_t>(pr.insmin) == i;
        res = first && !(present && !has_del) ? R_INSERTED : R_UPDATED;
        if (last_batch && static_cast<uint32_t>(im) - 1u == i) a.wtype[j] = kApplyInsert;
      }
    }
    if (a.status) a.status[i] = res;
    if (a.out_vals) a.out_vals[i] = out;
  }
}

// one clock reading for every batch of the window (called by a whole block)
__device__ void stamp_window(const ChainArgs& a) {
  __shared__ unsigned long long s_now;
  if (threadIdx.x == 0) s_now = wall_clock64();
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < a.nts; j += blockDim.x)
    __hip_atomic_store(a.ts + j, s_now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Chain pass 3, two launches: apply every chain's final state -- first erase
// the keys whose chain ends deleted (inserts = false), then upsert the final
// value of those whose chain ends with an insert (inserts = true).  A chain
// applies one op, so the ops of one launch are on distinct keys and run
// concurrently.  Erases and inserts stay in separate launches, as in every
// phase sequence of this library: an erase reads a key's hint word and its
// buckets with plain loads and no lock, so a concurrent cuckoo kick of that
// key (hint bit set, copy to the alternate bucket, removal from H0 between
// the erase's two reads) could hide it from the erase.  An insert the chain
// reports INSERTED takes the table's placement (STASHED) or failure.  The
// last block of the insert launch stamps the window's batches.
template <class P>
__global__ __launch_bounds__(kWinBlock) void k_chain_apply(TableView t, ChainArgs a, bool inserts) {
  const bool any = a.ctl.pending()[0] != 0u;
  int dtab = 0, dstash = 0, dslow = 0;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; any && j < a.n; j += gridDim.x * blockDim.x) {
    const uint8_t wt = a.wtype[j];
    if (!inserts) {
      if (wt == kApplyDelete) {
        int rt, rs;
        (void)erase_one<P>(t, a.keys[a.wlo + j], rt, rs);
        dtab -= rt;
        dstash -= rs;
      }
      continue;
    }
    if (wt != kApplyInsert) continue;
    const uint32_t i = a.wlo + j, key = a.keys[i], val = op_value(a, i);
    uint8_t r = fast_insert<P>(t, key, val, *t.holes == 0u);  // the erase pass ran in an earlier launch
    if (r == R_PENDING) {
      r = slow_insert<P>(t, key, val);
      ++dslow;
      dstash += (r == R_STASHED);
    }
    dtab += (r == R_INSERTED);
    if (a.status && (r == R_FAILED || (r == R_STASHED && a.status[i] == R_INSERTED))) a.status[i] = r;
  }
  wave_count(t, dtab, dstash, dslow);
  if (!inserts) return;
  __shared__ uint32_t s_last;
  wait_vmem();
  __syncthreads();
  if (threadIdx.x == 0)
    s_last = __hip_atomic_fetch_add(a.ctl.finished(), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1u;
  __syncthreads();
  if (s_last) stamp_window(a);
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
// win_set (u64, 13 cap): the key set [cap], then the chain records: pairs
// [cap PairRec], keys [cap KeyRec].
// win_buf (u32): per window buffer eidx, pbid, pe [cap / 2 each, by window
// position], chained [cap, by key entry], wtype [cap / 2 bytes], its control
// block on a 128-byte line, and the type lists of its phase kernels
// [partition_words(cap / 2)].  chained is never cleared: every chained key's
// word holds the window's epoch, and an old word that happens to hold it
// only costs a lookup nobody reads.
size_t ctl_offset(size_t cap) {
  const size_t half = cap / 2;
  return (3 * half + cap + (half + 3) / 4 + kLine - 1) / kLine * kLine;
}
size_t part_offset(size_t cap) { return ctl_offset(cap) + kCtlCounters * kLine; }
size_t buffer_words(size_t cap) { return (part_offset(cap) + partition_words(cap / 2) + kLine - 1) / kLine * kLine; }
constexpr size_t kSetWords = 1 + sizeof(PairRec) / 8 + sizeof(KeyRec) / 8;  // u64 words of win_set per key-set slot

struct WindowBuffer {
  uint32_t *eidx, *pbid, *pe, *chained, *part;
  uint8_t* wtype;
  WinCtl ctl;
};

WindowBuffer window_buffer(const hive_table* t, int p) {
  const size_t cap = t->win_cap, half = cap / 2;
  uint32_t* base = t->win_buf + static_cast<size_t>(p) * buffer_words(cap);
  WindowBuffer w;
  w.eidx = base;
  w.pbid = base + half;
  w.pe = base + 2 * half;
  w.chained = base + 3 * half;
  w.wtype = reinterpret_cast<uint8_t*>(w.chained + cap);
  w.ctl = WinCtl{base + ctl_offset(cap)};
  w.part = base + part_offset(cap);
  return w;
}

// the chain records: one set, since the chain passes of consecutive windows
// run in order on the table stream
ChainArgs chain_args(const hive_table* t) {
  const size_t cap = t->win_cap;
  ChainArgs a{};
  a.pr = reinterpret_cast<PairRec*>(t->win_set + cap);
  a.kr = reinterpret_cast<KeyRec*>(t->win_set + 5 * cap);
  a.mask = static_cast<uint32_t>(cap - 1);
  return a;
}
}  // namespace

// key set (2 entries per op, a power of two), chain records, and two window
// buffers (per-op arrays, chain marks, control block) for windows of up to
// `ops` ops
hive_status_t ensure_window(hive_table* t, size_t ops) {
  size_t cap = 1024;
  while (cap < 2 * ops) cap <<= 1;
  if (t->win_set && cap <= t->win_cap) return HIVE_OK;
  if (cap > (size_t(1) << 26)) return fail(HIVE_EINVAL, "window too large");
  const size_t words = 2 * buffer_words(cap);
  unsigned long long* set = nullptr;
  uint32_t* buf = nullptr;
  if (hipMalloc(&set, kSetWords * cap * 8) != hipSuccess) {
    (void)hipGetLastError();
    return fail(HIVE_ENOMEM, "window key set");
  }
  if (hipMalloc(&buf, words * 4) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(set);
    return fail(HIVE_ENOMEM, "window buffers");
  }
  if (hipMemset(set, 0, kSetWords * cap * 8) != hipSuccess || hipMemset(buf, 0, words * 4) != hipSuccess) {
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
  t->win_epoch = 0;  // every tagged word is zero: older than any window
  return HIVE_OK;
}

// One multi-batch window: ops [lo, lo + n), described by ntiles tiles from
// desc (batch numbers base .. base + nbatches - 1), in window buffer p.  The
// key-set passes run on ks, the table and chain passes on s; when they
// differ, s waits for the key-set passes through `ready`, and the caller
// orders ks after the buffer's last use.  ts/nts: completion stamps of the
// window's batches (nts may be 0).
hive_status_t enqueue_window(hive_table* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                             uint64_t lo, uint64_t n, uint32_t* out_vals, uint8_t* status, const uint4* desc,
                             uint32_t ntiles, uint32_t base, unsigned long long* ts, uint32_t nts, int p, int ncu,
                             hipStream_t ks, hipStream_t s, hipEvent_t ready, uint32_t phases) {
  const size_t cap = t->win_cap;
  unsigned long long* set = t->win_set;
  const uint32_t mask = static_cast<uint32_t>(cap - 1);
  const WindowBuffer wb = window_buffer(t, p);
  ChainArgs ca = chain_args(t);
  if (++t->win_epoch > 0xFFFFu) {  // epochs wrapped: free every entry and chain word
    k_fill_u64<<<fill_grid(cap), 256, 0, ks>>>(set, cap, 0ull);
    HIP_CHECK_LAUNCH();
    k_fill_u64<<<fill_grid((kSetWords - 1) * cap), 256, 0, s>>>(set + cap, (kSetWords - 1) * cap, 0ull);
    HIP_CHECK_LAUNCH();
    t->win_epoch = 1;
  }
  k_win_mark<<<ntiles, kWinBlock, 0, ks>>>(desc, base, types, keys, static_cast<uint32_t>(lo), set, mask,
                                           t->win_epoch, wb.eidx, wb.ctl);
  HIP_CHECK_LAUNCH();
  k_win_level<<<ntiles, kWinBlock, 0, ks>>>(desc, base, types, keys, static_cast<uint32_t>(lo), set, mask,
                                            t->win_epoch, wb.eidx, wb.wtype, wb.pbid, wb.pe, wb.chained, wb.ctl);
  HIP_CHECK_LAUNCH();
  // the type lists of the phase kernels are built with the key-set passes,
  // in the window's own buffer, so they too run ahead of the table stream
  const SrcNative src{keys + lo, vals ? vals + lo : nullptr, wb.wtype, OP_DEFERRED};
  const bool listed = (phases & (kPhaseDelete | kPhaseInsert)) && use_partition(t, n, wb.wtype);
  if (listed) {
    hive_status_t rp = partition_batch(src, n, wb.part, ks);
    if (rp != HIVE_OK) return rp;
  }
  if (ks != s) {
    HIP_TRY(hipEventRecord(ready, ks));
    HIP_TRY(hipStreamWaitEvent(s, ready, 0));
  }
  const OutNative out{status ? status + lo : nullptr, out_vals ? out_vals + lo : nullptr};
  const BatchClock none{nullptr, nullptr};
  hive_status_t r = listed ? launch_mixed_listed(t, src, n, out, wb.part, s, none, phases)
                           : launch_mixed(t, src, n, out, s, none, phases);
  if (r != HIVE_OK) return r;
  ca.types = types;
  ca.keys = keys;
  ca.vals = vals;
  ca.out_vals = out_vals;
  ca.status = status;
  ca.wtype = wb.wtype;
  ca.pbid = wb.pbid;
  ca.pe = wb.pe;
  ca.eidx = wb.eidx;
  ca.chained = wb.chained;
  ca.wlo = static_cast<uint32_t>(lo);
  ca.n = static_cast<uint32_t>(n);
  ca.epoch = t->win_epoch;
  ca.ctl = wb.ctl;
  ca.ts = ts;
  ca.nts = nts;
  // grid-stride loops over the window's ops
  const unsigned grid = std::max(1u, std::min(ntiles, static_cast<uint32_t>(std::max(ncu, 1)) * 8u));
  HIVE_POLICY_SWITCH(table_kind(t), (k_chain_note<P><<<grid, kWinBlock, 0, s>>>(t->v, ca)));
  HIP_CHECK_LAUNCH();
  k_chain_resolve<<<grid, kWinBlock, 0, s>>>(ca);
  HIP_CHECK_LAUNCH();
  HIVE_POLICY_SWITCH(table_kind(t), (k_chain_apply<P><<<grid, kWinBlock, 0, s>>>(t->v, ca, false)));
  HIP_CHECK_LAUNCH();
  HIVE_POLICY_SWITCH(table_kind(t), (k_chain_apply<P><<<grid, kWinBlock, 0, s>>>(t->v, ca, true)));
  HIP_CHECK_LAUNCH();
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

// Multi-batch windows alternate between two buffers.  The key-set passes
// (k_win_mark, k_win_level) touch only the inputs, the key set and the window
// buffer, so they run on a side stream one window ahead, overlapping the
// table work of the previous window; a window's phase kernels wait for its
// key-set passes, and a buffer's next key-set passes wait for its last
// window's chain passes.
hive_status_t enqueue_windows(hive_table* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                              const uint64_t* off, uint32_t* out_vals, uint8_t* status, unsigned int* ctrs,
                              unsigned long long* ts, const std::vector<WindowPlan>& plan, const uint4* d_desc,
                              int ncu, hipStream_t s) {
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
                                     p, ncu, pl.ms, s, pl.ready[p]);
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
