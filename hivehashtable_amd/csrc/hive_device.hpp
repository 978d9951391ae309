// hive_device.hpp -- device-side primitives of the MI355X Hive hash table:
// hash functions and policies, the table view, coherent loads and CAS, the
// per-bucket lock, and the insert / lookup / erase protocols.
//
// Memory-model notes for gfx950 (8 XCDs, one L2 per XCD):
//  * plain loads may return lines cached in this XCD's L2 before another XCD
//    wrote them during the same kernel; kernel boundaries make all writes
//    visible.  Every fast path therefore reads with plain loads and confirms
//    with CAS (performed at the coherence point, returns the true old value).
//  * slow paths read buckets with 16-byte volatile loads (sc0 sc1) and words
//    with agent-scope relaxed atomic loads (sc1), which bypass stale L2 lines,
//    and write only with atomics or coherent (sc1) stores.
//  * no acquire/release fences at agent scope: they emit buffer_inv/buffer_wbl2
//    over the whole L2.  The lock release waits for the critical section's
//    memory operations with s_waitcnt vmcnt(0) and releases with a coherent
//    store.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hive {

// Bucket width: slots of 8 B per bucket.  8 is the reference bucket
// (include/hive_hash_table_struct.cuh:20-22: 64 bytes of payload in a
// 128-byte struct); 16 fills the 128-byte line a bucket read fetches anyway.
// The width is a property of the policy type (P::kS), so every protocol
// function and kernel is instantiated per width.
constexpr int kRefSlots = 8;
constexpr int kStashSlots = 8;   // include/hive_stash_table.cuh:20
constexpr int kMaxKicks = 8;     // include/hive_hash_table.cuh:21
constexpr int kCounterLanes = 64;
constexpr int kCounterStride = 16;  // u64 words between counter lanes (128 B)

// result codes (include/hive.h)
constexpr uint8_t R_FAILED = 0, R_INSERTED = 1, R_UPDATED = 2, R_STASHED = 3, R_FOUND = 4,
                  R_NOT_FOUND = 5, R_ERASED = 6, R_INVALID_KEY = 7, R_NOOP = 8, R_PENDING = 0xFF;
constexpr int OP_INSERT = 0, OP_LOOKUP = 1, OP_DELETE = 2;
// internal op type of a windowed stream: an op resolved by the window's chain
// kernels; every phase kernel skips it and writes nothing for it
constexpr int OP_DEFERRED = 0xFE;
// per-bucket hint bits in TableView::meta: kHintBits per bucket, of which the
// low kHintBits - 1 are overflow fingerprint classes and the top one is the
// stash bit.  Overflow class bit f of bucket c: a key whose H0 is c and whose
// fingerprint class is f may live in another candidate bucket.  2 bits (one
// class) is the default: at 2^26 slots, 4 and 8 bits skip more H1 reads of
// absent keys but make the hint array 2-4x larger, and measured slower
// (insert to LF 0.9: 5.12 ms at 2 bits, 5.34 at 4, 5.31 at 8; lookups equal
// or slower).
#ifndef HIVE_HINT_BITS
#define HIVE_HINT_BITS 2
#endif
constexpr uint32_t kHintBits = HIVE_HINT_BITS;
static_assert(kHintBits == 2 || kHintBits == 4 || kHintBits == 8, "hint bits per bucket: 2, 4 or 8");
constexpr uint32_t kHintsPerWord = 32 / kHintBits;
constexpr uint32_t kOverflowClasses = kHintBits - 1;
constexpr uint32_t kOverflowMask = (1u << kOverflowClasses) - 1u;  // every overflow class
constexpr uint32_t kStashBit = 1u << kOverflowClasses;            // this bucket's stash holds pushed entries

struct Op16 {  // the reference's Operation record (include/utils.h:18-21)
  int32_t type;
  uint32_t pad;
  unsigned long long key;
};

// Table layouts.
//  Separate arrays (P::kInline false): slots[nb * S]; hint bits, locks and
//    stash counts in arrays of their own (meta packs kHintBits per bucket);
//    stash[nb * 8], one stash bucket per canonical bucket.
//  In-line header (P::kInline true, 8 slots): each bucket owns its 128-byte
//    line -- slots 0..7, then a 16-byte header {hint, lock, stash count, stash
//    ref} and padding -- so a probe's one line request brings the bucket's
//    hint, lock and stash fields with it.  The hint word has kInlineClasses
//    overflow classes and the stash bit.  Stash buckets come from a pool of
//    pool_cap 8-slot buckets taken on a bucket's first push (stash ref = pool
//    index + 1, 0 = none), so the stash occupies what it holds.
struct TableView {
  unsigned long long* slots;        // nb * S (separate) or nb * 16 (in-line header)
  unsigned long long* stash;        // nb * 8 (separate) or pool_cap * 8 (in-line header)
  uint32_t* stash_count;            // nb (separate layout)
  uint32_t* meta;                   // kHintBits per bucket, kHintsPerWord buckets per word (separate layout)
  uint32_t* locks;                  // nb (separate layout)
  unsigned long long* counters;     // [3][kCounterLanes][kCounterStride]: table, stash, slow-path ops
  uint32_t* pool_next;              // pool cursors, kPoolCursorStride apart (in-line header layout)
  uint32_t* holes;                  // nonzero once an erase has emptied a bucket slot (see note_hole)
  uint32_t nb;
  uint32_t S;                       // slots per bucket (8 or 16), for code not instantiated per layout
  uint32_t layout;                  // 0 separate arrays, 1 in-line header
  uint32_t pool_cap;                // stash buckets in the pool (in-line header layout)
};
constexpr uint32_t kLayoutSeparate = 0, kLayoutInline = 1;
// The in-line layout's stash pool is handed out by kPoolCursors cursors, each
// on its own 128-byte line and owning pool_cap / kPoolCursors buckets:
// atomics on one address serialize (about 0.2 G/s), and every bucket's first
// stash push takes a pool bucket.
constexpr uint32_t kPoolCursors = 64;
constexpr uint32_t kPoolCursorStride = 32;  // u32 between cursors
constexpr uint32_t kInlineClasses = 31;             // overflow classes of an in-line hint word
constexpr uint32_t kInlineStashBit = 1u << 31;
constexpr int kInlineStride = 16;                   // u64 per in-line bucket line
constexpr int kInlineHeader = 8;                    // u64 index of the header in the line

// ------------------------------------------------------------------ hashes
// include/hash.hpp:39-47
__host__ __device__ __forceinline__ uint32_t mix_hash1(uint32_t key) {
  key = ~key + (key << 15);
  key = key ^ (key >> 12);
  key = key + (key << 2);
  key = key ^ (key >> 4);
  key = key * 2057u;
  key = key ^ (key >> 16);
  return key;
}

// include/hash.hpp:64-78 (the lop3 0x96 steps are three-input XORs)
__host__ __device__ __forceinline__ uint32_t mix_hash2(uint32_t key) {
  key = (key + 0x7ed55d16u) + (key << 12);
  key = (key ^ 0xc761c23cu) ^ (key >> 19);
  key = (key + 0x165667b1u) + (key << 5);
  key = (key + 0xd3a2646cu) ^ (key << 9);
  key = (key + 0xfd7046c5u) + (key << 3);
  key = (key ^ 0xb55a4f09u) ^ (key >> 16);
  return key;
}

__host__ __device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// include/hash.hpp:110-138: MurmurHash3_x86_32, 4-byte key, seed 0
__host__ __device__ __forceinline__ uint32_t mix_murmur(uint32_t key) {
  uint32_t k = rotl32(key * 0xcc9e2d51u, 15) * 0x1b873593u;
  uint32_t h = rotl32(k, 13) * 5u + 0xe6546b64u;
  return fmix32(h ^ 4u);
}

// include/hash.hpp:80-108 (32-bit arithmetic throughout)
__host__ __device__ __forceinline__ uint32_t mix_city(uint32_t key) {
  uint32_t k = ((key ^ 0x9ae16a3bu) * 0xc2b2ae35u ^ 0x1b873593u) * 0x85ebca6bu;
  uint32_t h = ((k * 0xc2b2ae35u) ^ 0x1b873593u) * 0x85ebca6bu;
  return fmix32(h ^ 4u);
}

// include/hash.hpp:143-207: reflected CRC tables with the reference's polynomials
struct CrcTables {
  unsigned long long t64[256];
  uint32_t t32[256];
};

constexpr CrcTables make_crc_tables() {
  CrcTables c{};
  for (uint32_t i = 0; i < 256; ++i) {
    unsigned long long a = i;
    uint32_t b = i;
    for (int j = 0; j < 8; ++j) {
      a = (a & 1ull) ? (a >> 1) ^ 0x42F0E1EBA9EA3693ull : (a >> 1);
      b = (b & 1u) ? (b >> 1) ^ 0x04C11D87u : (b >> 1);
    }
    c.t64[i] = a;
    c.t32[i] = b;
  }
  return c;
}

__constant__ constexpr CrcTables kCrc = make_crc_tables();

__device__ __forceinline__ uint32_t mix_crc64(uint32_t key) {
  unsigned long long crc = ~0ull;
#pragma unroll
  for (int i = 0; i < 4; ++i) crc = (crc >> 8) ^ kCrc.t64[(crc ^ (key >> (8 * i))) & 0xFF];
  return static_cast<uint32_t>(crc ^ ~0ull);
}

__device__ __forceinline__ uint32_t mix_crc32(uint32_t key) {
  uint32_t crc = 0xFFFFFFFFu;
#pragma unroll
  for (int i = 0; i < 4; ++i) crc = (crc >> 8) ^ kCrc.t32[(crc ^ (key >> (8 * i))) & 0xFF];
  return crc ^ 0xFFFFFFFFu;
}

// ---------------------------------------------------------------- policies
// kCoherentReads = false: table reads outside locked sections are plain
// (cached) loads, valid because a kernel boundary separates every phase.
struct PolicyBase {
  static constexpr bool kCoherentReads = false;
  static constexpr int kS = kRefSlots;  // slots per bucket
  static constexpr bool kInline = false;  // bucket header in the bucket's line
};
// include/HashPolicies.cuh:10-65
struct PolicyDefault2 : PolicyBase {
  static constexpr int NH = 2;
  __device__ static __forceinline__ uint32_t raw(int i, uint32_t k) { return i == 0 ? mix_hash1(k) : mix_hash2(k); }
};
struct PolicyMurmurCity : PolicyBase {
  static constexpr int NH = 2;
  __device__ static __forceinline__ uint32_t raw(int i, uint32_t k) { return i == 0 ? mix_murmur(k) : mix_city(k); }
};
struct PolicyTriple : PolicyBase {
  static constexpr int NH = 3;
  __device__ static __forceinline__ uint32_t raw(int i, uint32_t k) {
    return i == 0 ? mix_hash1(k) : (i == 1 ? mix_hash2(k) : mix_city(k));
  }
};
struct PolicyMurmurCityBit : PolicyBase {
  static constexpr int NH = 3;
  __device__ static __forceinline__ uint32_t raw(int i, uint32_t k) {
    return i == 0 ? mix_murmur(k) : (i == 1 ? mix_city(k) : mix_hash1(k));
  }
};
struct PolicyLookup2 : PolicyBase {
  static constexpr int NH = 2;
  __device__ static __forceinline__ uint32_t raw(int i, uint32_t k) { return i == 0 ? mix_crc64(k) : mix_crc32(k); }
};
struct PolicyLookup3 : PolicyBase {
  static constexpr int NH = 3;
  __device__ static __forceinline__ uint32_t raw(int i, uint32_t k) {
    return i == 0 ? mix_crc64(k) : (i == 1 ? mix_crc32(k) : mix_hash1(k));
  }
};

// The same hashes over 16-slot buckets: one 128-byte line per bucket.
template <class P>
struct Wide16 : P {
  static constexpr int kS = 16;
};

// The same hashes over 8-slot buckets with the in-line header.
template <class P>
struct Inline8 : P {
  static constexpr bool kInline = true;
};

// u64 words from one bucket to the next
template <class P>
__host__ __device__ constexpr int bucket_stride() {
  return P::kInline ? kInlineStride : P::kS;
}

// The same hashes with every table read a device-coherent load: for code that
// runs with no kernel boundary after other XCDs' writes (the persistent stream
// kernel), so that it needs no L2 invalidation before a phase.
template <class P>
struct CoherentReads : P {
  static constexpr bool kCoherentReads = true;
};

template <class P>
__device__ __forceinline__ uint32_t bucket_of(int i, uint32_t key, uint32_t nb) {
  return P::raw(i, key) % nb;
}

// include/hive_hash_insert.cuh:58-71
template <class P>
__device__ __forceinline__ uint32_t alt_bucket(uint32_t key, uint32_t cur, uint32_t nb) {
#pragma unroll
  for (int i = 0; i < P::NH; ++i)
    if (bucket_of<P>(i, key, nb) == cur) return bucket_of<P>((i + 1) % P::NH, key, nb);
  return bucket_of<P>(0, key, nb);
}

// ------------------------------------------------------------------ memory
__device__ __forceinline__ uint32_t kv_key(unsigned long long kv) { return static_cast<uint32_t>(kv); }
__device__ __forceinline__ uint32_t kv_val(unsigned long long kv) { return static_cast<uint32_t>(kv >> 32); }
__device__ __forceinline__ unsigned long long kv_pack(uint32_t k, uint32_t v) {
  return (static_cast<unsigned long long>(v) << 32) | k;
}

__device__ __forceinline__ unsigned long long ld_sync(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_sync32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// returns the value held before the operation (== expect on success)
__device__ __forceinline__ unsigned long long cas64(unsigned long long* p, unsigned long long expect,
                                                    unsigned long long desired) {
  __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expect;
}

// Update a slot read as `cur` (holding kv's key) to kv.  When the CAS loses to
// a concurrent update of the same key, this update is ordered just before
// that one -- overwritten, but applied -- so it succeeds without a retry: hot
// keys updated by thousands of lanes in one batch cost one CAS each.  False
// if the slot no longer holds the key (a cuckoo move replaced it).
__device__ __forceinline__ bool update_slot(unsigned long long* p, unsigned long long cur, unsigned long long kv) {
  const unsigned long long seen = cas64(p, cur, kv);
  return seen == cur || kv_key(seen) == kv_key(kv);
}

// all of this lane's outstanding memory operations have completed
__device__ __forceinline__ void wait_vmem() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // s_waitcnt vmcnt(0)
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ bool try_lock(uint32_t* l) {
  uint32_t z = 0;
  return __hip_atomic_compare_exchange_strong(l, &z, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void unlock(uint32_t* l) {
  wait_vmem();  // the critical section's atomics and coherent stores are performed
  __hip_atomic_store(l, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// plain bucket read (S / 2 16-byte loads: four for 64 bytes, eight for 128)
template <int S>
__device__ __forceinline__ void load_bucket(const unsigned long long* p, unsigned long long (&s)[S]) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
  for (int j = 0; j < S / 2; ++j) {
    uint4 v = q[j];
    s[2 * j] = (static_cast<unsigned long long>(v.y) << 32) | v.x;
    s[2 * j + 1] = (static_cast<unsigned long long>(v.w) << 32) | v.z;
  }
}

// coherent bucket read: 16-byte volatile loads, which compile to
// system-coherent loads (sc0 sc1) that bypass stale L2 lines; each aligned
// 8-byte slot inside a 16-byte access is read whole.
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
template <int S>
__device__ __forceinline__ void load_bucket_sync(const unsigned long long* p, unsigned long long (&s)[S]) {
  const volatile u64x2* q = reinterpret_cast<const volatile u64x2*>(p);
#pragma unroll
  for (int j = 0; j < S / 2; ++j) {
    const u64x2 v = q[j];
    s[2 * j] = v.x;
    s[2 * j + 1] = v.y;
  }
}

// table reads outside locked sections: plain, or coherent under CoherentReads
template <class P, int S>
__device__ __forceinline__ void read_bucket(const unsigned long long* p, unsigned long long (&s)[S]) {
  if constexpr (P::kCoherentReads) load_bucket_sync(p, s);
  else load_bucket(p, s);
}
template <class P>
__device__ __forceinline__ uint32_t read_u32(const uint32_t* p) {
  if constexpr (P::kCoherentReads) return ld_sync32(p);
  else return *p;
}
template <class P>
__device__ __forceinline__ unsigned long long read_u64(const unsigned long long* p) {
  if constexpr (P::kCoherentReads) return ld_sync(p);
  else return *p;
}

template <int S>
__device__ __forceinline__ uint32_t match_mask(const unsigned long long (&s)[S], uint32_t key) {
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) m |= static_cast<uint32_t>(kv_key(s[j]) == key) << j;
  return m;  // key != 0, so an empty slot never matches
}

template <int S>
__device__ __forceinline__ uint32_t empty_mask(const unsigned long long (&s)[S]) {
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) m |= static_cast<uint32_t>(s[j] == 0ull) << j;
  return m;
}

// ------------------------------------------------------------- hint bits
// meta packs kHintBits per bucket (2 MiB for 2^23 buckets at 2 bits, small
// enough to stay L2-resident under the random bucket stream).
//  The overflow class bit ovf_bit(k) of bucket c is set before any key k
//  whose H0 is c is moved out of c by a cuckoo kick.  A key leaves H0 only
//  that way (Stage B places into H0; the kick that moves a victim to a
//  non-primary bucket sets the victim's class bit in its H0), so while key k's
//  class bit is clear no key with k's H0 and class lives in another candidate
//  bucket, and k's probes of H1/H2 can be skipped.  With kHintBits - 1 classes
//  a bucket whose displaced keys are all of other classes no longer costs
//  absent keys the H1 read.  The bits are conservative: deletes never clear
//  them; clear() and resize() rebuild them.
//  kStashBit of c is set when stash[c] receives its first entry (the
//  reference's freeMask bit 31, include/hive_hash_insert.cuh:512-514).
__device__ __forceinline__ uint32_t meta_bits(uint32_t word, uint32_t b) {
  return (word >> ((b % kHintsPerWord) * kHintBits)) & ((1u << kHintBits) - 1u);
}
template <class P>
__device__ __forceinline__ uint32_t meta_word(const TableView& t, uint32_t b) {
  return read_u32<P>(t.meta + b / kHintsPerWord);
}
__device__ __forceinline__ void meta_set(const TableView& t, uint32_t b, uint32_t bits) {
  __hip_atomic_fetch_or(t.meta + b / kHintsPerWord, bits << ((b % kHintsPerWord) * kHintBits), __ATOMIC_RELAXED,
                        __HIP_MEMORY_SCOPE_AGENT);
}
// a key's overflow class bit: a multiply-shift of a mix independent of the
// bucket hashes onto [0, kOverflowClasses)
__device__ __forceinline__ uint32_t ovf_bit(uint32_t key) {
  if constexpr (kOverflowClasses == 1) return 1u;
  const uint32_t x = fmix32(key ^ 0x2545F491u);
  return 1u << static_cast<uint32_t>((static_cast<unsigned long long>(x) * kOverflowClasses) >> 32);
}

template <class P>
__device__ __forceinline__ unsigned long long* bucket_ptr(const TableView& t, uint32_t b) {
  return t.slots + static_cast<size_t>(b) * bucket_stride<P>();
}

// ---- bucket fields through the layout
// Hdr: a bucket's hint bits, stash count and stash ref as one value.  The
// separate layout fills hint only (count is read on demand); the in-line
// layout reads the four header words with one 16-byte load from the line.
struct Hdr {
  uint32_t hint, count, ref;
};
template <class P>
__device__ __forceinline__ uint32_t* header_ptr(const TableView& t, uint32_t b) {
  return reinterpret_cast<uint32_t*>(t.slots + static_cast<size_t>(b) * kInlineStride + kInlineHeader);
}
template <class P>
__device__ __forceinline__ uint32_t stash_bit() {
  return P::kInline ? kInlineStashBit : kStashBit;
}
// a key's overflow class bit: a multiply-shift of a mix independent of the
// bucket hashes onto [0, classes)
template <class P>
__device__ __forceinline__ uint32_t ovf_bit_of(uint32_t key) {
  if constexpr (P::kInline) {
    const uint32_t x = fmix32(key ^ 0x2545F491u);
    return 1u << static_cast<uint32_t>((static_cast<unsigned long long>(x) * kInlineClasses) >> 32);
  } else {
    return ovf_bit(key);
  }
}
template <class P>
__device__ __forceinline__ uint32_t* lock_ptr(const TableView& t, uint32_t c) {
  if constexpr (P::kInline) return header_ptr<P>(t, c) + 1;
  else return t.locks + c;
}
template <class P>
__device__ __forceinline__ uint32_t* count_ptr(const TableView& t, uint32_t c) {
  if constexpr (P::kInline) return header_ptr<P>(t, c) + 2;
  else return t.stash_count + c;
}
// the bucket's header (table reads outside locked sections: plain, or
// coherent under CoherentReads)
template <class P>
__device__ __forceinline__ Hdr read_hdr(const TableView& t, uint32_t b) {
  if constexpr (P::kInline) {
    const uint32_t* h = header_ptr<P>(t, b);
    if constexpr (P::kCoherentReads) {
      const u64x2 v = *reinterpret_cast<const volatile u64x2*>(h);
      return Hdr{static_cast<uint32_t>(v.x), static_cast<uint32_t>(v.y), static_cast<uint32_t>(v.y >> 32)};
    } else {
      const uint4 v = *reinterpret_cast<const uint4*>(h);
      return Hdr{v.x, v.z, v.w};
    }
  } else {
    return Hdr{meta_bits(meta_word<P>(t, b), b), 0u, 0u};
  }
}
// coherent header read (locked sections)
template <class P>
__device__ __forceinline__ Hdr read_hdr_sync(const TableView& t, uint32_t b) {
  if constexpr (P::kInline) {
    const u64x2 v = *reinterpret_cast<const volatile u64x2*>(header_ptr<P>(t, b));
    return Hdr{static_cast<uint32_t>(v.x), static_cast<uint32_t>(v.y), static_cast<uint32_t>(v.y >> 32)};
  } else {
    return Hdr{meta_bits(ld_sync32(t.meta + b / kHintsPerWord), b), 0u, 0u};
  }
}
// Two-hash policies in the in-line layout publish a locked section's hint bits
// with the unlock.  There the hint word of bucket c is written only by holders
// of lock(c): a kick moves a victim out of c either to its H1 (its H0 is c,
// whose bit is set) or back to its own H0 (no bit), and stash pushes to c set
// c's stash bit.  The hint (word 0) and the lock (word 1) share the header's
// first 8 bytes, so one 8-byte store releases the lock and sets the bits: the
// next holder's coherent header read sees them.  Until then, inserts of c's
// displaced or stashed keys cannot go wrong for want of the bits: c is full
// (kicks and pushes happen only to a full c, and slots fill but never empty
// within an insert phase), so their fast path's claim in c fails and they
// take lock(c).  Lookups and erases run in later kernels.  With three hashes a
// kick from c can set the bit of another victim's H0, which that bucket's lock
// holder does not own; those keep the separate atomic.
template <class P>
__host__ __device__ constexpr bool hint_at_unlock() {
  return P::kInline && P::NH == 2;
}
template <class P>
__device__ __forceinline__ void unlock_with_hint(const TableView& t, uint32_t c, uint32_t hint) {
  wait_vmem();  // the critical section's atomics and coherent stores are performed
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(header_ptr<P>(t, c)), static_cast<unsigned long long>(hint),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // hint | lock 0 << 32
}

template <class P>
__device__ __forceinline__ void hint_set(const TableView& t, uint32_t b, uint32_t bits) {
  if constexpr (P::kInline)
    __hip_atomic_fetch_or(header_ptr<P>(t, b), bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    meta_set(t, b, bits);
}
// entries pushed to bucket c's stash (at most 8), plain or coherent
template <class P, bool Sync = false>
__device__ __forceinline__ uint32_t stash_count_of(const TableView& t, uint32_t c, const Hdr& h) {
  uint32_t n;
  if constexpr (P::kInline) n = h.count;
  else if constexpr (Sync) n = ld_sync32(t.stash_count + c);
  else n = read_u32<P>(t.stash_count + c);
  return n > kStashSlots ? kStashSlots : n;
}
// bucket c's stash bucket (in the in-line layout: h.ref != 0, which holds
// whenever the stash bit or a nonzero count was read)
template <class P>
__device__ __forceinline__ unsigned long long* stash_of(const TableView& t, uint32_t c, const Hdr& h) {
  if constexpr (P::kInline) return t.stash + static_cast<size_t>(h.ref - 1u) * kStashSlots;
  else return t.stash + static_cast<size_t>(c) * kStashSlots;
}

// Take a stash bucket from the in-line layout's pool for bucket c: from the
// share of c's cursor (c mod kPoolCursors), or of the next cursors when that
// share is used up.  Returns the pool index + 1, or 0 when the pool is used up.
__device__ __forceinline__ uint32_t pool_take(const TableView& t, uint32_t c) {
  const uint32_t share = t.pool_cap / kPoolCursors;
  for (uint32_t k = 0; k < kPoolCursors; ++k) {
    const uint32_t s = (c + k) % kPoolCursors;
    const uint32_t idx =
        __hip_atomic_fetch_add(t.pool_next + s * kPoolCursorStride, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (idx < share) return s * share + idx + 1u;
  }
  return 0u;
}

template <class P>
__device__ __forceinline__ bool is_dup(const uint32_t (&b)[P::NH], int i) {
  bool dup = false;
#pragma unroll
  for (int j = 0; j < i; ++j) dup |= b[j] == b[i];
  return dup;
}

// ------------------------------------------------------------------ insert
// Fast path of one insert, run by every lane with plain loads:
//   Stage A  (include/hive_hash_insert.cuh:911-930)  update in any candidate
//            (H1/H2 are read only when H0's overflow bit is set)
//   Stage B  (include/hive_hash_insert.cuh:271-310)  claim lowest empty H0 slot
// Returns R_PENDING when the stash must be consulted, the canonical bucket is
// full, or a CAS revealed a concurrent change; the op then takes slow_insert.
// full_hints: the table has the full-hint property (see note_hole).  An H0
// read with an empty slot then had no hint bits, and a claim that succeeds
// finds H0 still not full, so no copy of the key can live elsewhere: the
// hint word is read only for a full H0, in a second round.  Otherwise it is
// read in the same round as the slots.
template <class P>
__device__ __forceinline__ uint8_t fast_insert(const TableView& t, uint32_t key, uint32_t val,
                                               bool full_hints = false) {
  if (key == 0) return R_INVALID_KEY;
  const unsigned long long kv = kv_pack(key, val);
  uint32_t b[P::NH];
#pragma unroll
  for (int i = 0; i < P::NH; ++i) b[i] = bucket_of<P>(i, key, t.nb);
  uint32_t m0, e0, hint = 0;
  unsigned long long cur0 = 0;  // the lowest matching slot's value
  {
    unsigned long long s0[P::kS];
    read_bucket<P>(bucket_ptr<P>(t, b[0]), s0);
    if (!full_hints) hint = read_hdr<P>(t, b[0]).hint;  // the same round (in the in-line layout, the same line)
    m0 = match_mask(s0, key);
    e0 = empty_mask(s0);
#pragma unroll
    for (int q = 0; q < P::kS; ++q)
      if (m0 && q == __builtin_ctz(m0)) cur0 = s0[q];
  }
  if (m0) return update_slot(bucket_ptr<P>(t, b[0]) + __builtin_ctz(m0), cur0, kv) ? R_UPDATED : R_PENDING;
  if (full_hints && e0 == 0u) hint = read_hdr<P>(t, b[0]).hint;
  if (hint & ovf_bit_of<P>(key)) {
#pragma unroll
    for (int i = 1; i < P::NH; ++i) {
      if (is_dup<P>(b, i)) continue;
      unsigned long long s[P::kS];
      read_bucket<P>(bucket_ptr<P>(t, b[i]), s);
      const uint32_t m = match_mask(s, key);
      if (m) {
        const int j = __builtin_ctz(m);
        return update_slot(bucket_ptr<P>(t, b[i]) + j, s[j], kv) ? R_UPDATED : R_PENDING;
      }
    }
  }
  if (hint & stash_bit<P>()) return R_PENDING;  // Stage A' runs under the lock
  uint32_t e = e0;
  unsigned long long* cb = bucket_ptr<P>(t, b[0]);
  while (e) {
    const int j = __builtin_ctz(e);
    const unsigned long long seen = cas64(cb + j, 0ull, kv);
    if (seen == 0ull) return R_INSERTED;
    if (kv_key(seen) == key) return update_slot(cb + j, seen, kv) ? R_UPDATED : R_PENDING;  // it landed first
    e &= e - 1;
  }
  return R_PENDING;
}

// Stage A on one bucket (N slots: a bucket's S, or the 8 of a stash bucket)
// with coherent reads (all N slots requested at once, then scanned), among
// its first nslots slots: 1 updated, 0 absent, -1 raced.
template <int N>
__device__ __forceinline__ int locked_update(unsigned long long* bk, int nslots, uint32_t key,
                                             unsigned long long kv) {
  unsigned long long s[N];
  load_bucket_sync(bk, s);
  uint32_t m = match_mask(s, key) & ((1u << nslots) - 1u);
  if (m) {
    const int j = __builtin_ctz(m);
    return update_slot(bk + j, s[j], kv) ? 1 : -1;
  }
  return 0;
}

// Move the victim held in cb[vs] into slot ad[d] (already claimed with a copy
// of `victim`), replacing it with kv.  If a concurrent update changed the
// victim's value in the source slot, the copy is refreshed to the new value
// and the swap retried, so the destination slot never empties again.
// Returns true when kv is in place.
__device__ __forceinline__ bool finish_move(unsigned long long* src, unsigned long long* dst,
                                            unsigned long long victim, unsigned long long kv) {
  for (int spin = 0; spin < 64; ++spin) {
    const unsigned long long seen = cas64(src, victim, kv);
    if (seen == victim) return true;
    if (seen == 0ull || kv_key(seen) != kv_key(victim)) break;
    // source holds a newer value of the same key: carry it into the copy
    const unsigned long long prev = cas64(dst, victim, seen);
    if (prev != victim) break;
    victim = seen;
  }
  // withdraw the copy (only reachable if the source stopped holding the victim)
  unsigned long long cur = victim;
  for (int spin = 0; spin < 64; ++spin) {
    const unsigned long long seen = cas64(dst, cur, 0ull);
    if (seen == cur || seen == 0ull || kv_key(seen) != kv_key(victim)) break;
    cur = seen;
  }
  return false;
}

// The whole protocol for one key while holding lock(H0(key)):
//   Stage A / A' re-check   include/hive_hash_insert.cuh:522-541, :166-218
//   Stage B re-try          include/hive_hash_insert.cuh:586-614
//   Stage B at H1           not in the reference, whose new keys go to H0:
//                           an empty slot of the key's own H1, before any
//                           kick (8-slot buckets)
//   Stage C cuckoo kicks    include/hive_hash_insert.cuh:574-692
//   Stage D stash push      include/hive_hash_insert.cuh:489-517
// The victim is copied into its alternate bucket (claimed by CAS 0 -> victim)
// before its source slot is CAS-replaced, so it is never absent from both.
// pre_victim / pre_alt: slot 0 of H0 as read before the lock and its
// alternate bucket (~0u if none); that bucket is read in the same round as the
// coherent reload of H0, so the first kick usually needs no extra round trip.
// Under hint_at_unlock<P>() the hint bits the section adds to c are returned
// in `pend` (with the last coherent read of c's hint in `hint_now`) instead of
// being set with an atomic of their own; slow_insert publishes them with the
// unlock.
template <class P>
__device__ uint8_t locked_insert(const TableView& t, uint32_t key, uint32_t val, uint32_t c,
                                 unsigned long long pre_victim, uint32_t pre_alt, uint32_t& pend,
                                 uint32_t& hint_now) {
  constexpr int S = P::kS;
  constexpr int NV = kMaxKicks < S ? kMaxKicks : S;  // the kicks' victims: slots 0 .. NV-1
  const unsigned long long kv = kv_pack(key, val);
  unsigned long long* cb = bucket_ptr<P>(t, c);
  // Buckets are reduced to what the protocol needs as soon as they arrive
  // (match and empty masks, the matching slot, the victims), so that no more
  // than the two buckets of one round are live in registers.
  uint32_t pre_empty = 0;  // empty slots of pre_alt, read with the first coherent round
  bool have_pre = false;
  // 8-slot buckets: the key's own H1 (and its empty slots), read in the same
  // round.  A claim there places the key with one CAS instead of a kick's
  // two (copy, then replace); its hint bit rides on the unlock where
  // hint_at_unlock<P>().  (16-slot buckets keep their registers for the
  // round's two buckets and take the slow path half as often.)
  constexpr bool kOwnH1 = S == 8;
  const uint32_t own_alt = kOwnH1 ? bucket_of<P>(1, key, t.nb) : c;
  uint32_t own_empty = 0;
  for (int attempt = 0; attempt < 256; ++attempt) {
    bool retry = false;
    // one coherent round: the canonical bucket and its hint bits.  Only
    // holders of this lock move keys whose H0 is c out of c, or push to
    // stash[c], so the hint read here is exact for this key.
    uint32_t m, e;
    unsigned long long cur = 0;  // the lowest matching slot's value
    unsigned long long vic[NV];
    Hdr h;
    {
      unsigned long long cs[S];
      load_bucket_sync(cb, cs);
      h = read_hdr_sync<P>(t, c);
      if (attempt == 0 && pre_alt != ~0u) {
        unsigned long long pre[S];
        read_bucket<P>(bucket_ptr<P>(t, pre_alt), pre);
        pre_empty = empty_mask(pre);
        have_pre = true;
      }
      if (kOwnH1 && attempt == 0 && own_alt != c) {
        unsigned long long oa[S];
        read_bucket<P>(bucket_ptr<P>(t, own_alt), oa);
        own_empty = empty_mask(oa);
      }
      m = match_mask(cs, key);
      e = empty_mask(cs);
#pragma unroll
      for (int q = 0; q < S; ++q)
        if (m && q == __builtin_ctz(m)) cur = cs[q];
#pragma unroll
      for (int q = 0; q < NV; ++q) vic[q] = cs[q];
    }
    const uint32_t hint = h.hint;
    hint_now = hint;
    if (m) {
      if (update_slot(cb + __builtin_ctz(m), cur, kv)) return R_UPDATED;
      continue;
    }
    if (hint & ovf_bit_of<P>(key)) {
      uint32_t b[P::NH];
#pragma unroll
      for (int i = 0; i < P::NH; ++i) b[i] = bucket_of<P>(i, key, t.nb);
#pragma unroll
      for (int i = 1; i < P::NH; ++i) {
        if (retry || is_dup<P>(b, i)) continue;
        const int r = locked_update<S>(bucket_ptr<P>(t, b[i]), S, key, kv);
        if (r > 0) return R_UPDATED;
        if (r < 0) retry = true;
      }
    }
    if (retry) continue;
    uint32_t cnt = 0;
    if (hint & stash_bit<P>()) {
      cnt = stash_count_of<P, true>(t, c, h);
      if (cnt) {
        const int r = locked_update<kStashSlots>(stash_of<P>(t, c, h), static_cast<int>(cnt), key, kv);
        if (r > 0) return R_UPDATED;
        if (r < 0) continue;
      }
    }
    while (e) {
      const int j = __builtin_ctz(e);
      const unsigned long long seen = cas64(cb + j, 0ull, kv);
      if (seen == 0ull) return R_INSERTED;
#pragma unroll
      for (int q = 0; q < NV; ++q)
        if (q == j) vic[q] = seen;
      if (kv_key(seen) == key) { retry = true; break; }
      e &= e - 1;
    }
    if (retry) continue;
    if constexpr (kOwnH1) {
      // Stage B at H1.  Slots are claimed only by CAS 0 -> kv and c is full,
      // so c's lock holder is the only one placing this key; the stale plain
      // view of H1 can only overstate its empty slots, and the CAS decides.
      // The key's class bit in c is published before the unlock (with it, or
      // by an atomic the unlock waits for): an insert of the same key finds
      // c full and reads the bit under the lock.
      unsigned long long* ob = bucket_ptr<P>(t, own_alt);
      while (own_empty) {
        const int j = __builtin_ctz(own_empty);
        own_empty &= own_empty - 1;
        if (cas64(ob + j, 0ull, kv) == 0ull) {
          const uint32_t kbit = ovf_bit_of<P>(key);
          if constexpr (hint_at_unlock<P>()) pend |= kbit;
          else if (!(hint & kbit)) hint_set<P>(t, c, kbit);
          return R_INSERTED;
        }
      }
    }
#pragma unroll
    for (int kick = 0; kick < kMaxKicks; ++kick) {
      const int vs = kick % S;
      const unsigned long long victim = vic[vs];
      if (retry || victim == 0ull) continue;
      const uint32_t vkey = kv_key(victim);
      const uint32_t a = alt_bucket<P>(vkey, c, t.nb);
      if (a == c) continue;
      unsigned long long* ab = bucket_ptr<P>(t, a);
      // slots of the alternate bucket are only ever claimed by CAS 0 -> kv, so
      // a stale plain view can only overstate its empty slots; the CAS decides
      uint32_t ae;
      if (kick == 0 && have_pre && victim == pre_victim && a == pre_alt) {
        ae = pre_empty;
      } else {
        unsigned long long as[S];
        read_bucket<P>(ab, as);
        ae = empty_mask(as);
      }
      if (!ae) continue;
      const uint32_t vh0 = bucket_of<P>(0, vkey, t.nb);
      // before the victim can leave its H0 (already set when vh0 == c and the
      // coherent hint read above saw it)
      const uint32_t vbit = ovf_bit_of<P>(vkey);
      if (a != vh0 && !(vh0 == c && ((hint | pend) & vbit))) {
        if constexpr (hint_at_unlock<P>()) pend |= vbit;  // two hashes: a != vh0 means vh0 == c
        else hint_set<P>(t, vh0, vbit);
      }
      int d = -1;
      while (ae) {
        const int j = __builtin_ctz(ae);
        if (cas64(ab + j, 0ull, victim) == 0ull) { d = j; break; }
        ae &= ae - 1;
      }
      if (d < 0) continue;
      if (finish_move(cb + vs, ab + d, victim, kv)) return R_INSERTED;
      retry = true;
    }
    if (retry) continue;
    if (cnt < kStashSlots) {
      if constexpr (P::kInline) {
        if (h.ref == 0u) {  // the bucket's first push takes a stash bucket from the pool
          h.ref = pool_take(t, c);
          if (h.ref == 0u) return R_FAILED;  // pool exhausted
          __hip_atomic_store(header_ptr<P>(t, c) + 3, h.ref, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      __hip_atomic_store(stash_of<P>(t, c, h) + cnt, kv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(count_ptr<P>(t, c), cnt + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if constexpr (hint_at_unlock<P>()) pend |= stash_bit<P>();
      else hint_set<P>(t, c, stash_bit<P>());
      return R_STASHED;
    }
    return R_FAILED;
  }
  return R_FAILED;
}

// Acquire lock(H0) with a try-lock whose critical section completes inside the
// same loop iteration, so lanes of one wavefront that contend for the same
// lock cannot wait on each other (no independent thread scheduling on CDNA).
template <class P>
__device__ uint8_t slow_insert(const TableView& t, uint32_t key, uint32_t val) {
  const uint32_t c = bucket_of<P>(0, key, t.nb);
  // plain read of H0's slot 0 in the same round as the first lock attempt:
  // names the first victim, whose alternate bucket is then fetched with the
  // coherent reload
  const unsigned long long ps0 = read_u64<P>(bucket_ptr<P>(t, c));
  uint8_t result = R_PENDING;
  uint32_t backoff = 1;
  while (result == R_PENDING) {
    if (try_lock(lock_ptr<P>(t, c))) {
      uint32_t pa = ~0u;
      if (ps0 != 0ull) {
        pa = alt_bucket<P>(kv_key(ps0), c, t.nb);
        if (pa == c) pa = ~0u;
      }
      uint32_t pend = 0, hint_now = 0;
      result = locked_insert<P>(t, key, val, c, ps0, pa, pend, hint_now);
      if (hint_at_unlock<P>() && pend) unlock_with_hint<P>(t, c, hint_now | pend);
      else unlock(lock_ptr<P>(t, c));
    } else {
      for (uint32_t i = 0; i < backoff; ++i) __builtin_amdgcn_s_sleep(1);
      backoff = backoff < 64 ? backoff * 2 : 64;
    }
  }
  return result;
}

// ------------------------------------------------------------------ lookup
// include/hive_hash_lookup.cuh:126-155: H0, then the other candidates (only
// if H0's overflow bit is set), then stash[H0] (if its stash bit is set).
// The hint word is read only after an H0 miss.  Returns R_FOUND / R_NOT_FOUND.
// full_hints: the table has the full-hint property (no slot emptied by an
// erase since it was cleared or rebuilt; see note_hole), so a miss in an H0
// with an empty slot is final.
template <class P>
__device__ __forceinline__ uint8_t lookup_one(const TableView& t, uint32_t key, uint32_t& val,
                                              bool full_hints = false) {
  val = 0;
  if (key == 0) return R_NOT_FOUND;
  uint32_t b[P::NH];
#pragma unroll
  for (int i = 0; i < P::NH; ++i) b[i] = bucket_of<P>(i, key, t.nb);
  {
    unsigned long long s[P::kS];
    read_bucket<P>(bucket_ptr<P>(t, b[0]), s);
    const uint32_t m = match_mask(s, key);
    if (m) {
      val = kv_val(s[__builtin_ctz(m)]);
      return R_FOUND;
    }
    if (full_hints && empty_mask(s) != 0u) return R_NOT_FOUND;
  }
  // the hint word is read only after an H0 miss.  In the separate layouts it
  // lives in another line, so requesting it with H0 would cost every lookup
  // one more line request; in the in-line layout requesting it in the same
  // round (from the line the bucket read fetches) measured the same as this
  // dependent read (1.765-1.766 ms against 1.767-1.769 ms for 2^26 lookups)
  const Hdr h = read_hdr<P>(t, b[0]);
  const uint32_t hint = h.hint;
  if (hint & ovf_bit_of<P>(key)) {
#pragma unroll
    for (int i = 1; i < P::NH; ++i) {
      if (is_dup<P>(b, i)) continue;
      unsigned long long s[P::kS];
      read_bucket<P>(bucket_ptr<P>(t, b[i]), s);
      const uint32_t m = match_mask(s, key);
      if (m) {
        val = kv_val(s[__builtin_ctz(m)]);
        return R_FOUND;
      }
    }
  }
  if (hint & stash_bit<P>()) {
    const uint32_t cnt = stash_count_of<P>(t, b[0], h);
    const unsigned long long* sb = stash_of<P>(t, b[0], h);
    for (uint32_t j = 0; j < cnt; ++j) {
      const unsigned long long cur = read_u64<P>(sb + j);
      if (cur != 0ull && kv_key(cur) == key) {
        val = kv_val(cur);
        return R_FOUND;
      }
    }
  }
  return R_NOT_FOUND;
}

// ------------------------------------------------------------------- erase
// include/hive_hash_delete.cuh:168-194: lowest match in every candidate bucket
// (the others only if H0's overflow bit is set), then the canonical stash
// (count unchanged).  removed_tab / removed_stash receive the copies removed.
// remove the copy of key held in slot p (read as cur); a concurrent change to
// another key, or to empty, means there is nothing left to remove here
__device__ __forceinline__ int erase_slot(unsigned long long* p, unsigned long long cur, uint32_t key) {
  for (int spin = 0; spin < 64; ++spin) {
    const unsigned long long seen = cas64(p, cur, 0ull);
    if (seen == cur) return 1;
    if (seen == 0ull || kv_key(seen) != key) return 0;
    cur = seen;
  }
  return 0;
}

// lowest matching slot of a bucket: one 64-byte read, then the CAS
template <class P>
__device__ __forceinline__ int erase_bucket(unsigned long long* bk, uint32_t key) {
  unsigned long long s[P::kS];
  read_bucket<P>(bk, s);
  const uint32_t m = match_mask(s, key);
  if (!m) return 0;
  const int j = __builtin_ctz(m);
  return erase_slot(bk + j, s[j], key);
}

// lowest matching entry among the first cnt stash entries
// Full-hint property.  A bucket's hint bits (an overflow class or the stash
// bit) are set only while the bucket is full: a kick moves a key out of its
// full H0 (or, with three hashes, out of a bucket it reached after its full
// H0), a new key goes to its own H1 or the stash only when its H0 is full.
// Slots are emptied only by erases (a kick's transient copy is withdrawn only
// when its source no longer holds the victim, which takes an erase), so as
// long as no erase has emptied a slot since the table was cleared or rebuilt
// by a rehash, every bucket with a hint bit set is still full.  A lookup
// whose key is not in an H0 with an empty slot then has its answer without
// reading the hint word.  t.holes records the first such erase: zeroed with
// the counters (clear, a new view for a rehash), set by erase_one when it
// empties a bucket slot and by a raw import.  A plain load first, so that
// each L2 stores it about once.
__device__ __forceinline__ void note_hole(const TableView& t) {
  if (*t.holes == 0u) *t.holes = 1u;
}

template <class P>
__device__ __forceinline__ int erase_stash(unsigned long long* sb, uint32_t cnt, uint32_t key) {
  for (uint32_t j = 0; j < cnt; ++j) {
    const unsigned long long cur = read_u64<P>(sb + j);
    if (cur != 0ull && kv_key(cur) == key) return erase_slot(sb + j, cur, key);
  }
  return 0;
}

template <class P>
__device__ __forceinline__ uint8_t erase_one(const TableView& t, uint32_t key, int& removed_tab, int& removed_stash) {
  removed_tab = 0;
  removed_stash = 0;
  if (key == 0) return R_INVALID_KEY;
  uint32_t b[P::NH];
#pragma unroll
  for (int i = 0; i < P::NH; ++i) b[i] = bucket_of<P>(i, key, t.nb);
  const Hdr h = read_hdr<P>(t, b[0]);
  const uint32_t hint = h.hint;
  removed_tab += erase_bucket<P>(bucket_ptr<P>(t, b[0]), key);
  if (hint & ovf_bit_of<P>(key)) {
#pragma unroll
    for (int i = 1; i < P::NH; ++i) removed_tab += erase_bucket<P>(bucket_ptr<P>(t, b[i]), key);
  }
  if (hint & stash_bit<P>()) {
    const uint32_t cnt = stash_count_of<P>(t, b[0], h);
    removed_stash = erase_stash<P>(stash_of<P>(t, b[0], h), cnt, key);
  }
  if (removed_tab) note_hole(t);
  return (removed_tab + removed_stash) ? R_ERASED : R_NOT_FOUND;
}

}  // namespace hive
