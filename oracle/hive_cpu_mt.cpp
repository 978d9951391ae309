// hive_cpu_mt.cpp -- multithreaded CPU path of the Hive hash table.
//
// TEST INFRASTRUCTURE ONLY: this is the "reference CPU path" that bench.py's
// cpu_baseline leg times beside the GPU, and that tests check against the
// sequential oracle (hive_oracle.c).  The reference has no CPU Hive table, so
// this is a port of its protocol (include/hive_hash_insert.cuh:898-1042,
// include/hive_hash_lookup.cuh:126-155, include/hive_hash_delete.cuh:168-194)
// onto std::thread with GCC __atomic builtins, using the same table layout as
// the oracle and the device library:
//   fast path   relaxed loads of every candidate bucket, update by CAS of the
//               observed kv, claim of the lowest empty H0 slot by CAS 0->kv
//   slow path   per-canonical-bucket spin lock; re-check, cuckoo kicks and the
//               stash push all run under that lock
// Bulk operations of one call run concurrently; a mixed call runs its
// lookups, then its deletes, then its inserts (one admissible linearization
// of the reference's concurrent hive_mixed_kernel, include/hive_kernels.cuh:119-192).
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "hive_oracle.h"

namespace {

constexpr int S = HO_SLOTS;
constexpr int SS = HO_STASH_SLOTS;
constexpr int KICKS = HO_MAX_KICKS;

struct Table {
  uint32_t nb;
  int policy;
  uint64_t* slots;
  uint64_t* stash;
  uint32_t* stash_count;
  uint32_t* locks;
};

inline uint64_t ld(const uint64_t* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
inline uint32_t ld32(const uint32_t* p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
inline uint64_t cas(uint64_t* p, uint64_t expect, uint64_t desired) {
  __atomic_compare_exchange_n(p, &expect, desired, false, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE);
  return expect;  // value seen before the operation
}
inline uint32_t kkey(uint64_t kv) { return static_cast<uint32_t>(kv); }
inline uint64_t pack(uint32_t k, uint32_t v) { return (static_cast<uint64_t>(v) << 32) | k; }

inline void lock(Table& t, uint32_t b) {
  uint32_t spins = 0;
  for (;;) {
    uint32_t z = 0;
    if (__atomic_compare_exchange_n(&t.locks[b], &z, 1u, false, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED)) return;
    if (++spins > 64) std::this_thread::yield();
  }
}
inline void unlock(Table& t, uint32_t b) { __atomic_store_n(&t.locks[b], 0u, __ATOMIC_RELEASE); }

inline int nh(const Table& t) { return ho_num_hashes(t.policy); }
inline uint32_t bucket(const Table& t, int i, uint32_t key) { return ho_bucket(t.policy, i, key, t.nb); }

// Stage A on one bucket: update the lowest slot holding key.  Returns 1 on
// update, 0 if the key is not there, -1 if the slot changed under us.
int try_update(uint64_t* bk, uint32_t key, uint64_t kv) {
  for (int s = 0; s < S; ++s) {
    uint64_t cur = ld(&bk[s]);
    if (cur != 0 && kkey(cur) == key) {
      for (;;) {
        uint64_t seen = cas(&bk[s], cur, kv);
        if (seen == cur) return 1;
        if (seen == 0 || kkey(seen) != key) return -1;
        cur = seen;
      }
    }
  }
  return 0;
}

int slow_insert(Table& t, uint32_t key, uint32_t value) {
  const uint64_t kv = pack(key, value);
  const uint32_t c = bucket(t, 0, key);
  uint64_t* cb = t.slots + static_cast<size_t>(c) * S;
  lock(t, c);
  int result = HO_R_FAILED;
  for (int attempt = 0; attempt < 1024; ++attempt) {
    bool retry = false;
    // Stage A and A' under the lock
    for (int i = 0; i < nh(t) && !retry; ++i) {
      int r = try_update(t.slots + static_cast<size_t>(bucket(t, i, key)) * S, key, kv);
      if (r == 1) { result = HO_R_UPDATED; goto done; }
      if (r < 0) retry = true;
    }
    if (retry) continue;
    {
      uint32_t cnt = std::min<uint32_t>(ld32(&t.stash_count[c]), SS);
      uint64_t* sb = t.stash + static_cast<size_t>(c) * SS;
      for (uint32_t s = 0; s < cnt; ++s) {
        uint64_t cur = ld(&sb[s]);
        if (cur != 0 && kkey(cur) == key) {
          if (cas(&sb[s], cur, kv) == cur) { result = HO_R_UPDATED; goto done; }
          retry = true;
          break;
        }
      }
    }
    if (retry) continue;
    // Stage B (a slot may have freed since the fast path).  An occupied slot
    // is checked for the key too: a concurrent fast-path insert of the same
    // key may have landed after Stage A read the bucket.
    for (int s = 0; s < S; ++s) {
      const uint64_t cur = ld(&cb[s]);
      if (cur != 0) {
        if (kkey(cur) == key) { retry = true; break; }
        continue;
      }
      uint64_t seen = cas(&cb[s], 0, kv);
      if (seen == 0) { result = HO_R_INSERTED; goto done; }
      if (kkey(seen) == key) { retry = true; break; }
    }
    if (retry) continue;
    // Stage C
    for (int kick = 0; kick < KICKS; ++kick) {
      const int vs = kick % S;
      uint64_t victim = ld(&cb[vs]);
      if (victim == 0) continue;
      const uint32_t a = ho_alt_bucket(t.policy, kkey(victim), c, t.nb);
      if (a == c) continue;
      uint64_t* ab = t.slots + static_cast<size_t>(a) * S;
      int d = -1;
      for (int s = 0; s < S; ++s) {
        if (ld(&ab[s]) != 0) continue;
        if (cas(&ab[s], 0, victim) == 0) { d = s; break; }
      }
      if (d < 0) continue;
      // swap kv into the source slot; a concurrent update of the victim is
      // carried into the copy, so the claimed slot never empties again
      for (;;) {
        const uint64_t seen = cas(&cb[vs], victim, kv);
        if (seen == victim) { result = HO_R_INSERTED; goto done; }
        if (seen == 0 || kkey(seen) != kkey(victim)) break;
        if (cas(&ab[d], victim, seen) != victim) break;
        victim = seen;
      }
      // the source stopped holding the victim: withdraw the copy, retry
      for (uint64_t cur = victim;;) {
        uint64_t seen = cas(&ab[d], cur, 0);
        if (seen == cur || seen == 0 || kkey(seen) != kkey(victim)) break;
        cur = seen;
      }
      retry = true;
      break;
    }
    if (retry) continue;
    // Stage D
    {
      uint32_t cnt = ld32(&t.stash_count[c]);
      if (cnt < static_cast<uint32_t>(SS)) {
        __atomic_store_n(&t.stash[static_cast<size_t>(c) * SS + cnt], kv, __ATOMIC_RELAXED);
        __atomic_store_n(&t.stash_count[c], cnt + 1, __ATOMIC_RELEASE);
        result = HO_R_STASHED;
      } else {
        result = HO_R_FAILED;
      }
    }
    break;
  }
done:
  unlock(t, c);
  return result;
}

int insert_one(Table& t, uint32_t key, uint32_t value) {
  if (key == 0) return HO_R_INVALID_KEY;
  const uint64_t kv = pack(key, value);
  const uint32_t c = bucket(t, 0, key);
  for (int i = 0; i < nh(t); ++i) {
    int r = try_update(t.slots + static_cast<size_t>(bucket(t, i, key)) * S, key, kv);
    if (r == 1) return HO_R_UPDATED;
    if (r < 0) return slow_insert(t, key, value);
  }
  if (ld32(&t.stash_count[c]) != 0) return slow_insert(t, key, value);
  uint64_t* cb = t.slots + static_cast<size_t>(c) * S;
  // claim the lowest empty H0 slot; a slot that holds the key now (a
  // concurrent insert of the same key landed after Stage A) sends the op to
  // the locked path, which updates it
  for (int s = 0; s < S; ++s) {
    const uint64_t cur = ld(&cb[s]);
    if (cur != 0) {
      if (kkey(cur) == key) return slow_insert(t, key, value);
      continue;
    }
    uint64_t seen = cas(&cb[s], 0, kv);
    if (seen == 0) return HO_R_INSERTED;
    if (kkey(seen) == key) return slow_insert(t, key, value);
  }
  return slow_insert(t, key, value);
}

int lookup_one(const Table& t, uint32_t key, uint32_t* val) {
  *val = 0;
  if (key == 0) return HO_R_NOT_FOUND;
  uint32_t seen[3];
  for (int i = 0; i < nh(t); ++i) {
    const uint32_t b = bucket(t, i, key);
    seen[i] = b;
    bool dup = false;
    for (int j = 0; j < i; ++j) dup |= seen[j] == b;
    if (dup) continue;
    const uint64_t* bk = t.slots + static_cast<size_t>(b) * S;
    for (int s = 0; s < S; ++s) {
      uint64_t cur = ld(&bk[s]);
      if (cur != 0 && kkey(cur) == key) { *val = static_cast<uint32_t>(cur >> 32); return HO_R_FOUND; }
    }
  }
  const uint32_t c = seen[0];
  uint32_t cnt = std::min<uint32_t>(ld32(&t.stash_count[c]), SS);
  const uint64_t* sb = t.stash + static_cast<size_t>(c) * SS;
  for (uint32_t s = 0; s < cnt; ++s) {
    uint64_t cur = ld(&sb[s]);
    if (cur != 0 && kkey(cur) == key) { *val = static_cast<uint32_t>(cur >> 32); return HO_R_FOUND; }
  }
  return HO_R_NOT_FOUND;
}

int erase_in(uint64_t* bk, int n, uint32_t key) {
  for (int s = 0; s < n; ++s) {
    uint64_t cur = ld(&bk[s]);
    if (cur == 0 || kkey(cur) != key) continue;
    for (;;) {
      uint64_t seen = cas(&bk[s], cur, 0);
      if (seen == cur) return 1;
      if (seen == 0 || kkey(seen) != key) return 0;
      cur = seen;
    }
  }
  return 0;
}

int erase_one(Table& t, uint32_t key) {
  if (key == 0) return HO_R_INVALID_KEY;
  int removed = 0;
  for (int i = 0; i < nh(t); ++i) removed += erase_in(t.slots + static_cast<size_t>(bucket(t, i, key)) * S, S, key);
  const uint32_t c = bucket(t, 0, key);
  uint32_t cnt = std::min<uint32_t>(ld32(&t.stash_count[c]), SS);
  if (cnt) removed += erase_in(t.stash + static_cast<size_t>(c) * SS, static_cast<int>(cnt), key);
  return removed ? HO_R_ERASED : HO_R_NOT_FOUND;
}

template <typename F>
void parallel_for(size_t n, int threads, F f) {
  if (threads <= 1 || n < 4096) {
    f(size_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  size_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    size_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    pool.emplace_back([=] { f(lo, hi); });
  }
  for (auto& th : pool) th.join();
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

struct hcpu_table {
  Table t;
};

hcpu_table* hcpu_create(uint32_t nb, int policy) {
  auto* h = new hcpu_table();
  h->t.nb = nb;
  h->t.policy = policy;
  h->t.slots = static_cast<uint64_t*>(std::calloc(static_cast<size_t>(nb) * S, 8));
  h->t.stash = static_cast<uint64_t*>(std::calloc(static_cast<size_t>(nb) * SS, 8));
  h->t.stash_count = static_cast<uint32_t*>(std::calloc(nb, 4));
  h->t.locks = static_cast<uint32_t*>(std::calloc(nb, 4));
  if (!h->t.slots || !h->t.stash || !h->t.stash_count || !h->t.locks) {
    std::free(h->t.slots); std::free(h->t.stash); std::free(h->t.stash_count); std::free(h->t.locks);
    delete h;
    return nullptr;
  }
  return h;
}

void hcpu_destroy(hcpu_table* h) {
  if (!h) return;
  std::free(h->t.slots); std::free(h->t.stash); std::free(h->t.stash_count); std::free(h->t.locks);
  delete h;
}

double hcpu_clear(hcpu_table* h, int threads) {
  double t0 = now_s();
  Table& t = h->t;
  parallel_for(t.nb, threads, [&](size_t lo, size_t hi) {
    std::memset(t.slots + lo * S, 0, (hi - lo) * S * 8);
    std::memset(t.stash + lo * SS, 0, (hi - lo) * SS * 8);
    std::memset(t.stash_count + lo, 0, (hi - lo) * 4);
  });
  return now_s() - t0;
}

double hcpu_insert(hcpu_table* h, const uint32_t* keys, const uint32_t* vals, size_t n, uint8_t* status, int threads) {
  double t0 = now_s();
  Table& t = h->t;
  parallel_for(n, threads, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      int r = insert_one(t, keys[i], vals ? vals[i] : keys[i] + 1u);
      if (status) status[i] = static_cast<uint8_t>(r);
    }
  });
  return now_s() - t0;
}

double hcpu_lookup(hcpu_table* h, const uint32_t* keys, size_t n, uint32_t* vals, uint8_t* status, int threads) {
  double t0 = now_s();
  const Table& t = h->t;
  parallel_for(n, threads, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      uint32_t v;
      int r = lookup_one(t, keys[i], &v);
      if (vals) vals[i] = v;
      if (status) status[i] = static_cast<uint8_t>(r);
    }
  });
  return now_s() - t0;
}

double hcpu_erase(hcpu_table* h, const uint32_t* keys, size_t n, uint8_t* status, int threads) {
  double t0 = now_s();
  Table& t = h->t;
  parallel_for(n, threads, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      int r = erase_one(t, keys[i]);
      if (status) status[i] = static_cast<uint8_t>(r);
    }
  });
  return now_s() - t0;
}

// lookups, then deletes, then inserts; each phase runs on all threads
double hcpu_mixed(hcpu_table* h, const uint8_t* types, const uint32_t* keys, const uint32_t* vals, size_t n,
                  uint32_t* out_vals, uint8_t* status, int threads) {
  double t0 = now_s();
  Table& t = h->t;
  const uint8_t order[3] = {HO_OP_LOOKUP, HO_OP_DELETE, HO_OP_INSERT};
  parallel_for(n, threads, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i)
      if (types[i] > HO_OP_DELETE) {
        if (out_vals) out_vals[i] = 0;
        if (status) status[i] = HO_R_NOOP;
      }
  });
  for (uint8_t ty : order) {
    parallel_for(n, threads, [&](size_t lo, size_t hi) {
      for (size_t i = lo; i < hi; ++i) {
        if (types[i] != ty) continue;
        uint32_t v = 0;
        int r;
        if (ty == HO_OP_LOOKUP) r = lookup_one(t, keys[i], &v);
        else if (ty == HO_OP_DELETE) r = erase_one(t, keys[i]);
        else r = insert_one(t, keys[i], vals ? vals[i] : keys[i] + 1u);
        if (out_vals) out_vals[i] = v;
        if (status) status[i] = static_cast<uint8_t>(r);
      }
    });
  }
  return now_s() - t0;
}

void hcpu_export(const hcpu_table* h, uint64_t* slots, uint64_t* stash, uint32_t* stash_count) {
  const Table& t = h->t;
  std::memcpy(slots, t.slots, static_cast<size_t>(t.nb) * S * 8);
  std::memcpy(stash, t.stash, static_cast<size_t>(t.nb) * SS * 8);
  std::memcpy(stash_count, t.stash_count, static_cast<size_t>(t.nb) * 4);
}

int hcpu_hardware_threads(void) {
  unsigned n = std::thread::hardware_concurrency();
  return n ? static_cast<int>(n) : 1;
}

// Full-domain image size of a raw mix: thread j evaluates a contiguous share
// of the 2^32 inputs and sets each output's bit in one shared bitmap (atomic
// or), then the set bits are counted.
uint64_t ho_distinct_outputs(int fn, int threads) {
  uint32_t (*mix)(uint32_t) = nullptr;
  switch (fn) {
    case 0: mix = ho_hash1; break;
    case 1: mix = ho_hash2; break;
    case 2: mix = ho_murmur; break;
    case 3: mix = ho_city; break;
    default: return 0;
  }
  if (threads < 1) threads = 1;
  constexpr uint64_t kDomain = uint64_t(1) << 32;
  std::vector<uint64_t> bits(kDomain / 64, 0ull);
  uint64_t* bm = bits.data();
  std::vector<std::thread> pool;
  for (int j = 0; j < threads; ++j) {
    pool.emplace_back([=] {
      const uint64_t lo = kDomain * static_cast<uint64_t>(j) / static_cast<uint64_t>(threads);
      const uint64_t hi = kDomain * static_cast<uint64_t>(j + 1) / static_cast<uint64_t>(threads);
      for (uint64_t x = lo; x < hi; ++x) {
        const uint32_t y = mix(static_cast<uint32_t>(x));
        __atomic_fetch_or(&bm[y >> 6], 1ull << (y & 63u), __ATOMIC_RELAXED);
      }
    });
  }
  for (auto& th : pool) th.join();
  uint64_t n = 0;
  for (uint64_t w : bits) n += static_cast<uint64_t>(__builtin_popcountll(w));
  return n;
}

}  // extern "C"
