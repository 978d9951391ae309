/*
 * hive_oracle.c -- sequential CPU restatement of the Hive hash table.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  Nothing in
 * hivehashtable_amd/ links, loads or calls this file.
 *
 * Every function cites the reference lines it restates (paths relative to the
 * SabbirPulok/HiveHashTable tree).  The protocol is the reference's mixed-path
 * insert (hive_insert_one_coop), executed one operation at a time:
 *   Stage A  update in any candidate bucket     include/hive_hash_insert.cuh:911-930
 *   Stage A' update in the canonical stash      include/hive_hash_insert.cuh:932-938
 *   Stage B  claim lowest empty slot of H0      include/hive_hash_insert.cuh:271-310
 *   Stage C  one-level cuckoo, victim kick%8    include/hive_hash_insert.cuh:574-692
 *   Stage D  push to stash[H0] (8 slots)        include/hive_hash_insert.cuh:489-517
 * For distinct keys the insert-only path (hive_insert_one_only, :969-1042)
 * reduces to the same sequence.
 *
 * Stash reads cover slots [0, count): in every state the sequential protocol
 * reaches, slots at or above count are empty, so this equals the reference's
 * scan of all 8 slots (include/hive_hash_lookup.cuh:100-112).
 */
#include "hive_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- hashes */

/* include/hash.hpp:39-47 (Thomas Wang 32-bit mix) */
uint32_t ho_hash1(uint32_t key) {
  key = ~key + (key << 15);
  key = key ^ (key >> 12);
  key = key + (key << 2);
  key = key ^ (key >> 4);
  key = key * 2057u;
  key = key ^ (key >> 16);
  return key;
}

/* include/hash.hpp:64-78; the PTX "lop3.b32 d, d, t, C, 0x96" lines are the
 * three-input XOR d ^ t ^ C (closed form in the comments at :67 and :74). */
uint32_t ho_hash2(uint32_t key) {
  key = (key + 0x7ed55d16u) + (key << 12);
  key = (key ^ 0xc761c23cu) ^ (key >> 19);
  key = (key + 0x165667b1u) + (key << 5);
  key = (key + 0xd3a2646cu) ^ (key << 9);
  key = (key + 0xfd7046c5u) + (key << 3);
  key = (key ^ 0xb55a4f09u) ^ (key >> 16);
  return key;
}

static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

/* The MurmurHash3 32-bit finalizer.  include/hash.hpp:100-104 (cityhash)
 * and :130-134 (murmurhash) both end with these five steps; the published
 * MurmurHash3_x86_32 vectors of the empty string pin it directly, since for
 * len 0 the hash is fmix32(seed). */
uint32_t ho_fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

/* MurmurHash3_x86_32 over an arbitrary byte string (published algorithm);
 * used by tests to pin ho_murmur against the published vectors. */
uint32_t ho_murmur_seeded(const uint8_t* data, size_t len, uint32_t seed) {
  const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
  uint32_t h = seed;
  size_t nblocks = len / 4;
  for (size_t i = 0; i < nblocks; ++i) {
    uint32_t k = (uint32_t)data[4 * i] | ((uint32_t)data[4 * i + 1] << 8) |
                 ((uint32_t)data[4 * i + 2] << 16) | ((uint32_t)data[4 * i + 3] << 24);
    k *= c1;
    k = rotl32(k, 15);
    k *= c2;
    h ^= k;
    h = rotl32(h, 13);
    h = h * 5u + 0xe6546b64u;
  }
  const uint8_t* tail = data + nblocks * 4;
  uint32_t k1 = 0;
  switch (len & 3) {
    case 3: k1 ^= (uint32_t)tail[2] << 16; /* fallthrough */
    case 2: k1 ^= (uint32_t)tail[1] << 8;  /* fallthrough */
    case 1:
      k1 ^= tail[0];
      k1 *= c1;
      k1 = rotl32(k1, 15);
      k1 *= c2;
      h ^= k1;
  }
  h ^= (uint32_t)len;
  return ho_fmix32(h);
}

/* include/hash.hpp:110-138: MurmurHash3_x86_32 of the 4 little-endian key
 * bytes with seed 0. */
uint32_t ho_murmur(uint32_t key) {
  uint32_t k = key * 0xcc9e2d51u;
  k = rotl32(k, 15);
  k *= 0x1b873593u;
  uint32_t h = 0u ^ k;
  h = rotl32(h, 13);
  h = h * 5u + 0xe6546b64u;
  h ^= 4u;
  return ho_fmix32(h);
}

/* include/hash.hpp:80-108.  The constants k0..k3 are declared uint64_t there
 * and combined with uint32_t variables by compound assignment, which
 * truncates to 32 bits after every step; the sister project's copy
 * (UniformHashFunction/src/hash_function.cu:191-219) declares the same
 * constants uint32_t.  Both are 32-bit modular arithmetic: xor by a constant
 * and multiplication by an odd constant (both invertible on u32), then the
 * invertible fmix32, so the whole function is a permutation of u32. */
uint32_t ho_city_premix(uint32_t key) {
  uint32_t k = key;
  k ^= 0x9ae16a3bu;
  k *= 0xc2b2ae35u;
  k ^= 0x1b873593u;
  k *= 0x85ebca6bu;
  uint32_t h = 0u ^ k;
  h *= 0xc2b2ae35u;
  h ^= 0x1b873593u;
  h *= 0x85ebca6bu;
  h ^= 4u;
  return h;
}

uint32_t ho_city(uint32_t key) { return ho_fmix32(ho_city_premix(key)); }

/* include/hash.hpp:143-179: reflected table CRCs with the reference's own
 * polynomials, init all-ones, final xor all-ones, over the 4 LE key bytes. */
static uint64_t crc64_tab[256];
static uint32_t crc32_tab[256];
static int crc_ready = 0;

static void crc_init(void) {
  if (crc_ready) return;
  for (uint32_t i = 0; i < 256; ++i) {
    uint64_t c = i;
    for (int j = 0; j < 8; ++j) c = (c & 1u) ? (c >> 1) ^ 0x42F0E1EBA9EA3693ull : (c >> 1);
    crc64_tab[i] = c;
    uint32_t d = i;
    for (int j = 0; j < 8; ++j) d = (d & 1u) ? (d >> 1) ^ 0x04C11D87u : (d >> 1);
    crc32_tab[i] = d;
  }
  crc_ready = 1;
}

/* include/hash.hpp:181-194 */
uint32_t ho_crc64(uint32_t key) {
  crc_init();
  uint64_t crc = ~0ull;
  for (int i = 0; i < 4; ++i) {
    uint8_t b = (uint8_t)(key >> (8 * i));
    crc = (crc >> 8) ^ crc64_tab[(crc ^ b) & 0xFF];
  }
  return (uint32_t)(crc ^ ~0ull);
}

/* include/hash.hpp:196-207 */
uint32_t ho_crc32(uint32_t key) {
  crc_init();
  uint32_t crc = 0xFFFFFFFFu;
  for (int i = 0; i < 4; ++i) {
    uint8_t b = (uint8_t)(key >> (8 * i));
    crc = (crc >> 8) ^ crc32_tab[(crc ^ b) & 0xFF];
  }
  return crc ^ 0xFFFFFFFFu;
}

/* Bit-serial forms of the same CRCs (no table): a table-free computation that
 * must agree with the table-driven one for every input. */
uint32_t ho_crc64_bitwise(uint32_t key) {
  uint64_t crc = ~0ull;
  for (int i = 0; i < 4; ++i) {
    crc ^= (uint8_t)(key >> (8 * i));
    for (int j = 0; j < 8; ++j) crc = (crc & 1u) ? (crc >> 1) ^ 0x42F0E1EBA9EA3693ull : (crc >> 1);
  }
  return (uint32_t)(crc ^ ~0ull);
}

uint32_t ho_crc32_bitwise(uint32_t key) {
  uint32_t crc = 0xFFFFFFFFu;
  for (int i = 0; i < 4; ++i) {
    crc ^= (uint8_t)(key >> (8 * i));
    for (int j = 0; j < 8; ++j) crc = (crc & 1u) ? (crc >> 1) ^ 0x04C11D87u : (crc >> 1);
  }
  return crc ^ 0xFFFFFFFFu;
}

/* include/HashPolicies.cuh:10-65 */
int ho_num_hashes(int policy) {
  switch (policy) {
    case HO_POLICY_TRIPLE:
    case HO_POLICY_MURMURCITYBIT:
    case HO_POLICY_LOOKUP3: return 3;
    default: return 2;
  }
}

static uint32_t raw_hash(int policy, int i, uint32_t key) {
  switch (policy) {
    case HO_POLICY_DEFAULT2HASH: return i == 0 ? ho_hash1(key) : ho_hash2(key);
    case HO_POLICY_MURMURCITY: return i == 0 ? ho_murmur(key) : ho_city(key);
    case HO_POLICY_TRIPLE: return i == 0 ? ho_hash1(key) : (i == 1 ? ho_hash2(key) : ho_city(key));
    case HO_POLICY_MURMURCITYBIT: return i == 0 ? ho_murmur(key) : (i == 1 ? ho_city(key) : ho_hash1(key));
    case HO_POLICY_LOOKUP2: return i == 0 ? ho_crc64(key) : ho_crc32(key);
    case HO_POLICY_LOOKUP3: return i == 0 ? ho_crc64(key) : (i == 1 ? ho_crc32(key) : ho_hash1(key));
    default: return 0;
  }
}

uint32_t ho_bucket(int policy, int i, uint32_t key, uint32_t nb) { return raw_hash(policy, i, key) % nb; }

/* include/hive_hash_insert.cuh:58-71 */
uint32_t ho_alt_bucket(int policy, uint32_t key, uint32_t cur, uint32_t nb) {
  int nh = ho_num_hashes(policy);
  for (int i = 0; i < nh; ++i)
    if (ho_bucket(policy, i, key, nb) == cur) return ho_bucket(policy, (i + 1) % nh, key, nb);
  return ho_bucket(policy, 0, key, nb);
}

/* ---------------------------------------------------------------- table */

static uint32_t kv_key(uint64_t kv) { return (uint32_t)kv; }          /* :303-306 */
static uint32_t kv_val(uint64_t kv) { return (uint32_t)(kv >> 32); }  /* :308-311 */
static uint64_t kv_pack(uint32_t k, uint32_t v) { return ((uint64_t)v << 32) | k; } /* :298-301 */

ho_table_t* ho_create(uint32_t nb, int policy) { return ho_create_ex(nb, policy, HO_SLOTS); }

/* A table of nb buckets of S slots (8: the reference bucket; 16: the
 * 128-byte bucket of libhive's wide layout).  The protocol is the same with
 * SLOTS = S: Stage B takes the lowest empty of S slots, Stage C tries victim
 * slots kick % S for kick < HO_MAX_KICKS, the stash keeps 8 slots. */
ho_table_t* ho_create_ex(uint32_t nb, int policy, uint32_t S) {
  if (nb == 0 || (S != 8 && S != 16)) return NULL;
  ho_table_t* t = (ho_table_t*)calloc(1, sizeof(ho_table_t));
  if (!t) return NULL;
  t->nb = nb;
  t->S = S;
  t->policy = policy;
  t->slots = (uint64_t*)calloc((size_t)nb * t->S, sizeof(uint64_t));
  t->stash = (uint64_t*)calloc((size_t)nb * HO_STASH_SLOTS, sizeof(uint64_t));
  t->stash_count = (uint32_t*)calloc(nb, sizeof(uint32_t));
  if (!t->slots || !t->stash || !t->stash_count) {
    ho_destroy(t);
    return NULL;
  }
  return t;
}

void ho_destroy(ho_table_t* t) {
  if (!t) return;
  free(t->slots);
  free(t->stash);
  free(t->stash_count);
  free(t);
}

/* src/hive_memory.cu:203-218 resets buckets and masks; the stash is reset too
 * here (the reference leaves it, see SURVEY F-notes on hiveHashTableClear). */
void ho_clear(ho_table_t* t) {
  memset(t->slots, 0, (size_t)t->nb * t->S * sizeof(uint64_t));
  memset(t->stash, 0, (size_t)t->nb * HO_STASH_SLOTS * sizeof(uint64_t));
  memset(t->stash_count, 0, (size_t)t->nb * sizeof(uint32_t));
}

void ho_import(ho_table_t* t, const uint64_t* slots, const uint64_t* stash, const uint32_t* stash_count) {
  memcpy(t->slots, slots, (size_t)t->nb * t->S * sizeof(uint64_t));
  memcpy(t->stash, stash, (size_t)t->nb * HO_STASH_SLOTS * sizeof(uint64_t));
  memcpy(t->stash_count, stash_count, (size_t)t->nb * sizeof(uint32_t));
}

void ho_export(const ho_table_t* t, uint64_t* slots, uint64_t* stash, uint32_t* stash_count) {
  memcpy(slots, t->slots, (size_t)t->nb * t->S * sizeof(uint64_t));
  memcpy(stash, t->stash, (size_t)t->nb * HO_STASH_SLOTS * sizeof(uint64_t));
  memcpy(stash_count, t->stash_count, (size_t)t->nb * sizeof(uint32_t));
}

static uint32_t stash_limit(const ho_table_t* t, uint32_t b) {
  uint32_t c = t->stash_count[b];
  return c > HO_STASH_SLOTS ? HO_STASH_SLOTS : c;
}

/* lowest slot of bucket b holding key, or -1.  include/hive_hash_insert.cuh:96-113
 * (lowest lane, then x before y == lowest slot index). */
static int find_in_bucket(const uint64_t* bucket, uint32_t nslots, uint32_t key) {
  for (int s = 0; s < (int)nslots; ++s)
    if (kv_key(bucket[s]) == key && bucket[s] != 0) return s;
  return -1;
}

static int find_in_stash(const ho_table_t* t, uint32_t b, uint32_t key) {
  uint32_t lim = stash_limit(t, b);
  const uint64_t* sb = t->stash + (size_t)b * HO_STASH_SLOTS;
  for (uint32_t s = 0; s < lim; ++s)
    if (sb[s] != 0 && kv_key(sb[s]) == key) return (int)s;
  return -1;
}

static int lowest_empty(const uint64_t* bucket, uint32_t nslots) {
  for (int s = 0; s < (int)nslots; ++s)
    if (bucket[s] == 0) return s;
  return -1;
}

/* include/hive_hash_insert.cuh:898-966 (with :83-117, :166-218, :234-311,
 * :556-705, :489-517), one operation at a time. */
int ho_insert(ho_table_t* t, uint32_t key, uint32_t value) {
  if (key == 0) return HO_R_INVALID_KEY; /* :907, SENTINEL */
  const uint64_t kv = kv_pack(key, value);
  const int nh = ho_num_hashes(t->policy);
  /* Stage A: every candidate bucket, in policy order */
  for (int i = 0; i < nh; ++i) {
    uint32_t b = ho_bucket(t->policy, i, key, t->nb);
    uint64_t* bk = t->slots + (size_t)b * t->S;
    int s = find_in_bucket(bk, t->S, key);
    if (s >= 0) {
      bk[s] = kv;
      return HO_R_UPDATED;
    }
  }
  const uint32_t c = ho_bucket(t->policy, 0, key, t->nb);
  /* Stage A': canonical stash, gated by its count (freeMask bit 31 + count>0) */
  {
    int s = find_in_stash(t, c, key);
    if (s >= 0) {
      t->stash[(size_t)c * HO_STASH_SLOTS + s] = kv;
      return HO_R_UPDATED;
    }
  }
  uint64_t* cb = t->slots + (size_t)c * t->S;
  /* Stage B: lowest empty slot of the canonical bucket */
  {
    int s = lowest_empty(cb, t->S);
    if (s >= 0) {
      cb[s] = kv;
      return HO_R_INSERTED;
    }
  }
  /* Stage C: kick k tries victim slot k % SLOTS; the victim moves to the
   * lowest free slot of its alternate bucket. */
  for (int kick = 0; kick < HO_MAX_KICKS; ++kick) {
    int vs = kick % t->S;
    uint64_t victim = cb[vs];
    if (victim == 0) continue;
    uint32_t a = ho_alt_bucket(t->policy, kv_key(victim), c, t->nb);
    if (a == c) continue;
    uint64_t* ab = t->slots + (size_t)a * t->S;
    int d = lowest_empty(ab, t->S);
    if (d < 0) continue;
    ab[d] = victim;
    cb[vs] = kv;
    return HO_R_INSERTED;
  }
  /* Stage D: stash[c] at its monotonic push pointer */
  if (t->stash_count[c] < HO_STASH_SLOTS) {
    t->stash[(size_t)c * HO_STASH_SLOTS + t->stash_count[c]] = kv;
    t->stash_count[c] += 1;
    return HO_R_STASHED;
  }
  return HO_R_FAILED;
}

/* include/hive_hash_lookup.cuh:126-155 */
int ho_lookup(const ho_table_t* t, uint32_t key, uint32_t* value) {
  if (value) *value = 0;
  if (key == 0) return HO_R_NOT_FOUND;
  const int nh = ho_num_hashes(t->policy);
  for (int i = 0; i < nh; ++i) {
    uint32_t b = ho_bucket(t->policy, i, key, t->nb);
    const uint64_t* bk = t->slots + (size_t)b * t->S;
    int s = find_in_bucket(bk, t->S, key);
    if (s >= 0) {
      if (value) *value = kv_val(bk[s]);
      return HO_R_FOUND;
    }
  }
  uint32_t c = ho_bucket(t->policy, 0, key, t->nb);
  int s = find_in_stash(t, c, key);
  if (s >= 0) {
    if (value) *value = kv_val(t->stash[(size_t)c * HO_STASH_SLOTS + s]);
    return HO_R_FOUND;
  }
  return HO_R_NOT_FOUND;
}

/* include/hive_hash_delete.cuh:168-194: lowest match in EVERY candidate bucket
 * (no early exit), then the canonical stash; the stash count is unchanged. */
int ho_erase(ho_table_t* t, uint32_t key) {
  if (key == 0) return HO_R_INVALID_KEY;
  const int nh = ho_num_hashes(t->policy);
  int removed = 0;
  for (int i = 0; i < nh; ++i) {
    uint32_t b = ho_bucket(t->policy, i, key, t->nb);
    uint64_t* bk = t->slots + (size_t)b * t->S;
    int s = find_in_bucket(bk, t->S, key);
    if (s >= 0) {
      bk[s] = 0;
      removed++;
    }
  }
  uint32_t c = ho_bucket(t->policy, 0, key, t->nb);
  int s = find_in_stash(t, c, key);
  if (s >= 0) {
    t->stash[(size_t)c * HO_STASH_SLOTS + s] = 0;
    removed++;
  }
  return removed ? HO_R_ERASED : HO_R_NOT_FOUND;
}

void ho_insert_batch(ho_table_t* t, const uint32_t* keys, const uint32_t* vals, size_t n, uint8_t* status) {
  for (size_t i = 0; i < n; ++i) {
    int r = ho_insert(t, keys[i], vals ? vals[i] : keys[i] + 1u);
    if (status) status[i] = (uint8_t)r;
  }
}

void ho_lookup_batch(const ho_table_t* t, const uint32_t* keys, size_t n, uint32_t* vals, uint8_t* status) {
  for (size_t i = 0; i < n; ++i) {
    uint32_t v = 0;
    int r = ho_lookup(t, keys[i], &v);
    if (vals) vals[i] = v;
    if (status) status[i] = (uint8_t)r;
  }
}

void ho_erase_batch(ho_table_t* t, const uint32_t* keys, size_t n, uint8_t* status) {
  for (size_t i = 0; i < n; ++i) {
    int r = ho_erase(t, keys[i]);
    if (status) status[i] = (uint8_t)r;
  }
}

static int mixed_one(ho_table_t* t, uint8_t type, uint32_t key, uint32_t val, uint32_t* out) {
  *out = 0;
  switch (type) {
    case HO_OP_INSERT: return ho_insert(t, key, val);
    case HO_OP_LOOKUP: return ho_lookup(t, key, out);
    case HO_OP_DELETE: return ho_erase(t, key);
    default: return HO_R_NOOP; /* include/hive_kernels.cuh:181-184: NONE/ATOMIC_INC -> 0 */
  }
}

/* One mixed batch under the linearization the device uses: every lookup sees
 * the table as it was before the batch, then deletes apply in input order,
 * then inserts apply in input order.  All operations of a batch are
 * concurrent in the reference's hive_mixed_kernel (include/hive_kernels.cuh:119-192),
 * so this is one of its admissible outcomes. */
void ho_mixed_phased(ho_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                     size_t n, uint32_t* out_vals, uint8_t* status) {
  static const uint8_t order[3] = {HO_OP_LOOKUP, HO_OP_DELETE, HO_OP_INSERT};
  for (size_t i = 0; i < n; ++i) {
    if (types[i] != HO_OP_INSERT && types[i] != HO_OP_LOOKUP && types[i] != HO_OP_DELETE) {
      if (out_vals) out_vals[i] = 0;
      if (status) status[i] = HO_R_NOOP;
    }
  }
  for (int p = 0; p < 3; ++p) {
    for (size_t i = 0; i < n; ++i) {
      if (types[i] != order[p]) continue;
      uint32_t o;
      int r = mixed_one(t, types[i], keys[i], vals ? vals[i] : keys[i] + 1u, &o);
      if (out_vals) out_vals[i] = o;
      if (status) status[i] = (uint8_t)r;
    }
  }
}

/* Ops strictly in input order (the sequential single-op model). */
void ho_mixed_sequential(ho_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                         size_t n, uint32_t* out_vals, uint8_t* status) {
  for (size_t i = 0; i < n; ++i) {
    uint32_t o;
    int r = mixed_one(t, types[i], keys[i], vals ? vals[i] : keys[i] + 1u, &o);
    if (out_vals) out_vals[i] = o;
    if (status) status[i] = (uint8_t)r;
  }
}

/* Resize = rehash every live entry (buckets in index order, slots in order,
 * then every stash bucket) into a table of new_nb buckets with the insert
 * protocol.  The reference's linear-hashing split (include/hive_hash_resize.cuh:44-136)
 * addresses buckets with "% index_mask" while lookups use "% num_buckets",
 * so its layout is not reproduced; the observable contract is: same content
 * multiset, every key found with its value.  Returns 0, or -1 if an entry
 * could not be placed (the table is then left unchanged). */
int ho_resize(ho_table_t* t, uint32_t new_nb) {
  ho_table_t* n = ho_create_ex(new_nb, t->policy, t->S);
  if (!n) return -1;
  for (size_t i = 0; i < (size_t)t->nb * t->S; ++i) {
    uint64_t kv = t->slots[i];
    if (kv == 0) continue;
    int r = ho_insert(n, kv_key(kv), kv_val(kv));
    if (r == HO_R_FAILED) { ho_destroy(n); return -1; }
  }
  for (uint32_t b = 0; b < t->nb; ++b) {
    uint32_t lim = stash_limit(t, b);
    for (uint32_t s = 0; s < lim; ++s) {
      uint64_t kv = t->stash[(size_t)b * HO_STASH_SLOTS + s];
      if (kv == 0) continue;
      int r = ho_insert(n, kv_key(kv), kv_val(kv));
      if (r == HO_R_FAILED) { ho_destroy(n); return -1; }
    }
  }
  free(t->slots);
  free(t->stash);
  free(t->stash_count);
  t->nb = new_nb;
  t->slots = n->slots;
  t->stash = n->stash;
  t->stash_count = n->stash_count;
  free(n);
  return 0;
}

/* include/hive_stash_table.cuh:33-64 + :145-151: flatten live stash entries
 * (holes skipped, not copied), then reset the whole stash. */
size_t ho_drain_stash(ho_table_t* t, uint64_t* out, size_t cap) {
  size_t k = 0;
  for (uint32_t b = 0; b < t->nb; ++b) {
    uint32_t lim = stash_limit(t, b);
    for (uint32_t s = 0; s < lim; ++s) {
      uint64_t kv = t->stash[(size_t)b * HO_STASH_SLOTS + s];
      if (kv != 0) {
        if (out && k < cap) out[k] = kv;
        k++;
      }
    }
  }
  memset(t->stash, 0, (size_t)t->nb * HO_STASH_SLOTS * sizeof(uint64_t));
  memset(t->stash_count, 0, (size_t)t->nb * sizeof(uint32_t));
  return k;
}

size_t ho_dump(const ho_table_t* t, uint64_t* out, size_t cap) {
  size_t k = 0;
  for (size_t i = 0; i < (size_t)t->nb * t->S; ++i)
    if (t->slots[i] != 0) {
      if (out && k < cap) out[k] = t->slots[i];
      k++;
    }
  for (uint32_t b = 0; b < t->nb; ++b) {
    uint32_t lim = stash_limit(t, b);
    for (uint32_t s = 0; s < lim; ++s) {
      uint64_t kv = t->stash[(size_t)b * HO_STASH_SLOTS + s];
      if (kv != 0) {
        if (out && k < cap) out[k] = kv;
        k++;
      }
    }
  }
  return k;
}

uint64_t ho_count_live(const ho_table_t* t) { return (uint64_t)ho_dump(t, NULL, 0); }

uint64_t ho_count_stash(const ho_table_t* t) {
  uint64_t k = 0;
  for (uint32_t b = 0; b < t->nb; ++b) {
    uint32_t lim = stash_limit(t, b);
    for (uint32_t s = 0; s < lim; ++s)
      if (t->stash[(size_t)b * HO_STASH_SLOTS + s] != 0) k++;
  }
  return k;
}

static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* Structural invariants of a quiescent table:
 *  (i)   every live entry has key != 0
 *  (ii)  an entry in bucket b has b among its candidate buckets
 *  (iii) an entry in stash[b] has H0(key) == b
 *  (iv)  stash_count <= 8
 *  (v)   each live key appears exactly once over table + stash
 * Returns 0 if all hold, else the number of the first violated invariant. */
int ho_check_invariants(const ho_table_t* t, char* msg, size_t msglen) {
  const int nh = ho_num_hashes(t->policy);
  size_t live = ho_dump(t, NULL, 0);
  uint32_t* keys = (uint32_t*)malloc((live ? live : 1) * sizeof(uint32_t));
  size_t k = 0;
  int bad = 0;
  for (uint32_t b = 0; b < t->nb && !bad; ++b) {
    for (int s = 0; s < (int)t->S; ++s) {
      uint64_t kv = t->slots[(size_t)b * t->S + s];
      if (kv == 0) continue;
      uint32_t key = kv_key(kv);
      if (key == 0) { bad = 1; if (msg) snprintf(msg, msglen, "bucket %u slot %d: key 0 with value", b, s); break; }
      int ok = 0;
      for (int i = 0; i < nh; ++i) ok |= (ho_bucket(t->policy, i, key, t->nb) == b);
      if (!ok) { bad = 2; if (msg) snprintf(msg, msglen, "bucket %u slot %d: key %u not hashed here", b, s, key); break; }
      keys[k++] = key;
    }
  }
  for (uint32_t b = 0; b < t->nb && !bad; ++b) {
    if (t->stash_count[b] > HO_STASH_SLOTS) { bad = 4; if (msg) snprintf(msg, msglen, "stash %u count %u", b, t->stash_count[b]); break; }
    for (uint32_t s = 0; s < t->stash_count[b]; ++s) {
      uint64_t kv = t->stash[(size_t)b * HO_STASH_SLOTS + s];
      if (kv == 0) continue;
      uint32_t key = kv_key(kv);
      if (key == 0) { bad = 1; if (msg) snprintf(msg, msglen, "stash %u slot %u: key 0", b, s); break; }
      if (ho_bucket(t->policy, 0, key, t->nb) != b) { bad = 3; if (msg) snprintf(msg, msglen, "stash %u slot %u: key %u not canonical", b, s, key); break; }
      keys[k++] = key;
    }
  }
  if (!bad && k > 1) {
    qsort(keys, k, sizeof(uint32_t), cmp_u32);
    for (size_t i = 1; i < k; ++i)
      if (keys[i] == keys[i - 1]) { bad = 5; if (msg) snprintf(msg, msglen, "key %u stored twice", keys[i]); break; }
  }
  free(keys);
  if (!bad && msg && msglen) msg[0] = 0;
  return bad;
}

void ho_lookup_probe_count(const ho_table_t* t, const uint32_t* keys, size_t n,
                           uint64_t* bucket_reads, uint64_t* stash_reads) {
  const int nh = ho_num_hashes(t->policy);
  uint64_t br = 0, sr = 0;
  for (size_t j = 0; j < n; ++j) {
    uint32_t key = keys[j];
    if (key == 0) continue;
    int found = 0;
    uint32_t seen[3];
    for (int i = 0; i < nh && !found; ++i) {
      uint32_t b = ho_bucket(t->policy, i, key, t->nb);
      seen[i] = b;
      int dup = 0;
      for (int j2 = 0; j2 < i; ++j2) dup |= (seen[j2] == b);
      if (dup) continue; /* a bucket equal to an earlier candidate is not re-read */
      br++;
      if (find_in_bucket(t->slots + (size_t)b * t->S, t->S, key) >= 0) found = 1;
    }
    if (!found) {
      uint32_t c = ho_bucket(t->policy, 0, key, t->nb);
      if (t->stash_count[c] > 0) sr++;
    }
  }
  if (bucket_reads) *bucket_reads = br;
  if (stash_reads) *stash_reads = sr;
}
