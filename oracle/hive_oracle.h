/*
 * hive_oracle.h -- CPU restatement of the Hive hash table's single-operation
 * semantics.  TEST INFRASTRUCTURE ONLY: this library is the checker used by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.
 * The product path (hivehashtable_amd/, libhive.so) never links or calls it.
 *
 * Table state uses exactly the device layout of libhive.so, so a raw device
 * export can be loaded here and queried:
 *   slots[nb * S]       packed kv = (uint64)value << 32 | key, 0 = empty (S = 8 or 16)
 *   stash[nb * 8]       overflow stash, one 8-slot bucket per canonical bucket
 *   stash_count[nb]     monotonic push pointer of each stash bucket
 *
 * Citations are to the reference (SabbirPulok/HiveHashTable) source tree.
 */
#ifndef HIVE_ORACLE_H
#define HIVE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HO_SLOTS 8        /* include/hive_hash_table_struct.cuh:20-22 */
#define HO_STASH_SLOTS 8  /* include/hive_stash_table.cuh:20 */
#define HO_MAX_KICKS 8    /* include/hive_hash_table.cuh:21, MAX_EVICT */

/* Hash policies, include/HashPolicies.cuh:10-65 */
enum {
  HO_POLICY_DEFAULT2HASH = 0,     /* hash1, hash2 */
  HO_POLICY_MURMURCITY = 1,       /* murmur, city */
  HO_POLICY_TRIPLE = 2,           /* hash1, hash2, city */
  HO_POLICY_MURMURCITYBIT = 3,    /* murmur, city, hash1 */
  HO_POLICY_LOOKUP2 = 4,          /* crc64, crc32 */
  HO_POLICY_LOOKUP3 = 5           /* crc64, crc32, hash1 */
};

/* Per-op result codes (same numbering as include/hive.h) */
enum {
  HO_R_FAILED = 0, HO_R_INSERTED = 1, HO_R_UPDATED = 2, HO_R_STASHED = 3,
  HO_R_FOUND = 4, HO_R_NOT_FOUND = 5, HO_R_ERASED = 6, HO_R_INVALID_KEY = 7,
  HO_R_NOOP = 8
};

/* Op types, include/utils.h:10-16 */
enum { HO_OP_INSERT = 0, HO_OP_LOOKUP = 1, HO_OP_DELETE = 2, HO_OP_ATOMIC_INC = 3, HO_OP_NONE = 4 };

typedef struct {
  uint32_t nb;
  uint32_t S; /* slots per bucket: 8 (HO_SLOTS, the reference) or 16 */
  int policy;
  uint64_t* slots;
  uint64_t* stash;
  uint32_t* stash_count;
} ho_table_t;

/* raw 32-bit mixes (before the "% table_size" of include/hash.hpp) */
uint32_t ho_hash1(uint32_t key);
uint32_t ho_hash2(uint32_t key);
uint32_t ho_murmur(uint32_t key);
uint32_t ho_murmur_seeded(const uint8_t* data, size_t len, uint32_t seed);
uint32_t ho_city(uint32_t key);
uint32_t ho_city_premix(uint32_t key);  /* cityhash before its fmix32 finalizer */
uint32_t ho_fmix32(uint32_t h);         /* MurmurHash3 finalizer shared by murmur and city */
/* Number of distinct values of a raw mix over all 2^32 inputs (fn: 0 hash1,
 * 1 hash2, 2 murmur, 3 city), counted in a 512 MiB bitmap by `threads`
 * threads; 2^32 iff the mix is a permutation of u32.  0 on bad arguments. */
uint64_t ho_distinct_outputs(int fn, int threads);
uint32_t ho_crc64(uint32_t key);
uint32_t ho_crc32(uint32_t key);
uint32_t ho_crc64_bitwise(uint32_t key);
uint32_t ho_crc32_bitwise(uint32_t key);
int ho_num_hashes(int policy);
uint32_t ho_bucket(int policy, int i, uint32_t key, uint32_t nb);
uint32_t ho_alt_bucket(int policy, uint32_t key, uint32_t cur, uint32_t nb);

ho_table_t* ho_create(uint32_t nb, int policy);
ho_table_t* ho_create_ex(uint32_t nb, int policy, uint32_t slots_per_bucket);
void ho_destroy(ho_table_t* t);
void ho_clear(ho_table_t* t);
void ho_import(ho_table_t* t, const uint64_t* slots, const uint64_t* stash, const uint32_t* stash_count);
void ho_export(const ho_table_t* t, uint64_t* slots, uint64_t* stash, uint32_t* stash_count);

int ho_insert(ho_table_t* t, uint32_t key, uint32_t value);
int ho_lookup(const ho_table_t* t, uint32_t key, uint32_t* value);
int ho_erase(ho_table_t* t, uint32_t key);

void ho_insert_batch(ho_table_t* t, const uint32_t* keys, const uint32_t* vals, size_t n, uint8_t* status);
void ho_lookup_batch(const ho_table_t* t, const uint32_t* keys, size_t n, uint32_t* vals, uint8_t* status);
void ho_erase_batch(ho_table_t* t, const uint32_t* keys, size_t n, uint8_t* status);
void ho_mixed_phased(ho_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                     size_t n, uint32_t* out_vals, uint8_t* status);
void ho_mixed_sequential(ho_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                         size_t n, uint32_t* out_vals, uint8_t* status);

int ho_resize(ho_table_t* t, uint32_t new_nb);
size_t ho_drain_stash(ho_table_t* t, uint64_t* out, size_t cap);
size_t ho_dump(const ho_table_t* t, uint64_t* out, size_t cap);
uint64_t ho_count_live(const ho_table_t* t);
uint64_t ho_count_stash(const ho_table_t* t);
int ho_check_invariants(const ho_table_t* t, char* msg, size_t msglen);

/* buckets a lookup touches: p = 1 (hit in H0), 2..NH otherwise; stash reads counted apart */
void ho_lookup_probe_count(const ho_table_t* t, const uint32_t* keys, size_t n,
                           uint64_t* bucket_reads, uint64_t* stash_reads);

#ifdef __cplusplus
}
#endif
#endif
