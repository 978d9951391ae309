"""Pins the oracle (oracle/hive_oracle.c) -- CPU only.

The reference ships no tests, golden vectors or fixtures for this path, and
cannot be built here (CUDA only).  The oracle is pinned by:
  * published MurmurHash3_x86_32 vectors;
  * independent formulations of each hash (tests/hashref.py);
  * CRC mathematics: table-driven == bit-serial, and CRC affinity;
  * a brute-force map model for found/value/erase semantics;
  * hand-derived protocol cases from the reference lines (lowest slot,
    cuckoo victim rotation and alternate bucket, H0==H1 skip, stash order,
    stash overflow, stash holes, key 0);
  * the golden fixture tests/golden/hash_vectors.json and its generator.
"""
import json
import os
import struct

import numpy as np
import pytest

import hashref

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


@pytest.fixture(scope="module")
def L(O):
    return O.lib()


RNG = np.random.default_rng(12345)
SAMPLE = np.concatenate([
    np.array([0, 1, 2, 3, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF, 0x87654321, 0xDEADBEEF], np.uint32),
    RNG.integers(0, 2**32, size=20000, dtype=np.uint64).astype(np.uint32),
])


# ----------------------------------------------------------------- hashes
MURMUR_VECTORS = [  # published MurmurHash3_x86_32 test vectors (data, seed, hash)
    (b"", 0, 0x00000000),
    (b"", 1, 0x514E28B7),
    (b"", 0xFFFFFFFF, 0x81F16F39),
    (b"\x00\x00\x00\x00", 0, 0x2362F9DE),
    (b"aaaa", 0x9747B28C, 0x5A97808A),
    (b"aaa", 0x9747B28C, 0x283E0130),
    (b"aa", 0x9747B28C, 0x5D211726),
    (b"a", 0x9747B28C, 0x7FA09EA6),
    (b"abcd", 0x9747B28C, 0xF0478627),
    (b"Hello, world!", 0x9747B28C, 0x24884CBA),
    (b"\x21\x43\x65\x87", 0, 0xF55B516B),
    (b"\x21\x43\x65\x87", 0x5082EDEE, 0x2362F9DE),
]


@pytest.mark.parametrize("data,seed,expect", MURMUR_VECTORS)
def test_murmur_published_vectors(L, data, seed, expect):
    assert L.ho_murmur_seeded(data, len(data), seed) == expect


def test_murmur_key_form_matches_general_form(L):
    # include/hash.hpp:110-138 hashes the 4 little-endian key bytes with seed 0
    for k in SAMPLE[:3000].tolist():
        assert L.ho_murmur(k) == L.ho_murmur_seeded(struct.pack("<I", k), 4, 0)
    assert L.ho_murmur(0) == 0x2362F9DE and L.ho_murmur(0x87654321) == 0xF55B516B
    got = np.array([L.ho_murmur(k) for k in SAMPLE.tolist()], np.uint32)
    assert np.array_equal(got, hashref.murmur(SAMPLE))


def test_hash1_matches_shift_add_form(L):
    got = np.array([L.ho_hash1(k) for k in SAMPLE.tolist()], np.uint32)
    assert np.array_equal(got, hashref.hash1(SAMPLE))


def test_hash2_matches_ptx_lop3_sequence(L):
    # lop3 immLut 0x96 is F(0xF0,0xCC,0xAA) = 0xF0^0xCC^0xAA: three-input XOR
    assert 0xF0 ^ 0xCC ^ 0xAA == 0x96
    got = np.array([L.ho_hash2(k) for k in SAMPLE.tolist()], np.uint32)
    assert np.array_equal(got, hashref.hash2(SAMPLE))


def test_hash_mixes_are_bijections_on_a_sample(L):
    # Wang's and Jenkins' 32-bit mixes are permutations of u32: no collisions
    ks = np.arange(1, 1 << 18, dtype=np.uint32) * np.uint32(2654435761)
    for f in (hashref.hash1, hashref.hash2, hashref.murmur):
        h = f(ks)
        assert np.unique(h).size == ks.size


def test_city_matches_cxx_promotion_semantics(L):
    sub = SAMPLE[:3000]
    got = np.array([L.ho_city(k) for k in sub.tolist()], np.uint32)
    assert np.array_equal(got, hashref.city(sub))


# The published MurmurHash3_x86_32 vectors of the empty string are fmix32(seed)
# (no blocks, no tail, length 0), so they pin the finalizer on its own.
EMPTY_STRING_VECTORS = [(s, h) for d, s, h in MURMUR_VECTORS if d == b""]


def test_fmix32_pinned_by_published_vectors(L):
    assert len(EMPTY_STRING_VECTORS) == 3
    for seed, expect in EMPTY_STRING_VECTORS:
        assert L.ho_fmix32(seed) == expect


def test_city_is_fmix32_of_its_premix(L):
    # include/hash.hpp:100-104 ends cityhash with the same five finalizer steps
    # as murmurhash (:130-134), pinned above by the published vectors
    for k in SAMPLE[:3000].tolist():
        assert L.ho_city(k) == L.ho_fmix32(L.ho_city_premix(k))


def _city_mutant(keys, wide_state, wide_finalizer):
    """Two plausible misreadings of include/hash.hpp:80-108, on uint64
    arithmetic: the 64-bit constants carried into a 64-bit state with no
    truncation after the multiplies (wide_state), and the finalizer run on a
    64-bit value (wide_finalizer), truncated to 32 bits only at the end."""
    M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
    M32 = np.uint64(0xFFFFFFFF)
    k = keys.astype(np.uint64)
    with np.errstate(over="ignore"):
        k = k ^ np.uint64(0x9ae16a3b)
        k = k * np.uint64(0xc2b2ae35)
        k = k ^ np.uint64(0x1b873593)
        k = k * np.uint64(0x85ebca6b)
        if not wide_state:
            k &= M32
        h = k * np.uint64(0xc2b2ae35)
        h = h ^ np.uint64(0x1b873593)
        h = h * np.uint64(0x85ebca6b)
        h = h ^ np.uint64(4)
        h = h & (M64 if wide_finalizer else M32)
        h ^= h >> np.uint64(16)
        h = (h * np.uint64(0x85ebca6b)) & (M64 if wide_finalizer else M32)
        h ^= h >> np.uint64(13)
        h = (h * np.uint64(0xc2b2ae35)) & (M64 if wide_finalizer else M32)
        h ^= h >> np.uint64(16)
    return (h & M32).astype(np.uint32)


def test_city_collision_check_rejects_plausible_mistakes(L):
    # the full-domain check below fails a mix that is not a permutation of u32;
    # both misreadings collide already on 2^20 consecutive inputs (about 128
    # collisions expected for a random function), while ho_city does not
    ks = np.arange(1, (1 << 20) + 1, dtype=np.uint32)
    ref = np.array([L.ho_city(k) for k in ks[:2000].tolist()], np.uint32)
    assert np.array_equal(_city_mutant(ks[:2000], False, False), ref)
    # the low 32 bits of xors and products depend only on the low 32 bits of
    # their operands, so a 64-bit state truncated before the finalizer is the
    # same function; what changes the result is a finalizer that sees the
    # high bits (its right shifts bring them down)
    assert np.array_equal(_city_mutant(ks[:2000], True, False), ref)
    for wide_state in (False, True):
        m = _city_mutant(ks, wide_state, True)
        assert np.unique(m).size < ks.size, wide_state


def test_city_is_a_permutation_of_u32(L):
    # every step of cityhash is xor by a constant, multiplication by an odd
    # constant (mod 2^32) or fmix32, all invertible: all 2^32 inputs must give
    # 2^32 distinct outputs
    for c in (0x9ae16a3b, 0xc2b2ae35, 0x1b873593, 0x85ebca6b):
        assert c < 2**32
    for c in (0xc2b2ae35, 0x85ebca6b):
        assert c & 1
    threads = max(1, min(os.cpu_count() or 1, 16))
    assert L.ho_distinct_outputs(3, threads) == 1 << 32


def test_crc_table_equals_bit_serial(L):
    for k in SAMPLE[:5000].tolist():
        assert L.ho_crc64(k) == L.ho_crc64_bitwise(k)
        assert L.ho_crc32(k) == L.ho_crc32_bitwise(k)


def test_crc_is_affine(L):
    # CRC with fixed init/xorout over equal-length inputs: crc(a^b) = crc(a)^crc(b)^crc(0)
    z64, z32 = L.ho_crc64(0), L.ho_crc32(0)
    for a, b in zip(SAMPLE[:2000].tolist(), SAMPLE[2000:4000].tolist()):
        assert L.ho_crc64(a ^ b) == L.ho_crc64(a) ^ L.ho_crc64(b) ^ z64
        assert L.ho_crc32(a ^ b) == L.ho_crc32(a) ^ L.ho_crc32(b) ^ z32


def test_policy_buckets(L, O):
    nb = 1000003
    for k in SAMPLE[:2000].tolist():
        assert O.bucket("Default2Hash", 0, k, nb) == hashref.hash1(np.array([k]))[0] % nb
        assert O.bucket("Default2Hash", 1, k, nb) == hashref.hash2(np.array([k]))[0] % nb
        assert O.bucket("MurmurCityHash", 0, k, nb) == hashref.murmur(np.array([k]))[0] % nb
        assert O.bucket("TripleHash", 2, k, nb) == L.ho_city(k) % nb
        assert O.bucket("MurmurCityBitHash", 2, k, nb) == L.ho_hash1(k) % nb
        assert O.bucket("Lookup2Hash", 0, k, nb) == L.ho_crc64(k) % nb
        assert O.bucket("Lookup3Hash", 2, k, nb) == L.ho_hash1(k) % nb
    assert [L.ho_num_hashes(p) for p in range(6)] == [2, 2, 3, 3, 2, 3]


def test_golden_hash_fixture(L):
    with open(os.path.join(HERE, "golden", "hash_vectors.json")) as f:
        g = json.load(f)
    for row in g["vectors"]:
        k = row["key"]
        assert L.ho_hash1(k) == row["hash1"]
        assert L.ho_hash2(k) == row["hash2"]
        assert L.ho_murmur(k) == row["murmur"]
        assert L.ho_city(k) == row["city"]
        assert L.ho_crc64(k) == row["crc64"]
        assert L.ho_crc32(k) == row["crc32"]


# --------------------------------------------------------------- protocol
def keys_with_h0(policy, nb, target, count, start=1, h1_filter=None):
    """First `count` keys >= start whose H0 == target (optionally filtered on H1)."""
    out = []
    lo = start
    while len(out) < count:
        ks = np.arange(lo, lo + (1 << 20), dtype=np.uint64).astype(np.uint32)
        h0, h1 = hashref.buckets(policy, ks, nb)
        m = h0 == target
        if h1_filter is not None:
            m &= h1_filter(h1)
        out.extend(ks[m].tolist())
        lo += 1 << 20
    return out[:count]


def test_key_zero_rejected(O):
    o = O.Oracle(16)
    assert o.insert([0])[0] == O.R_INVALID_KEY
    v, s = o.lookup([0])
    assert s[0] == O.R_NOT_FOUND and v[0] == 0
    assert o.erase([0])[0] == O.R_INVALID_KEY
    assert o.count_live() == 0


def test_lowest_slot_and_update(O):
    nb = 64
    ks = keys_with_h0("Default2Hash", nb, 5, 8)
    o = O.Oracle(nb)
    assert (o.insert(ks) == O.R_INSERTED).all()
    raw = o.export()
    assert raw.slots.reshape(nb, 8)[5].tolist() == [(k + 1) << 32 | k for k in ks]  # slots 0..7 in order
    assert o.insert([ks[3]], [777])[0] == O.R_UPDATED
    v, s = o.lookup([ks[3]])
    assert s[0] == O.R_FOUND and v[0] == 777


def test_cuckoo_first_kick_moves_slot0_victim(O):
    # include/hive_hash_insert.cuh:616-692: kick 0 evicts slot 0 into the lowest
    # free slot of its alternate bucket; the new key takes slot 0.
    nb = 64
    c = 9
    ks = keys_with_h0("Default2Hash", nb, c, 9, h1_filter=lambda h1: h1 != c)
    o = O.Oracle(nb)
    st = o.insert(ks)
    assert (st == O.R_INSERTED).all()
    raw = o.export().slots.reshape(nb, 8)
    _, alt = hashref.buckets("Default2Hash", np.array([ks[0]], np.uint32), nb)
    assert raw[c, 0] & 0xFFFFFFFF == ks[8]
    assert ks[0] in (raw[int(alt[0])] & 0xFFFFFFFF).tolist()
    assert (raw[c, 1:] & 0xFFFFFFFF).tolist() == ks[1:8]


def test_cuckoo_skips_victim_whose_alt_is_the_same_bucket(O):
    # include/hive_hash_insert.cuh:642 (alt_bucket == bucket_idx): kick skipped
    nb = 16
    c = 3
    same = keys_with_h0("Default2Hash", nb, c, 1, h1_filter=lambda h1: h1 == c)
    other = keys_with_h0("Default2Hash", nb, c, 8, start=same[0] + 1, h1_filter=lambda h1: h1 != c)
    o = O.Oracle(nb)
    o.insert([same[0]] + other[:7])       # slot 0 holds a key with H0 == H1 == c
    assert o.insert([other[7]])[0] == O.R_INSERTED
    raw = o.export().slots.reshape(nb, 8)
    assert raw[c, 0] & 0xFFFFFFFF == same[0]          # not evicted
    assert raw[c, 1] & 0xFFFFFFFF == other[7]         # kick 1 evicted slot 1


def test_stash_order_overflow_and_holes(O):
    # every key maps H0 = H1 = c: no kick can move anything, so keys 9.. go
    # to stash[c] in order; the 17th fails (include/hive_hash_insert.cuh:489-517)
    nb = 8
    c = 2
    ks = keys_with_h0("Default2Hash", nb, c, 17, h1_filter=lambda h1: h1 == c)
    o = O.Oracle(nb)
    st = o.insert(ks)
    assert (st[:8] == O.R_INSERTED).all()
    assert (st[8:16] == O.R_STASHED).all()
    assert st[16] == O.R_FAILED
    raw = o.export()
    assert raw.stash_count[c] == 8
    assert (raw.stash.reshape(nb, 8)[c] & 0xFFFFFFFF).tolist() == ks[8:16]
    v, s = o.lookup(ks)
    assert (s[:16] == O.R_FOUND).all() and s[16] == O.R_NOT_FOUND
    assert (v[:16] == np.array(ks[:16], np.uint32) + 1).all()
    # update in the stash (Stage A', include/hive_hash_insert.cuh:932-938)
    assert o.insert([ks[10]], [5])[0] == O.R_UPDATED
    assert o.lookup([ks[10]])[0][0] == 5
    # delete from the stash leaves count unchanged; the hole is not reused
    assert o.erase([ks[9]])[0] == O.R_ERASED
    assert o.export().stash_count[c] == 8
    assert o.insert([ks[16]])[0] == O.R_FAILED
    r, msg = o.check_invariants()
    assert r == 0, msg


def test_erase_removes_lowest_copy_in_every_candidate(O):
    # include/hive_hash_delete.cuh:158-194: no early exit across buckets
    nb = 64
    k = keys_with_h0("Default2Hash", nb, 7, 1, h1_filter=lambda h1: h1 != 7)[0]
    _, h1 = hashref.buckets("Default2Hash", np.array([k], np.uint32), nb)
    raw = O.RawState(nb)
    s = raw.slots.reshape(nb, 8)
    s[7, 2] = (10 << 32) | k
    s[7, 5] = (11 << 32) | k
    s[int(h1[0]), 0] = (12 << 32) | k
    o = O.Oracle(nb)
    o.load(raw)
    v, st = o.lookup([k])
    assert st[0] == O.R_FOUND and v[0] == 10          # H0 first, lowest slot
    assert o.erase([k])[0] == O.R_ERASED
    after = o.export().slots.reshape(nb, 8)
    assert after[7, 2] == 0 and after[7, 5] != 0 and after[int(h1[0]), 0] == 0
    assert o.lookup([k])[0][0] == 11


def test_map_model_brute_force(O):
    # found / value / erase semantics equal a plain map while nothing fails
    rng = np.random.default_rng(7)
    for policy in ("Default2Hash", "MurmurCityHash", "TripleHash", "Lookup2Hash"):
        o = O.Oracle(128, policy)
        model = {}
        universe = rng.integers(1, 5000, size=400).astype(np.uint32)
        for step in range(3000):
            k = int(universe[rng.integers(0, universe.size)])
            op = rng.integers(0, 3)
            if op == 0:
                v = int(rng.integers(0, 2**32))
                r = o.insert([k], [v])[0]
                assert r != O.R_FAILED
                assert r == (O.R_UPDATED if k in model else r)
                model[k] = v
            elif op == 1:
                vals, st = o.lookup([k])
                if k in model:
                    assert st[0] == O.R_FOUND and vals[0] == model[k]
                else:
                    assert st[0] == O.R_NOT_FOUND
            else:
                r = o.erase([k])[0]
                assert r == (O.R_ERASED if k in model else O.R_NOT_FOUND)
                model.pop(k, None)
        r, msg = o.check_invariants()
        assert r == 0, msg
        assert o.count_live() == len(model)


def test_invariant_checker_catches_planted_faults(O):
    nb = 32
    ks = np.arange(1, 200, dtype=np.uint32)
    o = O.Oracle(nb)
    o.insert(ks)
    good = o.export()
    assert o.check_invariants()[0] == 0
    bad = O.RawState(nb, good.slots.copy(), good.stash.copy(), good.stash_count.copy())
    i = int(np.flatnonzero(bad.slots)[0])
    j = int(np.flatnonzero(bad.slots == 0)[0]) if (bad.slots == 0).any() else None
    # (2) an entry in a bucket that is not one of its candidates
    key = int(bad.slots[i] & 0xFFFFFFFF)
    h0, h1 = hashref.buckets("Default2Hash", np.array([key], np.uint32), nb)
    wrong = next(b for b in range(nb) if b not in (int(h0[0]), int(h1[0])))
    bad.slots.reshape(nb, 8)[wrong, 0] = bad.slots[i]
    o.load(bad)
    assert o.check_invariants()[0] in (2, 5)
    # (5) a key stored twice
    if j is not None:
        dup = O.RawState(nb, good.slots.copy(), good.stash.copy(), good.stash_count.copy())
        b = int(h0[0])
        row = dup.slots.reshape(nb, 8)[b]
        if (row == 0).any():
            row[int(np.flatnonzero(row == 0)[0])] = good.slots[i]
            o.load(dup)
            assert o.check_invariants()[0] == 5
    # (4) stash count above 8
    st = O.RawState(nb, good.slots.copy(), good.stash.copy(), good.stash_count.copy())
    st.stash_count[0] = 9
    o.load(st)
    assert o.check_invariants()[0] == 4


def test_mixed_phased_is_lookups_then_deletes_then_inserts(O):
    o = O.Oracle(64)
    o.insert([10, 20, 30])
    types = np.array([O.OP_INSERT, O.OP_LOOKUP, O.OP_DELETE, O.OP_LOOKUP, O.OP_NONE, O.OP_ATOMIC_INC], np.uint8)
    keys = np.array([40, 40, 20, 20, 5, 5], np.uint32)
    out, st = o.mixed(types, keys)
    assert st.tolist() == [O.R_INSERTED, O.R_NOT_FOUND, O.R_ERASED, O.R_FOUND, O.R_NOOP, O.R_NOOP]
    assert out[3] == 21
    seq = O.Oracle(64)
    seq.insert([10, 20, 30])
    out2, st2 = seq.mixed(types, keys, sequential=True)
    assert st2.tolist() == [O.R_INSERTED, O.R_FOUND, O.R_ERASED, O.R_NOT_FOUND, O.R_NOOP, O.R_NOOP]


def test_resize_and_drain(O):
    nb = 256
    ks = np.arange(1, 1900, dtype=np.uint32) * np.uint32(2654435761)
    ks = ks[ks != 0]
    o = O.Oracle(nb)
    st = o.insert(ks)
    before = o.dump()
    assert o.resize(2 * nb) == 0
    assert np.array_equal(o.dump(), before)
    v, s = o.lookup(ks[st != O.R_FAILED])
    assert (s == O.R_FOUND).all()
    assert o.check_invariants()[0] == 0
    assert o.resize(8) == -1 and o.nb == 2 * nb  # cannot shrink below content; unchanged
    n_st = o.count_stash()
    assert o.drain_stash() == n_st and o.count_stash() == 0


# ------------------------------------------- 16-slot buckets (the wide layout)
# The same protocol with SLOTS = 16: Stage B takes the lowest empty of 16
# slots; Stage C's kick k still evicts slot k % 16 for k < 8 (MAX_EVICT = 8);
# the stash keeps 8 slots per canonical bucket.
def test_wide_lowest_slot_fills_sixteen_before_kicking(O):
    nb = 64
    c = 5
    ks = keys_with_h0("Default2Hash", nb, c, 17, h1_filter=lambda h1: h1 != c)
    o = O.Oracle(nb, slots_per_bucket=16)
    assert (o.insert(ks[:16]) == O.R_INSERTED).all()
    raw = o.export()
    assert raw.slots.size == nb * 16
    assert raw.slots.reshape(nb, 16)[c].tolist() == [(k + 1) << 32 | k for k in ks[:16]]
    # the 17th key: kick 0 moves slot 0 to its alternate, the new key takes slot 0
    assert o.insert([ks[16]])[0] == O.R_INSERTED
    s = o.export().slots.reshape(nb, 16)
    _, alt = hashref.buckets("Default2Hash", np.array([ks[0]], np.uint32), nb)
    assert s[c, 0] & 0xFFFFFFFF == ks[16]
    assert ks[0] in (s[int(alt[0])] & 0xFFFFFFFF).tolist()
    assert (s[c, 1:] & 0xFFFFFFFF).tolist() == ks[1:16]


def test_wide_kicks_try_only_slots_zero_to_seven(O):
    # slots 0..7 hold keys whose H1 == H0 (no kick can move them); slots 8..15
    # hold movable keys, which MAX_EVICT = 8 kicks never reach: the 17th key
    # goes to the stash
    nb = 32
    c = 3
    fixed = keys_with_h0("Default2Hash", nb, c, 8, h1_filter=lambda h1: h1 == c)
    movable = keys_with_h0("Default2Hash", nb, c, 9, start=fixed[-1] + 1, h1_filter=lambda h1: h1 != c)
    o = O.Oracle(nb, slots_per_bucket=16)
    assert (o.insert(fixed + movable[:8]) == O.R_INSERTED).all()
    assert o.insert([movable[8]])[0] == O.R_STASHED
    raw = o.export()
    assert raw.stash_count[c] == 1 and raw.stash[c * 8] & 0xFFFFFFFF == movable[8]
    assert (raw.slots.reshape(nb, 16)[c] & 0xFFFFFFFF).tolist() == fixed + movable[:8]


def test_wide_map_model_brute_force(O):
    rng = np.random.default_rng(11)
    for policy in ("Default2Hash", "MurmurCityHash", "TripleHash"):
        o = O.Oracle(64, policy, slots_per_bucket=16)
        model = {}
        universe = rng.integers(1, 5000, size=600).astype(np.uint32)
        for step in range(3000):
            k = int(universe[rng.integers(0, universe.size)])
            op = rng.integers(0, 3)
            if op == 0:
                v = int(rng.integers(0, 2**32))
                r = o.insert([k], [v])[0]
                assert r != O.R_FAILED
                if k in model:
                    assert r == O.R_UPDATED
                model[k] = v
            elif op == 1:
                vals, st = o.lookup([k])
                assert (st[0] == O.R_FOUND and vals[0] == model[k]) if k in model else st[0] == O.R_NOT_FOUND
            else:
                r = o.erase([k])[0]
                assert r == (O.R_ERASED if k in model else O.R_NOT_FOUND)
                model.pop(k, None)
        r, msg = o.check_invariants()
        assert r == 0, msg
        assert o.count_live() == len(model)


def test_wide_and_reference_widths_hold_the_same_content(O):
    # the width changes placement, never the content: same key set, values,
    # lookups, and fewer keys outside their canonical bucket at 16 slots
    keys = np.arange(1, 30001, dtype=np.uint32) * np.uint32(2654435761)
    vals = keys ^ np.uint32(0x5A5A5A5A)
    slots = 1 << 15
    o8 = O.Oracle(slots // 8)
    o16 = O.Oracle(slots // 16, slots_per_bucket=16)
    s8, s16 = o8.insert(keys, vals), o16.insert(keys, vals)
    assert not (s8 == O.R_FAILED).any() and not (s16 == O.R_FAILED).any()
    assert np.array_equal(o8.dump(), o16.dump())
    lk = np.concatenate([keys[::3], keys[::7] + np.uint32(1)])
    assert all(np.array_equal(a, b) for a, b in zip(o8.lookup(lk), o16.lookup(lk)))
    away8 = o8.probe_count(keys)[0] - keys.size
    away16 = o16.probe_count(keys)[0] - keys.size
    assert away16 < away8
    assert o16.resize(slots // 16 * 2) == 0 and np.array_equal(o16.dump(), o8.dump())
