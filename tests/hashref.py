"""Independent numpy formulations of the reference's hash functions, used to
pin the oracle's C restatement (oracle/hive_oracle.c).

Each is written from a different description than the oracle's code:
  hash1  Thomas Wang's hash32shift in its shift-add form:
         (key << 15) - key - 1 and (key + (key << 3)) + (key << 11)
  hash2  the reference's PTX sequence (include/hash.hpp:49-62) evaluated
         through a generic LOP3 truth-table evaluator (PTX ISA: immLut is
         F(0xF0, 0xCC, 0xAA), bit index a*4 + b*2 + c)
  murmur the general MurmurHash3_x86_32 block/tail/finalizer structure
  city   C++ semantics of include/hash.hpp:80-108 with the 64-bit constants
         applied to a 32-bit variable (promotion to 64 bits, truncation on
         assignment)
"""
import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def _u64(x):
    return np.asarray(x, dtype=np.uint64)


def hash1(keys):
    k = _u64(keys)
    k = ((k << np.uint64(15)) - k - np.uint64(1)) & M32
    k = k ^ (k >> np.uint64(12))
    k = (k + (k << np.uint64(2))) & M32
    k = k ^ (k >> np.uint64(4))
    k = ((k + (k << np.uint64(3))) + (k << np.uint64(11))) & M32
    k = k ^ (k >> np.uint64(16))
    return k.astype(np.uint32)


def lop3(a, b, c, imm):
    """Bitwise 3-input LUT: result bit = imm[(a<<2)|(b<<1)|c] per bit position."""
    a, b, c = _u64(a), _u64(b), _u64(c)
    out = np.zeros_like(a)
    for idx in range(8):
        if (imm >> idx) & 1:
            ta = a if (idx >> 2) & 1 else (~a & M32)
            tb = b if (idx >> 1) & 1 else (~b & M32)
            tc = c if idx & 1 else (~c & M32)
            out |= ta & tb & tc
    return out & M32


def hash2(keys):
    k = _u64(keys)
    k = ((k + np.uint64(0x7ed55d16)) + (k << np.uint64(12))) & M32
    temp = k >> np.uint64(19)                          # shr.u32 temp, key, 19
    k = lop3(k, temp, np.uint64(0xc761c23c), 0x96)     # lop3.b32 key, key, temp, C, 0x96
    k = ((k + np.uint64(0x165667b1)) + (k << np.uint64(5))) & M32
    k = ((k + np.uint64(0xd3a2646c)) ^ (k << np.uint64(9))) & M32
    k = ((k + np.uint64(0xfd7046c5)) + (k << np.uint64(3))) & M32
    temp = k >> np.uint64(16)
    k = lop3(k, temp, np.uint64(0xb55a4f09), 0x96)
    return k.astype(np.uint32)


def _rotl(x, r):
    return ((x << np.uint64(r)) | (x >> np.uint64(32 - r))) & M32


def _fmix(h):
    h = h ^ (h >> np.uint64(16))
    h = (h * np.uint64(0x85ebca6b)) & M32
    h = h ^ (h >> np.uint64(13))
    h = (h * np.uint64(0xc2b2ae35)) & M32
    return h ^ (h >> np.uint64(16))


def murmur(keys, seed=0):
    k = _u64(keys)  # one 4-byte little-endian block == the key itself
    h = np.full_like(k, np.uint64(seed))
    k = (k * np.uint64(0xcc9e2d51)) & M32
    k = _rotl(k, 15)
    k = (k * np.uint64(0x1b873593)) & M32
    h = h ^ k
    h = _rotl(h, 13)
    h = (h * np.uint64(5) + np.uint64(0xe6546b64)) & M32
    h = h ^ np.uint64(4)
    return _fmix(h).astype(np.uint32)


def city(keys):
    k0, k1, k2, k3 = 0x9ae16a3b, 0xc2b2ae35, 0x1b873593, 0x85ebca6b
    out = []
    for key in np.asarray(keys, dtype=np.uint64).tolist():
        k = key
        k = (k ^ k0) & 0xFFFFFFFF
        k = (k * k1) & 0xFFFFFFFF
        k = (k ^ k2) & 0xFFFFFFFF
        k = (k * k3) & 0xFFFFFFFF
        h = 0 ^ k
        h = (h * k1) & 0xFFFFFFFF
        h = (h ^ k2) & 0xFFFFFFFF
        h = (h * k3) & 0xFFFFFFFF
        h ^= 4
        h ^= h >> 16
        h = (h * 0x85ebca6b) & 0xFFFFFFFF
        h ^= h >> 13
        h = (h * 0xc2b2ae35) & 0xFFFFFFFF
        h ^= h >> 16
        out.append(h)
    return np.asarray(out, dtype=np.uint32)


def buckets(policy, keys, nb):
    """(H0, H1) for the two-hash policies, numpy-vectorized."""
    keys = np.asarray(keys, np.uint32)
    if policy in ("Default2Hash", 0):
        return hash1(keys) % np.uint32(nb), hash2(keys) % np.uint32(nb)
    if policy in ("MurmurCityHash", 1):
        return murmur(keys) % np.uint32(nb), city(keys) % np.uint32(nb)
    raise ValueError(policy)
