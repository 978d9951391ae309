"""ctypes binding of the checker library (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py -- never by the hivehashtable_amd package.

``Oracle`` wraps the sequential restatement (hive_oracle.c); ``CpuTable``
wraps the multithreaded CPU path (hive_cpu_mt.cpp).  Both use the device
table layout: slots[nb*8] u64, stash[nb*8] u64, stash_count[nb] u32.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SLOTS = 8
STASH_SLOTS = 8

POLICIES = {
    "Default2Hash": 0,
    "MurmurCityHash": 1,
    "TripleHash": 2,
    "MurmurCityBitHash": 3,
    "Lookup2Hash": 4,
    "Lookup3Hash": 5,
}

R_FAILED, R_INSERTED, R_UPDATED, R_STASHED, R_FOUND, R_NOT_FOUND, R_ERASED, R_INVALID_KEY, R_NOOP = range(9)
OP_INSERT, OP_LOOKUP, OP_DELETE, OP_ATOMIC_INC, OP_NONE = range(5)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    u32, u64, sz, p = C.c_uint32, C.c_uint64, C.c_size_t, C.c_void_p
    for name in ("ho_hash1", "ho_hash2", "ho_murmur", "ho_city", "ho_city_premix", "ho_fmix32", "ho_crc64",
                 "ho_crc32", "ho_crc64_bitwise", "ho_crc32_bitwise"):
        getattr(L, name).restype = u32
        getattr(L, name).argtypes = [u32]
    L.ho_distinct_outputs.restype = u64
    L.ho_distinct_outputs.argtypes = [C.c_int, C.c_int]
    L.ho_murmur_seeded.restype = u32
    L.ho_murmur_seeded.argtypes = [C.c_char_p, sz, u32]
    L.ho_num_hashes.restype = C.c_int
    L.ho_num_hashes.argtypes = [C.c_int]
    L.ho_bucket.restype = u32
    L.ho_bucket.argtypes = [C.c_int, C.c_int, u32, u32]
    L.ho_alt_bucket.restype = u32
    L.ho_alt_bucket.argtypes = [C.c_int, u32, u32, u32]
    L.ho_create.restype = p
    L.ho_create.argtypes = [u32, C.c_int]
    L.ho_create_ex.restype = p
    L.ho_create_ex.argtypes = [u32, C.c_int, u32]
    L.ho_destroy.argtypes = [p]
    L.ho_clear.argtypes = [p]
    L.ho_import.argtypes = [p, p, p, p]
    L.ho_export.argtypes = [p, p, p, p]
    L.ho_insert.restype = C.c_int
    L.ho_insert.argtypes = [p, u32, u32]
    L.ho_lookup.restype = C.c_int
    L.ho_lookup.argtypes = [p, u32, C.POINTER(u32)]
    L.ho_erase.restype = C.c_int
    L.ho_erase.argtypes = [p, u32]
    L.ho_insert_batch.argtypes = [p, p, p, sz, p]
    L.ho_lookup_batch.argtypes = [p, p, sz, p, p]
    L.ho_erase_batch.argtypes = [p, p, sz, p]
    L.ho_mixed_phased.argtypes = [p, p, p, p, sz, p, p]
    L.ho_mixed_sequential.argtypes = [p, p, p, p, sz, p, p]
    L.ho_resize.restype = C.c_int
    L.ho_resize.argtypes = [p, u32]
    L.ho_drain_stash.restype = sz
    L.ho_drain_stash.argtypes = [p, p, sz]
    L.ho_dump.restype = sz
    L.ho_dump.argtypes = [p, p, sz]
    L.ho_count_live.restype = u64
    L.ho_count_live.argtypes = [p]
    L.ho_count_stash.restype = u64
    L.ho_count_stash.argtypes = [p]
    L.ho_check_invariants.restype = C.c_int
    L.ho_check_invariants.argtypes = [p, C.c_char_p, sz]
    L.ho_lookup_probe_count.argtypes = [p, p, sz, C.POINTER(u64), C.POINTER(u64)]
    L.hcpu_create.restype = p
    L.hcpu_create.argtypes = [u32, C.c_int]
    L.hcpu_destroy.argtypes = [p]
    L.hcpu_clear.restype = C.c_double
    L.hcpu_clear.argtypes = [p, C.c_int]
    L.hcpu_insert.restype = C.c_double
    L.hcpu_insert.argtypes = [p, p, p, sz, p, C.c_int]
    L.hcpu_lookup.restype = C.c_double
    L.hcpu_lookup.argtypes = [p, p, sz, p, p, C.c_int]
    L.hcpu_erase.restype = C.c_double
    L.hcpu_erase.argtypes = [p, p, sz, p, C.c_int]
    L.hcpu_mixed.restype = C.c_double
    L.hcpu_mixed.argtypes = [p, p, p, p, sz, p, p, C.c_int]
    L.hcpu_export.argtypes = [p, p, p, p]
    L.hcpu_hardware_threads.restype = C.c_int
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def policy_id(policy):
    if isinstance(policy, int):
        return policy
    return POLICIES[policy]


def hash1(k):
    return lib().ho_hash1(k)


def hash2(k):
    return lib().ho_hash2(k)


def murmur(k):
    return lib().ho_murmur(k)


def city(k):
    return lib().ho_city(k)


def bucket(policy, i, key, nb):
    return lib().ho_bucket(policy_id(policy), i, key, nb)


class RawState:
    """The device layout of one table, as numpy arrays (S slots per bucket)."""

    def __init__(self, nb, slots=None, stash=None, stash_count=None, slots_per_bucket=None):
        self.nb = nb
        if slots_per_bucket is None:
            slots_per_bucket = SLOTS if slots is None else len(slots) // nb
        self.S = slots_per_bucket
        self.slots = (np.zeros(nb * self.S, np.uint64) if slots is None
                      else np.ascontiguousarray(slots, np.uint64))
        self.stash = np.zeros(nb * STASH_SLOTS, np.uint64) if stash is None else np.ascontiguousarray(stash, np.uint64)
        self.stash_count = (np.zeros(nb, np.uint32) if stash_count is None
                            else np.ascontiguousarray(stash_count, np.uint32))

    def live_kv(self):
        """Sorted array of every live packed kv (table + stash[0:count])."""
        parts = [self.slots[self.slots != 0]]
        cnt = np.minimum(self.stash_count, STASH_SLOTS)
        mask = np.arange(STASH_SLOTS)[None, :] < cnt[:, None]
        st = self.stash.reshape(self.nb, STASH_SLOTS)[mask]
        parts.append(st[st != 0])
        return np.sort(np.concatenate(parts))


class Oracle:
    """Sequential single-operation model (hive_oracle.c), S slots per bucket."""

    def __init__(self, nb, policy="Default2Hash", slots_per_bucket=SLOTS):
        self.L = lib()
        self.nb = nb
        self.S = slots_per_bucket
        self.policy = policy_id(policy)
        self.h = self.L.ho_create_ex(nb, self.policy, self.S)
        if not self.h:
            raise MemoryError("ho_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ho_destroy(self.h)
            self.h = None

    def clear(self):
        self.L.ho_clear(self.h)

    def insert(self, keys, vals=None):
        keys = np.ascontiguousarray(keys, np.uint32)
        vals = None if vals is None else np.ascontiguousarray(vals, np.uint32)
        st = np.empty(len(keys), np.uint8)
        self.L.ho_insert_batch(self.h, _ptr(keys), _ptr(vals), len(keys), _ptr(st))
        return st

    def lookup(self, keys):
        keys = np.ascontiguousarray(keys, np.uint32)
        vals = np.empty(len(keys), np.uint32)
        st = np.empty(len(keys), np.uint8)
        self.L.ho_lookup_batch(self.h, _ptr(keys), len(keys), _ptr(vals), _ptr(st))
        return vals, st

    def erase(self, keys):
        keys = np.ascontiguousarray(keys, np.uint32)
        st = np.empty(len(keys), np.uint8)
        self.L.ho_erase_batch(self.h, _ptr(keys), len(keys), _ptr(st))
        return st

    def mixed(self, types, keys, vals=None, sequential=False):
        types = np.ascontiguousarray(types, np.uint8)
        keys = np.ascontiguousarray(keys, np.uint32)
        vals = None if vals is None else np.ascontiguousarray(vals, np.uint32)
        out = np.empty(len(keys), np.uint32)
        st = np.empty(len(keys), np.uint8)
        fn = self.L.ho_mixed_sequential if sequential else self.L.ho_mixed_phased
        fn(self.h, _ptr(types), _ptr(keys), _ptr(vals), len(keys), _ptr(out), _ptr(st))
        return out, st

    def resize(self, new_nb):
        r = self.L.ho_resize(self.h, new_nb)
        if r == 0:
            self.nb = new_nb
        return r

    def drain_stash(self):
        n = self.L.ho_drain_stash(self.h, None, 0)
        # the first call drained already; report through a dump taken before
        return n

    def dump(self):
        n = self.L.ho_dump(self.h, None, 0)
        out = np.empty(n, np.uint64)
        self.L.ho_dump(self.h, _ptr(out), n)
        return np.sort(out)

    def count_live(self):
        return int(self.L.ho_count_live(self.h))

    def count_stash(self):
        return int(self.L.ho_count_stash(self.h))

    def check_invariants(self):
        buf = C.create_string_buffer(256)
        r = self.L.ho_check_invariants(self.h, buf, 256)
        return r, buf.value.decode()

    def export(self):
        st = RawState(self.nb, slots_per_bucket=self.S)
        self.L.ho_export(self.h, _ptr(st.slots), _ptr(st.stash), _ptr(st.stash_count))
        return st

    def load(self, raw):
        if raw.nb != self.nb or raw.S != self.S:
            self.L.ho_destroy(self.h)
            self.h = self.L.ho_create_ex(raw.nb, self.policy, raw.S)
            self.nb = raw.nb
            self.S = raw.S
        self.L.ho_import(self.h, _ptr(raw.slots), _ptr(raw.stash), _ptr(raw.stash_count))

    def probe_count(self, keys):
        keys = np.ascontiguousarray(keys, np.uint32)
        br, sr = C.c_uint64(0), C.c_uint64(0)
        self.L.ho_lookup_probe_count(self.h, _ptr(keys), len(keys), C.byref(br), C.byref(sr))
        return int(br.value), int(sr.value)


class CpuTable:
    """Multithreaded CPU path (hive_cpu_mt.cpp) -- the timed CPU baseline."""

    def __init__(self, nb, policy="Default2Hash", threads=None):
        self.L = lib()
        self.nb = nb
        self.policy = policy_id(policy)
        self.threads = threads or self.L.hcpu_hardware_threads()
        self.h = self.L.hcpu_create(nb, self.policy)
        if not self.h:
            raise MemoryError("hcpu_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.hcpu_destroy(self.h)
            self.h = None

    def clear(self):
        return self.L.hcpu_clear(self.h, self.threads)

    def insert(self, keys, vals=None):
        keys = np.ascontiguousarray(keys, np.uint32)
        vals = None if vals is None else np.ascontiguousarray(vals, np.uint32)
        st = np.empty(len(keys), np.uint8)
        secs = self.L.hcpu_insert(self.h, _ptr(keys), _ptr(vals), len(keys), _ptr(st), self.threads)
        return st, secs

    def lookup(self, keys):
        keys = np.ascontiguousarray(keys, np.uint32)
        vals = np.empty(len(keys), np.uint32)
        st = np.empty(len(keys), np.uint8)
        secs = self.L.hcpu_lookup(self.h, _ptr(keys), len(keys), _ptr(vals), _ptr(st), self.threads)
        return vals, st, secs

    def erase(self, keys):
        keys = np.ascontiguousarray(keys, np.uint32)
        st = np.empty(len(keys), np.uint8)
        secs = self.L.hcpu_erase(self.h, _ptr(keys), len(keys), _ptr(st), self.threads)
        return st, secs

    def mixed(self, types, keys, vals=None):
        types = np.ascontiguousarray(types, np.uint8)
        keys = np.ascontiguousarray(keys, np.uint32)
        vals = None if vals is None else np.ascontiguousarray(vals, np.uint32)
        out = np.empty(len(keys), np.uint32)
        st = np.empty(len(keys), np.uint8)
        secs = self.L.hcpu_mixed(self.h, _ptr(types), _ptr(keys), _ptr(vals), len(keys), _ptr(out), _ptr(st),
                                 self.threads)
        return out, st, secs

    def export(self):
        st = RawState(self.nb, slots_per_bucket=SLOTS)  # the CPU path keeps the reference's 8-slot bucket
        self.L.hcpu_export(self.h, _ptr(st.slots), _ptr(st.stash), _ptr(st.stash_count))
        return st
