"""Synthetic workload generators (seeded, numpy).

Restates the reference harness's inputs (src/hashTableBenchmark.cc:21-100,
156-272, 275-432) with the seeded, duplicate-free key sets the benchmark
configurations call for: keys are distinct u32 drawn uniformly from
[1, 2^32-2] (0 is the empty sentinel; 0xFFFFFFFF is kept out of generated
keys), values are uniform u32.
"""
import numpy as np

KEY_LO = 1
KEY_HI = 0xFFFFFFFE  # inclusive
OP_INSERT_CODE, OP_LOOKUP_CODE, OP_DELETE_CODE = 0, 1, 2  # Operation.type (include/utils.h:10-16)

LF_POINTS = (0.65, 0.80, 0.90, 0.95, 1.0)


def keys_for_lf(num_slots, lf):
    """Number of keys that fill num_slots to load factor lf (truncating, as table_size * LF)."""
    return int(num_slots * lf)


def _bijection_params(seed):
    rng = np.random.default_rng([seed, 0x5EED])
    mult = (rng.integers(1, 2**31, size=3, dtype=np.uint64) * 2 + 1).astype(np.uint64)  # odd multipliers
    xors = rng.integers(0, 2**32, size=3, dtype=np.uint64)
    return mult, xors


def key_stream(seed, start, count, bits=32):
    """f(start), ..., f(start+count-1) for a seeded bijection f of [0, 2^bits),
    so the outputs are distinct for distinct indices: rounds of xor with a
    constant, xor-shift right and multiplication by an odd constant, each
    invertible modulo 2^bits."""
    mult, xors = _bijection_params(seed)
    mask = np.uint64((1 << bits) - 1)
    x = np.arange(start, start + count, dtype=np.uint64) & mask
    s1, s2 = np.uint64(bits // 2), np.uint64(bits // 2 - 3)
    for r in range(3):
        x ^= xors[r] & mask
        x ^= x >> s1
        x = (x * mult[r]) & mask
        x ^= x >> s2
    return x.astype(np.uint32)


def distinct_keys(n, seed, exclude=None, start=0, bits=32):
    """n distinct u32 keys, a seeded pseudo-random permutation prefix
    (sampling without replacement): in [1, 2^32-2] for bits=32, in
    [1, 2^bits] for bits < 32.  Keys in `exclude` are skipped (small sets
    only); `start` continues the same stream.  With bits=32 the stream holds
    0 and 0xFFFFFFFF once each, so the first n keys come from positions below
    n + 2 and distinct_keys(m, seed, start=n + 2) never repeats
    distinct_keys(n, seed)."""
    if bits < 32:
        if start + n > (1 << bits):
            raise ValueError("key space exhausted")
        cand = key_stream(seed, start, n + (0 if exclude is None else 0), bits).astype(np.uint64) + 1
        out = cand.astype(np.uint32)
        if exclude is not None:
            raise ValueError("exclude is only supported for bits=32")
        return np.ascontiguousarray(out)
    out = np.empty(0, np.uint32)
    excl = None if exclude is None else np.unique(np.asarray(exclude, np.uint32))
    pos = start
    while out.size < n:
        need = n - out.size
        take = need + (need >> 6) + 64
        cand = key_stream(seed, pos, take)
        pos += take
        keep = (cand >= KEY_LO) & (cand <= KEY_HI)
        if excl is not None and excl.size:
            keep &= ~np.isin(cand, excl, assume_unique=False)
        out = np.concatenate([out, cand[keep]])
    return np.ascontiguousarray(out[:n])


def values(n, seed):
    return np.random.default_rng(seed).integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)


def lookup_set(inserted, n, exist_ratio, seed, key_seed=None):
    """n lookup keys: round(n*exist_ratio) drawn from `inserted`, the rest fresh
    absent keys, shuffled (src/hashTableBenchmark.cc:289-431).  When
    `inserted` is distinct_keys(len, key_seed), the absent keys continue that
    stream (absent by construction); otherwise they are filtered against it.
    Returns (keys, expect_present bool mask)."""
    rng = np.random.default_rng(seed)
    n_hit = int(round(n * exist_ratio))
    hit = inserted[rng.integers(0, inserted.size, size=n_hit)] if n_hit else np.empty(0, np.uint32)
    n_miss = n - n_hit
    if not n_miss:
        miss = np.empty(0, np.uint32)
    elif key_seed is not None:
        miss = distinct_keys(n_miss, key_seed, start=inserted.size + (inserted.size >> 6) + 64)
    else:
        miss = distinct_keys(n_miss, seed + 7919, exclude=inserted)
    keys = np.concatenate([hit, miss]).astype(np.uint32)
    present = np.concatenate([np.ones(n_hit, bool), np.zeros(n_miss, bool)])
    perm = rng.permutation(n)
    return np.ascontiguousarray(keys[perm]), present[perm]


def cfg1(seed=42):
    """2^20 slots, LF 0.8: 838,860 distinct keys + values; 50/50 lookups."""
    slots = 1 << 20
    n = keys_for_lf(slots, 0.8)
    keys = distinct_keys(n, seed)
    vals = values(n, seed + 1)
    lk, present = lookup_set(keys, n, 0.5, seed + 2, key_seed=seed)
    return slots, keys, vals, lk, present


def log_uniform_sizes(count, lo, hi, seed):
    """Batch sizes log-uniform in [lo, hi] (integers)."""
    rng = np.random.default_rng(seed)
    e = rng.uniform(np.log(lo), np.log(hi), size=count)
    return np.clip(np.round(np.exp(e)).astype(np.int64), lo, hi)


def mixed_stream(live_keys, key_space, sizes, mix=(0.5, 0.4, 0.1), seed=0):
    """A stream of mixed batches, all at once (vectorized).

    Each op is INSERT/LOOKUP/DELETE with probabilities `mix`; insert and
    lookup keys are uniform over [1, key_space]; a delete draws uniformly from
    the keys inserted before its batch (the preload `live_keys` plus inserts
    of earlier batches).  Returns (types u8, keys u32, vals u32, offsets u64
    of len(sizes) + 1)."""
    rng = np.random.default_rng(seed)
    sizes = np.asarray(sizes, np.int64)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    total = int(offsets[-1])
    types = rng.choice(3, size=total, p=list(mix)).astype(np.uint8)
    keys = rng.integers(1, key_space + 1, size=total, dtype=np.uint64).astype(np.uint32)
    vals = rng.integers(0, 2**32, size=total, dtype=np.uint64).astype(np.uint32)
    live = np.asarray(live_keys, np.uint32)
    ins_pos = np.flatnonzero(types == 0)
    pool = np.concatenate([live, keys[ins_pos]])
    batch_of = np.repeat(np.arange(sizes.size), sizes)
    # inserts of batches before each batch
    ins_per_batch = np.bincount(batch_of[ins_pos], minlength=sizes.size)
    ins_before = np.concatenate([[0], np.cumsum(ins_per_batch)[:-1]])
    dpos = np.flatnonzero(types == 2)
    if dpos.size:
        avail = live.size + ins_before[batch_of[dpos]]
        has = avail > 0
        pick = (rng.random(dpos.size) * np.maximum(avail, 1)).astype(np.int64)
        keys[dpos[has]] = pool[pick[has]]
    return types, keys, vals, offsets


def mixed_batches(live_keys, key_space, sizes, mix=(0.5, 0.4, 0.1), seed=0):
    """mixed_stream, yielded batch by batch as (types, keys, vals)."""
    types, keys, vals, off = mixed_stream(live_keys, key_space, sizes, mix, seed)
    for b in range(len(off) - 1):
        lo, hi = int(off[b]), int(off[b + 1])
        yield types[lo:hi], keys[lo:hi], vals[lo:hi]


def zipf_sampler(n_items, theta, seed):
    """YCSB-style Zipfian sampler over ranks [0, n_items) (Gray et al.,
    "Quickly generating billion-record synthetic databases", SIGMOD 1994),
    returned as a function of a sample count."""
    rng = np.random.default_rng(seed)
    # zeta(n, theta) by integral approximation with an exact head
    head = min(n_items, 1 << 20)
    zeta_head = np.sum(1.0 / np.arange(1, head + 1, dtype=np.float64) ** theta)
    if n_items > head:
        a, b = head + 0.5, n_items + 0.5
        zeta_tail = (b ** (1 - theta) - a ** (1 - theta)) / (1 - theta)
    else:
        zeta_tail = 0.0
    zetan = zeta_head + zeta_tail
    zeta2 = 1.0 + 0.5 ** theta
    alpha = 1.0 / (1.0 - theta)
    eta = (1 - (2.0 / n_items) ** (1 - theta)) / (1 - zeta2 / zetan)

    def sample(k):
        u = rng.random(k)
        uz = u * zetan
        r = (n_items * (eta * u - eta + 1) ** alpha).astype(np.int64)
        r = np.where(uz < 1.0, 0, np.where(uz < zeta2, 1, r))
        return np.clip(r, 0, n_items - 1)

    return sample


# cfg5 request mixes: fractions of (read, update of a preloaded key, insert of
# a new key).  A/B/C follow YCSB workloads A, B, C (Zipfian over the preloaded
# keys); D is YCSB D (reads skewed to the latest inserted keys, new inserts).
SERVE_MIXES = {"A": (0.5, 0.5, 0.0), "B": (0.95, 0.05, 0.0), "C": (1.0, 0.0, 0.0), "D": (0.95, 0.0, 0.05)}


def serving_trace(mix, preload, n_requests, offered_mops, seed, theta=0.99, size_lo=1 << 8, size_hi=1 << 18,
                  fresh=None):
    """Client requests for the serving loop.

    Request sizes are log-uniform in [size_lo, size_hi]; arrivals are a
    Poisson process whose rate offers `offered_mops` million ops per second
    on average (0: every request arrives at time 0, a closed-loop saturation
    run).  Reads are LOOKUP ops, updates and new keys are INSERT ops.  Keys of
    mixes A-C are Zipfian (theta) over `preload` (rank i = preload[i]); mix D
    takes new keys from `fresh` in order and draws reads Zipfian over the keys
    inserted before the request, most recent first.
    Returns (types u8, keys u32, vals u32, req_offsets u64, arrival_us f64)."""
    pr, pu, _ = SERVE_MIXES[mix]
    rng = np.random.default_rng(seed)
    sizes = log_uniform_sizes(n_requests, size_lo, size_hi, seed + 1)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    total = int(off[-1])
    if offered_mops and offered_mops > 0:
        rate = offered_mops * 1e6 / sizes.mean()  # requests per second
        gaps = rng.exponential(1e6 / rate, n_requests)
        arrivals = np.cumsum(gaps) - gaps[0]
    else:
        arrivals = np.zeros(n_requests)
    u = rng.random(total)
    kind = (u >= pr).astype(np.int8) + (u >= pr + pu).astype(np.int8)  # 0 read, 1 update, 2 new
    types = np.where(kind == 0, OP_LOOKUP_CODE, OP_INSERT_CODE).astype(np.uint8)
    vals = rng.integers(0, 2**32, size=total, dtype=np.uint64).astype(np.uint32)
    pre = np.asarray(preload, np.uint32)
    keys = np.empty(total, np.uint32)
    if mix != "D":
        z = zipf_sampler(pre.size, theta, seed + 2)
        keys[:] = pre[z(total)]
        return types, keys, vals, off, arrivals
    new_pos = np.flatnonzero(kind == 2)
    if fresh is None or len(fresh) < new_pos.size:
        raise ValueError(f"mix D needs {new_pos.size} fresh keys")
    fresh = np.asarray(fresh, np.uint32)[: new_pos.size]
    keys[new_pos] = fresh
    req_of = np.repeat(np.arange(n_requests), sizes)
    new_per_req = np.bincount(req_of[new_pos], minlength=n_requests)
    new_before = np.concatenate([[0], np.cumsum(new_per_req)[:-1]])
    read_pos = np.flatnonzero(kind != 2)
    known = pre.size + new_before[req_of[read_pos]]
    z = zipf_sampler(pre.size + fresh.size, theta, seed + 2)
    rank = np.minimum(z(read_pos.size), known - 1)
    pool = np.concatenate([pre, fresh])
    keys[read_pos] = pool[known - 1 - rank]
    return types, keys, vals, off, arrivals
