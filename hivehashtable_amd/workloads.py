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

LF_POINTS = (0.65, 0.80, 0.90, 0.95, 1.0)


def keys_for_lf(num_slots, lf):
    """Number of keys that fill num_slots to load factor lf (truncating, as table_size * LF)."""
    return int(num_slots * lf)


def _bijection_params(seed):
    rng = np.random.default_rng([seed, 0x5EED])
    mult = (rng.integers(1, 2**31, size=3, dtype=np.uint64) * 2 + 1).astype(np.uint64)  # odd multipliers
    xors = rng.integers(0, 2**32, size=3, dtype=np.uint64)
    return mult, xors


def key_stream(seed, start, count):
    """f(start), ..., f(start+count-1) for a seeded bijection f of u32, so the
    outputs are distinct for distinct indices: rounds of xor with a constant,
    xor-shift right and multiplication by an odd constant (each invertible
    modulo 2^32)."""
    mult, xors = _bijection_params(seed)
    x = np.arange(start, start + count, dtype=np.uint64) & np.uint64(0xFFFFFFFF)
    m32 = np.uint64(0xFFFFFFFF)
    for r in range(3):
        x ^= xors[r]
        x ^= x >> np.uint64(16)
        x = (x * mult[r]) & m32
        x ^= x >> np.uint64(13)
    return x.astype(np.uint32)


def distinct_keys(n, seed, exclude=None, start=0):
    """n distinct u32 keys in [1, 2^32-2], a seeded pseudo-random permutation
    prefix (sampling without replacement).  Keys in `exclude` are skipped
    (small sets only); `start` continues the same stream, so
    distinct_keys(m, seed, start=n) never repeats distinct_keys(n, seed)."""
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


def mixed_batches(live_keys, key_space, sizes, mix=(0.5, 0.4, 0.1), seed=0):
    """Mixed batches: each op INSERT/LOOKUP/DELETE with probabilities `mix`;
    insert and lookup keys uniform over [1, key_space]; deletes drawn from
    keys inserted so far (the preload plus earlier batches).  Yields
    (types u8, keys u32, vals u32) per batch."""
    rng = np.random.default_rng(seed)
    pool = list(np.asarray(live_keys, np.uint32))
    pool_arr = np.asarray(pool, np.uint32)
    for sz in sizes:
        types = rng.choice(3, size=int(sz), p=list(mix)).astype(np.uint8)
        keys = rng.integers(1, key_space + 1, size=int(sz), dtype=np.uint64).astype(np.uint32)
        dl = types == 2
        if dl.any() and pool_arr.size:
            keys[dl] = pool_arr[rng.integers(0, pool_arr.size, size=int(dl.sum()))]
        vals = rng.integers(0, 2**32, size=int(sz), dtype=np.uint64).astype(np.uint32)
        ins = keys[types == 0]
        if ins.size:
            pool_arr = np.concatenate([pool_arr, ins])
        yield types, keys, vals


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
