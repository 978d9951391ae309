"""Windowed batches at the edges the other window tests leave open.

* high load factor: chains that end in deletes beside chains whose final
  insert takes the slow path (cuckoo kicks), in the same window;
* windows of at least 2^18 ops with type lists (the listed phase kernels),
  in the windowed stream and in ordered serving;
* deferred batches that touch one key more than once (delete + insert,
  lookup + insert, two inserts, two deletes), whose outcomes the chain
  kernels fix in op order, exactly as the oracle's batch does;
* caller type bytes that are not op types (5, 252..254) in plain and window
  paths: NOOP with value 0.

Every comparison is against the oracle run batch by batch (or request by
request), through the C-ABI, status, value and the final content with values.
"""
import numpy as np
import pytest

from conftest import gpu_available
from hivehashtable_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


@pytest.fixture(scope="module")
def H():
    import hivehashtable_amd as H
    return H


def np_(t, dtype=None):
    a = t.cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def norm(st):
    return np.where(st == 3, 1, st)  # STASHED is a placement detail: count it as INSERTED


def oracle_batches(O, slots, pre, pre_v, types, keys, vals, off, policy="Default2Hash", start=None):
    """The oracle run batch by batch, from the preload or from `start`: a GPU
    table's content (packed kv), where a preload at high load factor may
    leave a few keys unplaced -- a placement outcome of concurrent inserts."""
    o = O.Oracle(slots // 8, policy)
    if start is not None:
        o.insert((start & np.uint64(0xFFFFFFFF)).astype(np.uint32), (start >> np.uint64(32)).astype(np.uint32))
    elif pre is not None and pre.size:
        o.insert(pre, pre_v)
    ov = np.empty(keys.size, np.uint32)
    os_ = np.empty(keys.size, np.uint8)
    for b in range(off.size - 1):
        lo, hi = int(off[b]), int(off[b + 1])
        if hi > lo:
            ov[lo:hi], os_[lo:hi] = o.mixed(types[lo:hi], keys[lo:hi], vals[lo:hi])
    return o, ov, os_


def content(t):
    return np.sort(np_(t.dump()).view(np.uint64))


def high_lf_chains(seed):
    """A 2^15-slot table at LF 0.86 and 64 batches of 300 distinct keys each
    from a pool of 4000 (half of it preloaded): 45 % inserts, 20 % lookups,
    35 % deletes, so most keys have chains across the window's batches, some
    ending deleted and some ending with an insert at LF 0.87 (about one
    insert in ten takes the slow path)."""
    slots = 1 << 15
    rng = np.random.default_rng(seed)
    pool = W.distinct_keys(4000, seed)
    back = W.distinct_keys(int(slots * 0.86) - 2000, seed, start=4002)
    pre = np.concatenate([back, pool[:2000]])
    pre_v = W.values(pre.size, seed + 2)
    parts = [rng.choice(pool, 300, replace=False) for _ in range(64)]
    keys = np.concatenate(parts).astype(np.uint32)
    types = rng.choice(3, size=keys.size, p=[0.45, 0.2, 0.35]).astype(np.uint8)
    vals = W.values(keys.size, seed + 3)
    off = (np.arange(65) * 300).astype(np.uint64)
    return slots, pre, pre_v, types, keys, vals, off


def test_high_lf_windows_with_deleting_and_kicking_chains(H, O):
    """Final deletes and final inserts of a window's chains run in two
    launches: an erase never meets a concurrent cuckoo move.  Regression test
    for the single apply pass, at LF 0.86-0.87 where about one insert in ten
    takes the slow path."""
    print("[case] windows at LF 0.86: chains ending in deletes and slow-path inserts", flush=True)
    slots, pre, pre_v, types, keys, vals, off = high_lf_chains(4101)
    for win in (1 << 12, 1 << 15):
        print(f"[case] windows at LF 0.86, window {win} ops", flush=True)
        t = H.HiveTable(slots)
        t.set_stream_window(win, 1 << 10)
        t.insert(pre, pre_v)
        o, ov, os_ = oracle_batches(O, slots, pre, pre_v, types, keys, vals, off, start=content(t))
        assert not (os_ == 0).any()  # no insert fails in the oracle: every op has one outcome
        gv, gs, done, _ = t.mixed_stream(types, keys, vals, off, 3)
        gv, gs = np_(gv, np.uint32), np_(gs)
        assert np.array_equal(norm(gs), norm(os_)), win
        assert np.array_equal(gv, ov), win
        assert np.array_equal(content(t), np.sort(o.dump())), win
        st = t.stats()
        assert st["slow_path_ops"] > 0
        t.close()


def test_big_windows_run_the_listed_phase_kernels(H, O):
    """Windows of 2^19 ops (above the 2^18-op threshold of the type lists)
    with split_by_type on: the window's phase kernels run over insert and
    delete lists.  Ops whose key repeats inside a batch are excluded from the
    per-op comparison (two inserts of one key in one batch may land in either
    order) together with that key's later ops; the key set must match."""
    print("[case] windowed stream: 2^19-op windows with type lists", flush=True)
    slots = 1 << 20
    pre = W.distinct_keys(slots // 2, 4201, bits=21)
    sizes = np.full(40, 1 << 14)
    types, keys, vals, off = W.mixed_stream(pre, 1 << 21, sizes, seed=4202)
    o, ov, os_ = oracle_batches(O, slots, pre, pre + 1, types, keys, vals, off)
    nb = off.size - 1
    batch = np.repeat(np.arange(nb), np.diff(off).astype(np.int64))
    comb = (batch.astype(np.uint64) << np.uint64(32)) | keys.astype(np.uint64)
    _, inv, cnt = np.unique(comb, return_inverse=True, return_counts=True)
    dup = cnt[inv] > 1
    first_dup = {}
    for k, b in zip(keys[dup].tolist(), batch[dup].tolist()):
        first_dup[k] = min(b, first_dup.get(k, nb))
    fd = np.array([first_dup.get(k, nb) for k in keys.tolist()])
    clean = (~dup) & (batch < fd)
    assert clean.mean() > 0.9
    keyset = np.sort(o.dump() & 0xFFFFFFFF)
    for split in (1, 0):
        print(f"[case] 2^19-op windows, split_by_type={split}", flush=True)
        t = H.HiveTable(slots)
        for op in (0, 1, 2):
            t.set_launch(op, 256, split_by_type=split)
        t.set_stream_window(1 << 19, 1 << 15)
        t.insert(pre)
        gv, gs, done, _ = t.mixed_stream(types, keys, vals, off, 3)
        gv, gs = np_(gv, np.uint32), np_(gs)
        assert np.unique(done).size == 2  # a window of 32 batches (2^19 ops, listed), then one of 8
        assert np.array_equal(norm(gs)[clean], norm(os_)[clean]), split
        assert np.array_equal(gv[clean], ov[clean]), split
        assert np.array_equal(np.sort(content(t) & np.uint64(0xFFFFFFFF)), keyset.astype(np.uint64)), split
        t.close()


def test_ordered_serving_big_batches_run_the_listed_phase_kernels(H, O):
    """Ordered serving of 150 requests of 4096 ops (no key twice inside a
    request, keys colliding across requests), all arrived at once: batches of
    up to 2^20 ops run as windows above the type-list threshold, and every
    request gets exactly its result alone after the requests before it."""
    print("[case] ordered serving: windows above 2^18 ops with type lists", flush=True)
    slots = 1 << 20
    rng = np.random.default_rng(4301)
    pre = W.distinct_keys(slots // 2, 4302, bits=20)
    pre_v = W.values(pre.size, 4303)
    nreq, size = 150, 1 << 12
    keys = np.concatenate([rng.choice(np.arange(1, 1 << 20, dtype=np.uint32), size, replace=False)
                           for _ in range(nreq)]).astype(np.uint32)
    types = rng.choice(3, size=keys.size, p=[0.4, 0.4, 0.2]).astype(np.uint8)
    vals = W.values(keys.size, 4304)
    off = (np.arange(nreq + 1) * size).astype(np.uint64)
    o, ov, os_ = oracle_batches(O, slots, pre, pre_v, types, keys, vals, off)
    t = H.HiveTable(slots)
    t.set_launch(0, 1024, inline_slow_path=1, combine_duplicates=1, split_by_type=1)
    t.insert(pre, pre_v)
    gv, gs, lat, rep = t.serve(types, keys, vals, off, np.zeros(nreq), max_batch_ops=1 << 20, ordered=1)
    gv, gs = np_(gv, np.uint32), np_(gs)
    print(f"[case] ordered serving big batches: {rep['batches']} batches", flush=True)
    assert rep["batches"] <= 4 and rep["mean_batch_ops"] >= (1 << 18)
    assert np.array_equal(norm(gs), norm(os_))
    assert np.array_equal(gv, ov)
    assert np.array_equal(content(t), np.sort(o.dump()))
    t.close()


def repeated_key_batches(seed, nbatches=40, pool_size=64):
    """Batch 0 inserts every pool key once (it is every key's first written
    batch in a window holding all batches); every later batch has 48 ops on
    16 pool keys in patterns that repeat a key: delete + insert, lookup +
    insert, insert + insert, delete + delete, delete + lookup + insert, each
    in a shuffled order."""
    rng = np.random.default_rng(seed)
    pool = W.distinct_keys(pool_size, seed)
    ty = [np.zeros(pool_size, np.uint8)]
    ks = [pool.copy()]
    patterns = [(2, 0), (1, 0), (0, 0), (2, 2), (2, 1, 0), (0, 2), (0, 1, 0), (1, 2)]
    for _ in range(nbatches - 1):
        bt, bk = [], []
        for k in rng.choice(pool, 16, replace=False):
            p = patterns[rng.integers(len(patterns))]
            bt.extend(p)
            bk.extend([k] * len(p))
        perm = rng.permutation(len(bt))
        ty.append(np.array(bt, np.uint8)[perm])
        ks.append(np.array(bk, np.uint32)[perm])
    sizes = [a.size for a in ks]
    types = np.concatenate(ty)
    keys = np.concatenate(ks).astype(np.uint32)
    vals = W.values(keys.size, seed + 1)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return types, keys, vals, off


def test_deferred_batches_with_repeated_keys_match_the_oracle(H, O):
    """In a deferred batch (after the key's first written batch of the
    window), every op on a repeated key has one outcome in op order: deletes
    before inserts, the first delete and the first insert find the key as it
    was, the last insert's value stays.  Status, value and final content
    (with values) equal the oracle run batch by batch, in the windowed stream
    and in ordered serving (requests = the batches)."""
    print("[case] windows: repeated keys inside deferred batches", flush=True)
    types, keys, vals, off = repeated_key_batches(4401)
    slots = 1 << 12
    o, ov, os_ = oracle_batches(O, slots, None, None, types, keys, vals, off)
    ref_content = np.sort(o.dump())
    t = H.HiveTable(slots)
    t.set_stream_window(1 << 13, 1 << 10)
    gv, gs, done, _ = t.mixed_stream(types, keys, vals, off, 3)
    assert np.unique(done).size == 1  # one window holds every batch
    assert np.array_equal(norm(np_(gs)), norm(os_))
    assert np.array_equal(np_(gv, np.uint32), ov)
    assert np.array_equal(content(t), ref_content)
    t.close()
    nreq = off.size - 1
    t = H.HiveTable(slots)
    gv, gs, lat, rep = t.serve(types, keys, vals, off, np.zeros(nreq), max_batch_ops=1 << 20, ordered=1)
    print(f"[case] ordered serving, repeated keys: {rep['batches']} batches", flush=True)
    assert np.array_equal(norm(np_(gs)), norm(os_))
    assert np.array_equal(np_(gv, np.uint32), ov)
    assert np.array_equal(content(t), ref_content)
    t.close()


def test_caller_type_bytes_outside_the_op_types_are_noops(H, O):
    """Type bytes 3 (ATOMIC_INC), 4 (NONE), 5 and 252..254 -- the last three
    are the window's internal codes -- are no-ops with result 0 from every
    entry point: hive_mixed (plain, and split by type at 2^18 ops),
    hive_mixed_ops, and windows."""
    print("[case] caller type bytes 3, 4, 5, 252, 253, 254", flush=True)
    slots = 1 << 19
    pre = W.distinct_keys(1 << 16, 4501)
    rng = np.random.default_rng(4502)
    n = 1 << 18
    keys = np.concatenate([pre, W.distinct_keys(n - pre.size, 4501, start=pre.size + 2)])[rng.permutation(n)]
    types = rng.choice(np.array([0, 1, 2, 3, 4, 5, 252, 253, 254], np.uint8), size=n)
    vals = W.values(n, 4503)
    o = O.Oracle(slots // 8)
    o.insert(pre)
    ov, os_ = o.mixed(types, keys, vals)
    odd = ~np.isin(types, (0, 1, 2))
    assert (os_[odd] == 8).all() and (ov[odd] == 0).all()
    for split in (0, 1):
        t = H.HiveTable(slots)
        t.set_launch(0, 256, split_by_type=split)
        t.insert(pre)
        gv, gs = t.mixed(types, keys, vals)
        gv, gs = np_(gv, np.uint32), np_(gs)
        assert np.array_equal(norm(gs), norm(os_)), split
        assert np.array_equal(gv, ov), split
        t.close()
    # the reference's records: int32 type, 0/1 results
    ops = np.zeros(4096, dtype=[("type", "<i4"), ("pad", "<u4"), ("key", "<u8")])
    ops["type"] = types[:4096].astype(np.int32)
    ops["key"] = keys[:4096]
    t = H.HiveTable(slots)
    t.insert(pre)
    res = np_(t.mixed_ops(ops))
    assert (res[odd[:4096]] == 0).all()
    expect = np.isin(os_[:4096], (1, 2, 3, 4, 6)).astype(np.int64)
    # the record kernels insert value key + 1: outcomes, not values, are compared
    assert np.array_equal(res, expect)
    t.close()
    # a window of two batches: the same bytes in a deferred and a plain position
    off = np.array([0, n // 2, n], np.uint64)
    _, wv, ws = oracle_batches(O, slots, pre, pre + 1, types, keys, vals, off)
    t = H.HiveTable(slots)
    t.set_stream_window(1 << 20, 1 << 19)
    t.insert(pre)
    gv, gs, done, _ = t.mixed_stream(types, keys, vals, off, 3)
    assert np.unique(done).size == 1
    assert np.array_equal(norm(np_(gs)), norm(ws))
    assert np.array_equal(np_(gv, np.uint32), wv)
    t.close()
