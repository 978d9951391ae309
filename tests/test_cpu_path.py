"""The multithreaded CPU path (oracle/hive_cpu_mt.cpp, the timed CPU
baseline) against the sequential oracle -- CPU only."""
import numpy as np
import pytest

from hivehashtable_amd import workloads as W


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


@pytest.mark.parametrize("policy", ["Default2Hash", "MurmurCityHash", "TripleHash"])
def test_cpu_bulk_insert_lookup_matches_oracle(O, policy):
    slots = 1 << 16
    n = W.keys_for_lf(slots, 0.8)
    keys = W.distinct_keys(n, 42)
    vals = W.values(n, 43)
    lk, present = W.lookup_set(keys, n, 0.5, 44)
    o = O.Oracle(slots // 8, policy)
    so = o.insert(keys, vals)
    c = O.CpuTable(slots // 8, policy, threads=8)
    sc, _ = c.insert(keys, vals)
    ok_o = np.isin(so, [O.R_INSERTED, O.R_STASHED, O.R_UPDATED])
    ok_c = np.isin(sc, [O.R_INSERTED, O.R_STASHED, O.R_UPDATED])
    assert ok_o.all() and ok_c.all()
    raw = c.export()
    assert np.array_equal(raw.live_kv(), o.dump())
    chk = O.Oracle(slots // 8, policy)
    chk.load(raw)
    assert chk.check_invariants()[0] == 0
    v1, s1 = o.lookup(lk)
    v2, s2, _ = c.lookup(lk)
    assert np.array_equal(s1, s2) and np.array_equal(v1, v2)
    assert np.array_equal(s2 == O.R_FOUND, present)


def test_cpu_high_load_keeps_invariants(O):
    slots = 1 << 14
    keys = W.distinct_keys(slots, 5)
    c = O.CpuTable(slots // 8, threads=8)
    st, _ = c.insert(keys)
    raw = c.export()
    chk = O.Oracle(slots // 8)
    chk.load(raw)
    r, msg = chk.check_invariants()
    assert r == 0, msg
    ok = np.isin(st, [O.R_INSERTED, O.R_STASHED])
    assert raw.live_kv().size == ok.sum()
    v, s, _ = c.lookup(keys)
    assert np.array_equal(s == O.R_FOUND, ok)


def test_cpu_mixed_matches_phased_oracle(O):
    slots = 1 << 15
    pre = W.distinct_keys(8000, 9)
    o = O.Oracle(slots // 8)
    c = O.CpuTable(slots // 8, threads=8)
    o.insert(pre)
    c.insert(pre)
    conflicted = set()
    for types, keys, vals in W.mixed_batches(pre, 1 << 17, [5000, 3000], seed=3):
        out_o, st_o = o.mixed(types, keys, vals)
        out_c, st_c, _ = c.mixed(types, keys, vals)
        # keys appearing once in a batch have exactly one admissible outcome
        uniq, cnt = np.unique(keys, return_counts=True)
        single = np.isin(keys, uniq[cnt == 1])
        conflicted.update(uniq[cnt > 1].tolist())
        norm = lambda s: np.where(s == O.R_STASHED, O.R_INSERTED, s)
        assert np.array_equal(norm(st_o)[single], norm(st_c)[single])
        assert np.array_equal(out_o[single], out_c[single])
    a, b = c.export().live_kv(), o.dump()
    assert np.array_equal(np.sort(a & 0xFFFFFFFF), np.sort(b & 0xFFFFFFFF))   # same key set
    clean = ~np.isin((b & 0xFFFFFFFF).astype(np.uint32), np.array(sorted(conflicted), np.uint32))
    ka = dict(zip((a & 0xFFFFFFFF).tolist(), (a >> 32).tolist()))
    for kv in b[clean].tolist():
        assert ka[kv & 0xFFFFFFFF] == kv >> 32


def test_cpu_path_is_race_free_under_tsan():
    """The multithreaded CPU path, built with ThreadSanitizer, under contended
    inserts and mixed batches on a small table: no race report, and the state
    after each step passes the oracle's invariant checker."""
    import os
    import shutil
    import subprocess
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    b = subprocess.run(["make", "-s", "-C", here, "tsan"], capture_output=True, text=True)
    if b.returncode != 0:
        pytest.skip(f"ThreadSanitizer build unavailable: {b.stderr[-300:]}")
    r = subprocess.run([os.path.join(here, "tsan_driver")], capture_output=True, text=True, timeout=300)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-2000:]
    assert r.returncode == 0, r.stderr[-2000:]
    assert "all checks passed" in r.stdout


def test_cpu_duplicate_heavy_inserts_store_each_key_once(O):
    """Many threads inserting copies of the same keys at once: every key ends
    up stored exactly once (the claim of an empty slot re-checks the key),
    with one of the values written for it."""
    slots = 1 << 12
    rng = np.random.default_rng(11)
    for rep in range(12):
        uniq = W.distinct_keys(300, 100 + rep)
        keys = np.repeat(uniq, 48)
        vals = (np.arange(keys.size, dtype=np.uint64) % 7 + 1).astype(np.uint32) + keys
        perm = rng.permutation(keys.size)
        keys, vals = keys[perm], vals[perm]
        c = O.CpuTable(slots // 8, threads=8)
        st, _ = c.insert(keys, vals)
        assert np.isin(st, [O.R_INSERTED, O.R_STASHED, O.R_UPDATED]).all()
        raw = c.export()
        chk = O.Oracle(slots // 8)
        chk.load(raw)
        r, msg = chk.check_invariants()
        assert r == 0, msg
        kv = raw.live_kv()
        assert np.array_equal(np.sort((kv & 0xFFFFFFFF).astype(np.uint32)), np.sort(uniq))
        got = dict(zip((kv & 0xFFFFFFFF).tolist(), (kv >> 32).tolist()))
        written = {}
        for k, v in zip(keys.tolist(), vals.tolist()):
            written.setdefault(k, set()).add(v)
        assert all(got[k] in written[k] for k in got)
        # exactly one INSERTED/STASHED per key
        fresh = np.isin(st, [O.R_INSERTED, O.R_STASHED])
        assert np.array_equal(np.sort(keys[fresh]), np.sort(uniq))
