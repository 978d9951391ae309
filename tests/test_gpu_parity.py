"""GPU parity: libhive.so (through its C-ABI) against the oracle.

Bar: bit-exact on integers.  Concurrent placement is not deterministic (the
reference's kernels are not either), so the exact comparisons are on the
observables the sequential semantics fix: per-op found/value/erase results,
insert outcome classes (placed / updated / failed), the content multiset,
and -- for read-only and delete kernels, which move nothing -- the raw table
state itself.  GPU-built tables are also loaded back into the oracle for its
structural invariant check.
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


OK = (1, 2, 3)  # INSERTED, UPDATED, STASHED
LAYOUT_NAMES = ["bucket8", "bucket16", "inline8"]
SPB = {"bucket8": 8, "bucket16": 16, "inline8": 8}  # slots per bucket of each layout


def placed(st):
    return np.isin(st, OK)


def gpu_raw(tbl, O):
    slots, stash, cnt = tbl.export_raw()
    return O.RawState(tbl.num_buckets, slots, stash, cnt)


def check_gpu_state(tbl, O, policy):
    raw = gpu_raw(tbl, O)
    chk = O.Oracle(raw.nb, policy)
    chk.load(raw)
    r, msg = chk.check_invariants()
    assert r == 0, msg
    return raw, chk


@pytest.mark.parametrize("layout", LAYOUT_NAMES)
@pytest.mark.parametrize("policy", ["Default2Hash", "MurmurCityHash"])
def test_cfg1_insert_then_lookup(H, O, policy, layout):
    """cfg1 (2^20 slots, LF 0.8) against the oracle.  Which keys a full
    bucket kicks or stashes is a placement outcome of concurrent execution.
    In a table this small nearly the whole batch is in flight at once, so
    staged kicks run against an almost full table and stash pushes cluster:
    now and then a canonical stash bucket fills and a key fails on the GPU
    that the sequential oracle placed (the reference's Stage D,
    include/hive_hash_insert.cuh:489-517).  Such a key's canonical stash
    must hold 8 entries.  Every other key's outcome class must match the
    oracle's; the content must be exactly the placed keys with their values;
    lookups must match the oracle on every key placed in both, and the
    oracle run on the GPU's own state on every key."""
    width = SPB[layout]
    print(f"[case] cfg1 2^20 slots LF 0.8 {policy} layout {layout}", flush=True)
    slots, keys, vals, lk, present = W.cfg1()
    nb = slots // width
    o = O.Oracle(nb, policy, slots_per_bucket=width)
    so = o.insert(keys, vals)
    t = H.HiveTable(slots, policy, layout=layout)
    sg = np_(t.insert(keys, vals))
    assert not (sg == 2).any()  # distinct keys: no updates
    gok, ook = placed(sg), placed(so)
    assert np.isin(sg[~gok], (0,)).all()  # the only other outcome: FAILED
    raw, chk = check_gpu_state(t, O, policy)
    odd = np.flatnonzero(gok != ook)
    assert odd.size <= 8, odd.size
    oraw = o.export()
    for i in odd.tolist():
        c = O.bucket(policy, 0, int(keys[i]), nb)
        full = raw.stash_count[c] if not gok[i] else oraw.stash_count[c]
        assert full >= 8, (int(keys[i]), c, int(full))  # failed only with its canonical stash full
    if odd.size:
        print(f"[case] cfg1 {policy} {layout}: {odd.size} keys failed with a full canonical stash", flush=True)
    # content: exactly the placed keys with their values
    want = np.sort((vals[gok].astype(np.uint64) << np.uint64(32)) | keys[gok].astype(np.uint64))
    assert np.array_equal(np.sort(raw.live_kv()), want)
    gv, gs = t.lookup(lk)
    gv, gs = np_(gv, np.uint32), np_(gs)
    ov, os_ = o.lookup(lk)
    both = ~np.isin(lk, keys[odd])
    assert np.array_equal(gs[both], os_[both]) and np.array_equal(gv[both], ov[both])  # found flag and value
    cv, cs = chk.lookup(lk)                                    # oracle on the GPU's own state
    assert np.array_equal(gs, cs) and np.array_equal(gv, cv)
    assert np.array_equal(gs == 4, present & ~np.isin(lk, keys[~gok]))
    a, b = t.count_exact()
    st = t.stats()
    assert (a, b) == (st["table_live"], st["stash_live"]) and a + b == gok.sum()


@pytest.mark.parametrize("layout", LAYOUT_NAMES)
@pytest.mark.parametrize("policy", ["Default2Hash", "MurmurCityHash", "TripleHash", "MurmurCityBitHash",
                                    "Lookup2Hash", "Lookup3Hash"])
def test_lookup_and_erase_on_oracle_built_table(H, O, policy, layout):
    # LF 0.97 so the stash is populated; lookups and deletes move nothing,
    # so results and the raw state after deletes must be identical.
    width = SPB[layout]
    print(f"[case] oracle-built table {policy} layout {layout}", flush=True)
    nb = 1 << 12
    keys = W.distinct_keys(int(nb * width * 0.97), 11)
    vals = W.values(keys.size, 12)
    o = O.Oracle(nb, policy, slots_per_bucket=width)
    so = o.insert(keys, vals)
    assert (so == 3).any()
    raw = o.export()
    t = H.HiveTable(nb * width, policy, layout=layout)
    t.import_raw(raw.slots, raw.stash, raw.stash_count)
    probe = np.concatenate([keys, W.distinct_keys(5000, 13, exclude=keys), np.array([0], np.uint32)])
    ov, os_ = o.lookup(probe)
    # lane-per-key, cooperative (4 or 8 lanes per key) and several keys per
    # lane (2 or 4), block sizes whose last block is partly empty
    for lanes, kpl, blk in ((1, 1, 256), (4, 1, 256), (8, 1, 256), (4, 1, 1024), (8, 1, 64), (1, 2, 256),
                            (1, 4, 64), (1, 4, 1024), (1, 2, 512)):
        t.set_launch(1, blk, keys_per_lane=kpl, lanes_per_key=lanes)
        gv, gs = t.lookup(probe)
        assert np.array_equal(np_(gs), os_) and np.array_equal(np_(gv, np.uint32), ov), (lanes, kpl, blk)
    t.set_launch(1, 256)
    rng = np.random.default_rng(14)
    dk = np.concatenate([rng.choice(keys, 4000, replace=False), W.distinct_keys(500, 15, exclude=keys),
                         np.array([0], np.uint32)])
    ge = np_(t.erase(dk))
    oe = o.erase(dk)
    assert np.array_equal(ge, oe)
    after = gpu_raw(t, O)
    want = o.export()
    assert np.array_equal(after.slots, want.slots)
    assert np.array_equal(after.stash_count, want.stash_count)
    assert np.array_equal(after.live_kv(), want.live_kv())


@pytest.mark.parametrize("layout", LAYOUT_NAMES)
@pytest.mark.parametrize("policy", ["Default2Hash", "TripleHash", "MurmurCityBitHash", "Lookup3Hash"])
def test_insert_at_full_load_keeps_invariants(H, O, layout, policy):
    """LF 1.0: most of the last keys take the locked slow path (own H1,
    kicks, stash pushes), in two- and three-hash policies.  The oracle's
    structural check on the GPU's state (every key once, in a candidate
    bucket or its canonical stash), and every placed key found by the GPU's
    lookup, which follows the hint bits."""
    width = SPB[layout]
    print(f"[case] LF 1.0 insert, 2^15 slots, {policy} layout {layout}", flush=True)
    slots = 1 << 15
    keys = W.distinct_keys(slots, 21)
    o = O.Oracle(slots // width, policy, slots_per_bucket=width)
    so = o.insert(keys)
    t = H.HiveTable(slots, policy, layout=layout)
    sg = np_(t.insert(keys))
    raw, chk = check_gpu_state(t, O, policy)
    ok = placed(sg)
    assert raw.live_kv().size == ok.sum()
    gv, gs = t.lookup(keys)
    gs, gv = np_(gs), np_(gv, np.uint32)
    assert np.array_equal(gs == 4, ok)
    assert np.array_equal(gv[ok], keys[ok] + 1)
    # failure count is placement dependent; both stay within a few percent
    assert abs(int((~ok).sum()) - int((~placed(so)).sum())) <= max(64, 0.02 * slots)


@pytest.mark.parametrize("layout", LAYOUT_NAMES)
def test_imported_table_with_holes_is_probed_in_full(H, O, layout):
    """An imported state may hold hint bits on buckets with empty slots (here
    the oracle's table at LF 0.97 after 40 % of its keys were erased), so an
    import turns the lookups' stop at an H0 with room off: every lookup must
    equal the oracle's."""
    width = SPB[layout]
    print(f"[case] imported table with holes, layout {layout}", flush=True)
    nb = 1 << 12
    keys = W.distinct_keys(int(nb * width * 0.97), 16)
    vals = W.values(keys.size, 17)
    o = O.Oracle(nb, "Default2Hash", slots_per_bucket=width)
    o.insert(keys, vals)
    rng = np.random.default_rng(18)
    o.erase(rng.choice(keys, int(keys.size * 0.4), replace=False))
    raw = o.export()
    t = H.HiveTable(nb * width, "Default2Hash", layout=layout)
    t.import_raw(raw.slots, raw.stash, raw.stash_count)
    probe = np.concatenate([keys, W.distinct_keys(5000, 19, exclude=keys)])
    ov, os_ = o.lookup(probe)
    for kpl in (1, 2):
        t.set_launch(1, 256, keys_per_lane=kpl)
        gv, gs = t.lookup(probe)
        assert np.array_equal(np_(gs), os_) and np.array_equal(np_(gv, np.uint32), ov), kpl


@pytest.mark.parametrize("layout", LAYOUT_NAMES)
def test_lookups_after_erases_read_the_hints(H, O, layout):
    """Lookups stop at an H0 with an empty slot only while no erase has
    emptied a slot since the table was cleared or rehashed (every bucket with
    a hint bit is then full).  At LF 0.97 many full buckets hold the hint
    bits of displaced and stashed keys; erasing 40 % of the keys opens holes
    in them, through a bulk erase and through a mixed batch.  Every remaining
    key must still be found with its value, and after a clear and after a
    resize the table works as before."""
    width = SPB[layout]
    print(f"[case] lookups after erases, layout {layout}", flush=True)
    slots = 1 << 16
    keys = W.distinct_keys(int(slots * 0.97), 61)
    vals = W.values(keys.size, 62)
    rng = np.random.default_rng(63)
    for how in ("erase", "mixed"):
        t = H.HiveTable(slots, layout=layout)
        st = np_(t.insert(keys, vals))
        ok = placed(st)
        assert t.count_exact()[1] > 0  # a populated stash
        gv, gs = t.lookup(keys)
        assert np.array_equal(np_(gs) == 4, ok)
        gone = rng.random(keys.size) < 0.4
        if how == "erase":
            es = np_(t.erase(keys[gone]))
        else:
            ty = np.full(int(gone.sum()), 2, np.uint8)
            es = np_(t.mixed(ty, keys[gone])[1])
        assert np.array_equal(es == 6, ok[gone])
        gv, gs = t.lookup(keys)
        gv, gs = np_(gv, np.uint32), np_(gs)
        keep = ok & ~gone
        assert np.array_equal(gs == 4, keep)
        assert np.array_equal(gv[keep], vals[keep])
        raw, chk = check_gpu_state(t, O, "Default2Hash")
        cv, cs = chk.lookup(keys)
        assert np.array_equal(gs, cs) and np.array_equal(gv, cv)
        # a resize rebuilds the table: lookups as before
        t.resize(slots * 2)
        gv, gs = t.lookup(keys)
        assert np.array_equal(np_(gs) == 4, keep)
        assert np.array_equal(np_(gv, np.uint32)[keep], vals[keep])
        # a clear empties it; refilled, the same keys are found again
        t.clear()
        st = np_(t.insert(keys, vals))
        ok2 = placed(st)
        gv, gs = t.lookup(keys)
        assert np.array_equal(np_(gs) == 4, ok2)
        assert np.array_equal(np_(gv, np.uint32)[ok2], vals[ok2])
        t.close()


@pytest.mark.parametrize("combine", [0, 1])
def test_duplicates_in_one_batch_and_invalid_keys(H, O, combine):
    print(f"[case] duplicate keys / key 0, combine={combine}", flush=True)
    t = H.HiveTable(1 << 12)
    t.set_launch(0, 256, combine_duplicates=combine)
    keys = np.array([5, 5, 5, 9, 0, 9, 0xFFFFFFFF], np.uint32)
    vals = np.array([1, 2, 3, 4, 5, 6, 7], np.uint32)
    st = np_(t.insert(keys, vals))
    assert st[4] == 7                                     # key 0 rejected
    assert sorted(st[[0, 1, 2]].tolist()) == [1, 2, 2]    # one insert, two updates
    assert sorted(st[[3, 5]].tolist()) == [1, 2]
    assert st[6] == 1                                     # 0xFFFFFFFF is an ordinary key
    gv, gs = t.lookup(np.array([5, 9, 0, 0xFFFFFFFF, 6], np.uint32))
    gv, gs = np_(gv, np.uint32), np_(gs)
    assert gs.tolist() == [4, 4, 5, 4, 5]
    assert gv[0] in (1, 2, 3) and gv[1] in (4, 6) and gv[3] == 7
    assert t.count_exact() == (3, 0)
    # empty batches are no-ops
    t.insert(np.empty(0, np.uint32))
    t.lookup(np.empty(0, np.uint32))
    t.erase(np.empty(0, np.uint32))
    assert t.count_exact() == (3, 0)


@pytest.mark.parametrize("combine,layout", [(0, "bucket8"), (1, "bucket8"), (0, "bucket16"), (0, "inline8")])
def test_contended_same_key_inserts_are_linearizable(H, O, combine, layout):
    # Litmus for cross-XCD coherence of the claim/update protocol: 2^20 inserts
    # over only 4096 distinct keys, shuffled so every key is written by lanes
    # of many workgroups (and so of every XCD) in the same launch.  Exactly one
    # insert per key may report INSERTED, the table must hold each key once,
    # and its value must be one of the values written for it.
    width = SPB[layout]
    print(f"[case] contended same-key inserts, combine={combine} layout={layout}", flush=True)
    rng = np.random.default_rng(81)
    base = W.distinct_keys(4096, 82)
    keys = np.tile(base, 256)
    rng.shuffle(keys)
    vals = (keys ^ np.uint32(0x5bd1e995)).astype(np.uint32)  # value is a function of the key
    t = H.HiveTable(1 << 16, layout=layout)
    t.set_launch(0, 1024, combine_duplicates=combine)
    st = np_(t.insert(keys, vals))
    assert ((st == 1) | (st == 3)).sum() == base.size
    assert (st == 2).sum() == keys.size - base.size
    raw, _ = check_gpu_state(t, O, "Default2Hash")
    kv = raw.live_kv()
    assert np.array_equal(np.sort((kv & 0xFFFFFFFF).astype(np.uint32)), np.sort(base))
    assert np.array_equal((kv >> 32).astype(np.uint32), (kv & 0xFFFFFFFF).astype(np.uint32) ^ np.uint32(0x5bd1e995))
    # same with a table small enough that most keys take the locked slow path
    t = H.HiveTable(1 << 9, layout=layout)
    t.set_launch(0, 1024, combine_duplicates=combine)
    small = base[:600]
    keys = np.tile(small, 64)
    rng.shuffle(keys)
    st = np_(t.insert(keys, keys + 1))
    raw, _ = check_gpu_state(t, O, "Default2Hash")
    kv = raw.live_kv()
    k_in = np.sort((kv & 0xFFFFFFFF).astype(np.uint32))
    assert np.unique(k_in).size == k_in.size
    assert ((st == 1) | (st == 3)).sum() == k_in.size


@pytest.mark.parametrize("combine,layout", [(0, "bucket8"), (1, "bucket8"), (0, "bucket16"), (1, "bucket16"),
                                            (0, "inline8"), (1, "inline8")])
def test_mixed_batches_match_phased_oracle(H, O, combine, layout):
    width = SPB[layout]
    print(f"[case] mixed batches 50/40/10, combine={combine} layout={layout}", flush=True)
    slots = 1 << 18
    pre = W.distinct_keys(slots // 2, 31)
    o = O.Oracle(slots // width, slots_per_bucket=width)
    o.insert(pre)
    t = H.HiveTable(slots, layout=layout)
    t.set_launch(0, 512, combine_duplicates=combine)
    t.insert(pre)
    conflicted = set()
    sizes = W.log_uniform_sizes(12, 1 << 8, 1 << 15, 32)
    for types, keys, vals in W.mixed_batches(pre, 1 << 19, sizes, seed=33):
        ov, os_ = o.mixed(types, keys, vals)
        gv, gs = t.mixed(types, keys, vals)
        gv, gs = np_(gv, np.uint32), np_(gs)
        uniq, cnt = np.unique(keys, return_counts=True)
        conflicted.update(uniq[cnt > 1].tolist())
        # one admissible outcome: the key occurs once in this batch and never
        # had concurrent inserts in an earlier one (whose final value is open)
        single = np.isin(keys, uniq[cnt == 1]) & ~np.isin(keys, np.array(sorted(conflicted), np.uint32))
        norm = lambda s: np.where(s == 3, 1, s)
        assert np.array_equal(norm(gs)[single], norm(os_)[single])
        assert np.array_equal(gv[single], ov[single])
    raw, _ = check_gpu_state(t, O, "Default2Hash")
    a, b = raw.live_kv(), o.dump()
    assert np.array_equal(np.sort(a & 0xFFFFFFFF), np.sort(b & 0xFFFFFFFF))
    ka = dict(zip((a & 0xFFFFFFFF).tolist(), (a >> 32).tolist()))
    for kv in b.tolist():
        if (kv & 0xFFFFFFFF) not in conflicted:
            assert ka[kv & 0xFFFFFFFF] == kv >> 32


def test_big_mixed_batches_split_by_type(H, O):
    # batches of >= 2^18 ops are split by op type before their phase kernels
    print("[case] big mixed batches (type-split path)", flush=True)
    slots = 1 << 20
    pre = W.distinct_keys(slots // 2, 35)
    sizes = np.array([300_000, 1 << 18, 5_000, 400_000])
    types, keys, vals, off = W.mixed_stream(pre, 1 << 22, sizes, seed=36)
    rng = np.random.default_rng(37)
    types[rng.random(types.size) < 0.01] = 4  # NONE ops: NOOP, value 0
    clean = _clean_ops(types, keys, off)
    o = O.Oracle(slots // 8)
    o.insert(pre)
    ov = np.empty(keys.size, np.uint32)
    os_ = np.empty(keys.size, np.uint8)
    for b in range(len(sizes)):
        lo, hi = int(off[b]), int(off[b + 1])
        ov[lo:hi], os_[lo:hi] = o.mixed(types[lo:hi], keys[lo:hi], vals[lo:hi])
    norm = lambda s: np.where(s == 3, 1, s)
    keyset = np.sort(o.dump() & 0xFFFFFFFF)
    for how, split, kpl in (("mixed", 1, 1), ("launches", 1, 1), ("graph", 1, 1), ("mixed", 0, 1), ("mixed", 1, 2),
                            ("launches", 0, 4)):
        t = H.HiveTable(slots)
        t.set_launch(0, 256, split_by_type=split)
        t.set_launch(1, 256, keys_per_lane=kpl)  # the lookup phase with several keys per lane
        t.insert(pre)
        if how == "mixed":
            outs = [t.mixed(types[int(off[b]):int(off[b + 1])], keys[int(off[b]):int(off[b + 1])],
                            vals[int(off[b]):int(off[b + 1])]) for b in range(len(sizes))]
            gv = np.concatenate([np_(v, np.uint32) for v, _ in outs])
            gs = np.concatenate([np_(s_) for _, s_ in outs])
        else:
            gv, gs, done, _ = t.mixed_stream(types, keys, vals, off, 0 if how == "launches" else 1)
            gv, gs = np_(gv, np.uint32), np_(gs)
        assert (gs[types == 4] == 8).all() and (gv[types == 4] == 0).all(), how
        assert np.array_equal(norm(gs)[clean], norm(os_)[clean]), how
        assert np.array_equal(gv[clean], ov[clean]), how
        live = np.sort(np_(t.dump()).view(np.uint64) & np.uint64(0xFFFFFFFF))
        assert np.array_equal(live, keyset), how
        t.close()


@pytest.mark.parametrize("layout", LAYOUT_NAMES)
def test_resize_growth_sequence(H, O, layout):
    # cfg3 shape at small scale: grow x2 whenever the next batch would exceed LF 0.9
    width = SPB[layout]
    print(f"[case] resize sequence, layout {layout}", flush=True)
    slots = 1 << 12
    t = H.HiveTable(slots, layout=layout)
    o = O.Oracle(slots // width, slots_per_bucket=width)
    all_keys = W.distinct_keys(64 * 512, 41)
    for i in range(64):
        batch = all_keys[i * 512:(i + 1) * 512]
        if (i + 1) * 512 > 0.9 * slots:
            slots *= 2
            t.resize(slots)
            assert o.resize(slots // width) == 0
            raw, _ = check_gpu_state(t, O, "Default2Hash")
            assert np.array_equal(raw.live_kv(), o.dump())
        assert placed(np_(t.insert(batch))).all()
        o.insert(batch)
        lk = all_keys[np.random.default_rng(i).integers(0, (i + 1) * 512, size=512)]
        gv, gs = t.lookup(lk)
        assert (np_(gs) == 4).all() and np.array_equal(np_(gv, np.uint32), lk + 1)
    assert t.num_buckets == slots // width
    raw, _ = check_gpu_state(t, O, "Default2Hash")
    assert np.array_equal(raw.live_kv(), o.dump())
    with pytest.raises(H.HiveError):
        t.resize(1 << 10)                 # cannot hold the content: table unchanged
    assert t.num_buckets == slots // width


@pytest.mark.parametrize("mode", [0, 1])
def test_insert_sequence_with_growth(H, O, mode):
    # cfg3 shape at small scale through hive_insert_sequence: plain per-batch
    # calls (host-side resize decisions) or one hipGraph with the rehashes inside
    print(f"[case] insert sequence mode={mode}", flush=True)
    nbatch, bsz = 48, 512
    keys = W.distinct_keys(nbatch * bsz, 61)
    vals = W.values(keys.size, 62)
    rng = np.random.default_rng(63)
    lk = np.concatenate([keys[rng.integers(0, (b + 1) * bsz, size=256)] for b in range(nbatch)])
    import torch
    st = torch.empty(keys.size, dtype=torch.uint8, device="cuda")
    ls = torch.empty(lk.size, dtype=torch.uint8, device="cuda")
    t = H.HiveTable(1 << 12)
    rep = t.insert_sequence(keys, bsz, vals=vals, lookup_keys=lk, lookups_per_batch=256, max_load_factor=0.9,
                            growth=2.0, mode=mode, status=st, lookup_status=ls)
    print(f"[case] insert sequence report {rep}", flush=True)
    assert rep["resizes"] == 3 and rep["final_slots"] == 1 << 15 and rep["rehash_failed"] == 0
    assert placed(np_(st)).all() and (np_(ls) == 4).all()
    assert t.num_buckets == (1 << 15) // 8 and t.stats()["capacity"] == 1 << 15
    o = O.Oracle((1 << 15) // 8)
    o.insert(keys, vals)
    assert np.array_equal(np.sort(np_(t.dump()).view(np.uint64)), np.sort(o.dump()))
    check_gpu_state(t, O, "Default2Hash")
    gv, gs = t.lookup(keys)
    assert (np_(gs) == 4).all() and np.array_equal(np_(gv, np.uint32), vals)


def test_dump_is_the_content_multiset(H, O):
    print("[case] dump", flush=True)
    nb = 1 << 10
    keys = W.distinct_keys(int(nb * 8 * 0.97), 57)
    o = O.Oracle(nb, "MurmurCityHash")
    o.insert(keys, W.values(keys.size, 58))
    raw = o.export()
    t = H.HiveTable(nb * 8, "MurmurCityHash")
    t.import_raw(raw.slots, raw.stash, raw.stash_count)
    assert o.count_stash() > 0
    d = np_(t.dump()).view(np.uint64)
    assert np.array_equal(np.sort(d), np.sort(o.dump()))
    with pytest.raises(H.HiveError):
        t.dump(cap=10)                    # HIVE_ERANGE: nothing written
    t.clear()
    assert t.dump().numel() == 0


@pytest.mark.parametrize("src_slots,n_keys,dst_slots", [
    (1 << 22, 3_850_000, 9_000_000),   # LF 0.92 with a stash; 1 125 000 buckets: 3 row chunks, a partial row
    (1 << 9, 200, 296),                # 37 destination buckets: one partial row
])
def test_resize_keeps_content_at_odd_sizes(H, src_slots, n_keys, dst_slots):
    """Resize moves every entry, bucket slots and stash slots, whatever the
    bucket count: the content multiset is unchanged and every key is found
    with its value."""
    print(f"[case] resize {src_slots} -> {dst_slots} slots, {n_keys} keys", flush=True)
    keys = W.distinct_keys(n_keys, 808)
    vals = W.values(n_keys, 809)
    t = H.HiveTable(src_slots)
    st = np_(t.insert(keys, vals))
    ok = placed(st)  # a key fails only with its canonical stash full (placement outcome)
    assert (~ok).sum() <= 8 and (st[~ok] == 0).all()
    keys, vals = keys[ok], vals[ok]
    before = np.sort(np_(t.dump()))
    stash = t.stats()["stash_live"]
    print(f"[case] stash entries before: {stash}", flush=True)
    if n_keys > 1000:
        assert stash > 0
    t.resize(dst_slots)
    assert t.num_buckets == dst_slots // 8
    assert np.array_equal(np.sort(np_(t.dump())), before)
    gv, gs = t.lookup(keys)
    assert (np_(gs) == 4).all() and np.array_equal(np_(gv, np.uint32), vals)
    t.close()


@pytest.mark.parametrize("layout", LAYOUT_NAMES)
def test_resize_and_layout_change_keep_content_with_a_stash(H, O, layout):
    """A table of the given layout at LF 0.93 (a populated stash) is resized
    to an odd slot count, then re-laid out in each other layout and back:
    every time the content multiset is unchanged, every key is found with its
    value and the invariants hold."""
    print(f"[case] resize and layout change from layout {layout}", flush=True)
    slots = 1 << 20
    keys = W.distinct_keys(int(slots * 0.93), 901)
    vals = W.values(keys.size, 902)
    t = H.HiveTable(slots, "MurmurCityHash", layout=layout)
    st = np_(t.insert(keys, vals))
    ok = placed(st)
    before = np.sort(np_(t.dump()).view(np.uint64))
    assert t.stats()["stash_live"] > 0
    t.resize(1_300_000)
    others = [x for x in LAYOUT_NAMES if x != layout]
    for step, lay in enumerate([layout] + others + [layout]):
        if step:
            t.set_layout(lay)
        assert t.layout == H.LAYOUTS[lay]
        assert t.num_buckets == 1_300_000 // SPB[lay]
        assert np.array_equal(np.sort(np_(t.dump()).view(np.uint64)), before)
        gv, gs = t.lookup(keys)
        assert np.array_equal(np_(gs) == 4, ok) and np.array_equal(np_(gv, np.uint32)[ok], vals[ok])
        check_gpu_state(t, O, "MurmurCityHash")
    t.close()


def test_stash_drain_and_reinsert(H, O):
    print("[case] stash drain / reinsert", flush=True)
    nb = 1 << 10
    keys = W.distinct_keys(int(nb * 8 * 0.98), 51)
    o = O.Oracle(nb)
    o.insert(keys)
    raw = o.export()
    t = H.HiveTable(nb * 8)
    t.import_raw(raw.slots, raw.stash, raw.stash_count)
    n_stash = o.count_stash()
    assert n_stash > 0
    before = gpu_raw(t, O).live_kv()
    drained = np.sort(np_(t.stash_drain()).view(np.uint64))
    assert drained.size == n_stash
    assert t.stats()["stash_live"] == 0 and t.count_exact()[1] == 0
    t.insert((drained & 0xFFFFFFFF).astype(np.uint32), (drained >> 32).astype(np.uint32))
    assert np.array_equal(gpu_raw(t, O).live_kv(), before)
    # reinsert path: import again, free some bucket space, reinsert the stash
    t2 = H.HiveTable(nb * 8)
    t2.import_raw(raw.slots, raw.stash, raw.stash_count)
    t2.erase(keys[:2000])
    o.erase(keys[:2000])
    n = t2.stash_reinsert()
    assert n == o.count_stash()
    got = gpu_raw(t2, O).live_kv()
    assert np.array_equal(got, o.dump())
    check_gpu_state(t2, O, "Default2Hash")


def test_reference_record_kernels_and_dispatch(H, O):
    print("[case] reference Operation records / dispatch pair", flush=True)
    slots = 1 << 14
    keys = W.distinct_keys(int(slots * 0.8), 61)
    ins = H.make_ops(np.zeros(keys.size, np.int32), keys)
    res, ms = H.dispatch(ins, slots, num_iterations=2)
    o = O.Oracle(slots // 8)
    expect = placed(o.insert(keys)).astype(np.uint64)
    assert np.array_equal(res, expect) and ms > 0
    look = H.make_ops(np.ones(keys.size, np.int32), keys)
    res, _ = H.dispatch(look, slots, num_iterations=3)
    assert (res == 1).all()
    types = np.random.default_rng(62).choice([0, 1, 2], size=keys.size, p=[0.5, 0.4, 0.1]).astype(np.int32)
    mixed = H.make_ops(types, keys)
    res, _ = H.dispatch(mixed, slots, num_iterations=1)
    # on an empty table with distinct keys: lookups/deletes miss, inserts succeed
    assert np.array_equal(res, (types == 0).astype(np.uint64))
    with pytest.raises(H.HiveError):
        H.dispatch(ins, slots, policy="NoSuchHash")
    pre = H.make_ops(np.zeros(keys.size, np.int32), keys)
    wl = H.make_ops(np.ones(1000, np.int32), keys[:1000])
    res, _ = H.dispatch_ycsb(pre, wl, slots, num_iterations=2)
    assert (res == 1).all()
    t = H.HiveTable(slots)
    r = np_(t.insert_ops(ins))
    assert np.array_equal(r.astype(np.uint64), expect)
    r = np_(t.lookup_ops(look))
    assert (r == 1).all()
    r = np_(t.mixed_ops(H.make_ops(np.full(10, 2, np.int32), keys[:10])))
    assert (r == 1).all()
    assert not (np_(t.lookup(keys[:10])[1]) == 4).any()


def test_launch_shapes_give_identical_results(H, O):
    """Every launch shape gives the same lookup results and content.  A key
    may fail in one shape and not another only when its canonical stash is
    full (which keys a full bucket stashes is a placement outcome); such
    keys are left out of the comparison."""
    print("[case] launch shapes", flush=True)
    slots = 1 << 16
    keys = W.distinct_keys(int(slots * 0.85), 71)
    lk, _ = W.lookup_set(keys, keys.size, 0.5, 72)
    runs = []
    for blk, inl, comb in ((64, 0, 0), (256, 0, 0), (1024, 0, 0), (64, 1, 0), (512, 1, 0), (256, 0, 1), (1024, 1, 1),
                           (256, 2, 0), (1024, 2, 0), (64, 2, 0)):
        t = H.HiveTable(slots)
        for op in (0, 1, 2):  # inline_slow_path 2 (a second launch) is for inserts only
            t.set_launch(op, blk, inline_slow_path=inl if op == 0 or inl < 2 else 0, combine_duplicates=comb)
        assert t.get_launch(0) == (blk, 1, inl, comb, 1, 1)
        st = np_(t.insert(keys))
        raw = gpu_raw(t, O)
        failed = keys[~placed(st)]
        for k in failed.tolist():
            assert raw.stash_count[O.bucket("Default2Hash", 0, k, slots // 8)] >= 8
        gv, gs = t.lookup(lk)
        runs.append((np_(gv), np_(gs), np.sort(raw.live_kv()), failed))
        t.close()
    anyfail = np.unique(np.concatenate([r[3] for r in runs]))
    assert anyfail.size <= 16
    keep = ~np.isin(lk, anyfail)
    ref = runs[0]
    for gv, gs, live, _ in runs[1:]:
        assert np.array_equal(gv[keep], ref[0][keep]) and np.array_equal(gs[keep], ref[1][keep])
        drop = lambda x: x[~np.isin((x & np.uint64(0xFFFFFFFF)).astype(np.uint32), anyfail)]
        assert np.array_equal(drop(live), drop(ref[2]))


@pytest.mark.parametrize("policy,lf,layout,slow", [("Default2Hash", 0.9, "bucket8", 0),
                                                  ("MurmurCityHash", 0.95, "bucket8", 0),
                                                  ("Default2Hash", 1.0, "bucket8", 0), ("Default2Hash", 0.9, "bucket16", 0),
                                                  ("MurmurCityHash", 1.0, "bucket16", 0),
                                                  ("Default2Hash", 0.9, "inline8", 0),
                                                  ("MurmurCityHash", 1.0, "inline8", 0),
                                                  ("Default2Hash", 0.9, "inline8", 2),
                                                  ("Default2Hash", 1.0, "inline8", 2)])
def test_full_size_properties(H, policy, lf, layout, slow):
    """2^26 slots at the sweep's load factors: size-independent properties
    (every placed key found with its value, absent keys not found, exact
    counts, no key stored twice, re-insert = update); slow = 2: the insert
    slow path in a second launch."""
    import torch
    width = SPB[layout]
    print(f"[case] 2^26 slots LF {lf} {policy} layout {layout} slow path {slow} full size", flush=True)
    slots = 1 << 26
    n = W.keys_for_lf(slots, lf)
    keys = W.distinct_keys(n, 44)
    vals = W.values(n, 45)
    absent = W.distinct_keys(1 << 22, 44, start=n + (n >> 6) + 64)
    t = H.HiveTable(slots, policy, layout=layout)
    t.set_launch(0, 256, inline_slow_path=slow)
    st = t.insert(keys, vals)
    ok = placed(np_(st))
    assert ok.mean() > 0.9999
    assert t.stats()["slow_path_ops"] > 0
    assert (np_(st)[~ok] == 0).all()  # the only other outcome: FAILED (stash full)
    kv = np_(t.dump()).view(np.uint64)
    assert kv.size == ok.sum() and np.unique(kv & np.uint64(0xFFFFFFFF)).size == kv.size
    dk = torch.from_numpy(keys.view(np.int32)).cuda()
    gv, gs = t.lookup(dk)
    gs = np_(gs)
    assert np.array_equal(gs == 4, ok)
    assert np.array_equal(np_(gv, np.uint32)[ok], vals[ok])
    gv, gs = t.lookup(absent)
    assert (np_(gs) == 5).all()
    a, b = t.count_exact()
    s = t.stats()
    assert a + b == ok.sum() and (a, b) == (s["table_live"], s["stash_live"])
    # re-inserting the same keys updates in place: content count unchanged
    st2 = np_(t.insert(dk, torch.from_numpy(vals.view(np.int32)).cuda()))
    assert (st2[ok] == 2).all()
    newly = int(np.isin(st2[~ok], (1, 3)).sum())
    assert sum(t.count_exact()) == a + b + newly


def _clean_ops(types, keys, off):
    """Ops with one admissible outcome: key unique in its batch and never
    inserted twice concurrently in an earlier batch."""
    nb = len(off) - 1
    batch = np.repeat(np.arange(nb), np.diff(off).astype(np.int64))
    comb = (batch.astype(np.uint64) << np.uint64(32)) | keys.astype(np.uint64)
    _, inv, cnt = np.unique(comb, return_inverse=True, return_counts=True)
    dup = cnt[inv] > 1
    first_dup = {}
    for k, b in zip(keys[dup].tolist(), batch[dup].tolist()):
        if k not in first_dup or b < first_dup[k]:
            first_dup[k] = b
    fd = np.array([first_dup.get(k, nb) for k in keys.tolist()])
    return (~dup) & (batch < fd)


@pytest.mark.parametrize("layout", LAYOUT_NAMES)
def test_stream_modes_agree_with_phased_oracle(H, O, layout):
    width = SPB[layout]
    print(f"[case] mixed stream: launches / graph / persistent, layout {layout}", flush=True)
    slots = 1 << 18
    pre = W.distinct_keys(slots // 2, 91)
    sizes = W.log_uniform_sizes(40, 1 << 6, 1 << 14, 92)
    sizes[5] = 0  # an empty batch
    types, keys, vals, off = W.mixed_stream(pre, 1 << 19, sizes, seed=93)
    o = O.Oracle(slots // width, slots_per_bucket=width)
    o.insert(pre)
    ov = np.empty(keys.size, np.uint32)
    os_ = np.empty(keys.size, np.uint8)
    for b in range(len(sizes)):
        lo, hi = int(off[b]), int(off[b + 1])
        ov[lo:hi], os_[lo:hi] = o.mixed(types[lo:hi], keys[lo:hi], vals[lo:hi])
    clean = _clean_ops(types, keys, off)
    norm = lambda s: np.where(s == 3, 1, s)
    keyset = np.sort(o.dump() & 0xFFFFFFFF)
    prev = None
    # (mode, plain-kernel slow path inline, persistent tile ops, blocks per CU, persistent inline, window ops)
    shapes = [(0, 0, 0, 0, 0, 0), (0, 1, 0, 0, 0, 0), (1, 0, 0, 0, 0, 0), (2, 0, 0, 0, 0, 0), (2, 0, 256, 8, 1, 0),
              (2, 0, 1024, 2, 0, 0), (2, 0, 0, 4, 1, 0), (3, 0, 0, 0, 0, (1 << 10, 0)), (3, 1, 0, 0, 0, (1 << 13, 0)),
              (3, 0, 0, 0, 0, (1 << 16, 0)), (3, 0, 0, 0, 0, (1 << 13, 1 << 9))]
    for mode, inl, tile, bpc, sinl, win in shapes:
        print(f"[case] stream shape mode={mode} inline={inl} tile={tile} bpc={bpc} stream_inline={sinl} "
              f"window={win}", flush=True)
        t = H.HiveTable(slots, layout=layout)
        for op in (0, 1, 2):
            t.set_launch(op, 256, inline_slow_path=inl)
        t.set_stream_shape(tile, bpc, sinl)
        t.set_stream_window(*(win or (0, 0)))
        t.insert(pre)
        gv, gs, done, setup = t.mixed_stream(types, keys, vals, off, mode)
        gv, gs = np_(gv, np.uint32), np_(gs)
        assert done.size == len(sizes) and (np.diff(done) >= 0).all() and done[-1] > 0
        assert np.array_equal(norm(gs)[clean], norm(os_)[clean]), mode
        assert np.array_equal(gv[clean], ov[clean]), mode
        raw, _ = check_gpu_state(t, O, "Default2Hash")
        assert np.array_equal(np.sort(raw.live_kv() & 0xFFFFFFFF), keyset), mode
        if prev is not None:
            assert np.array_equal(norm(gs)[clean], norm(prev)[clean])
        prev = gs


def test_stream_autotune_leaves_table_unchanged(H, O):
    print("[case] stream autotune", flush=True)
    slots = 1 << 16
    pre = W.distinct_keys(slots // 2, 95)
    sizes = W.log_uniform_sizes(12, 1 << 6, 1 << 12, 96)
    types, keys, vals, off = W.mixed_stream(pre, 1 << 18, sizes, seed=97)
    t = H.HiveTable(slots)
    t.insert(pre)
    before = gpu_raw(t, O).live_kv()
    windows = tuple(1 << e for e in range(15, 23))
    coalesce = tuple(1 << e for e in range(10, 20))
    # windows 2^15..2^22 x coalescing below 1/32..1/4 of them (at most 2^19):
    # 31 pairs, of which the 19 with windows of >= 2^18 ops are also timed
    # with type compaction
    n_win = 31 + 19
    # default: no persistent kernel; the choice is bounded by the launches' p99
    res = t.autotune_stream(types, keys, vals, off)
    assert res["candidates"] == 8 + 1 + n_win  # launches (no big batches: split only), graph, windows
    assert res["window_ops"] in windows and res["windowed_ms"] > 0 and res["windowed_p99_us"] > 0
    assert res["window_batch_ops"] in coalesce and res["window_batch_ops"] <= res["window_ops"] // 4
    assert res["best_mode"] in (0, 1, 3) and res["persistent_ms"] == 0
    assert res["block_threads"] in (128, 256, 512, 1024) and res["inline_slow_path"] in (0, 1)
    assert min(res["launches_ms"], res["graph_ms"]) > 0 and res["launches_p99_us"] > 0
    assert res["p99_bound_us"] == pytest.approx(res["launches_p99_us"] * 1.1)
    if res["bounded_window_ops"]:
        assert res["bounded_windowed_p99_us"] <= res["p99_bound_us"]
        assert res["bounded_windowed_ms"] >= res["windowed_ms"]
    assert set(res["windowed_ms_by_size"]) == set(windows)
    assert min(res["windowed_ms_by_size"].values()) == pytest.approx(res["windowed_ms"])
    assert all(v > 0 for v in res["windowed_p99_us_by_size"].values())
    chosen_p99 = {0: res["launches_p99_us"], 1: res["graph_p99_us"], 3: res["bounded_windowed_p99_us"]}
    assert chosen_p99[res["best_mode"]] <= res["p99_bound_us"] * (1 + 1e-9)
    assert np.array_equal(gpu_raw(t, O).live_kv(), before)
    assert t.get_launch(0) == (res["block_threads"], 1, res["inline_slow_path"], 0, res["split_by_type"], 1)
    # with the persistent kernel and no latency bound: the fastest form
    res = t.autotune_stream(types, keys, vals, off, try_persistent=True, max_p99_us=0.0)
    assert res["candidates"] == 8 + 1 + 24 + n_win and res["p99_bound_us"] == 0
    assert res["tile_ops"] in (0, 256, 1024, 4096) and res["blocks_per_cu"] in (2, 4, 8)
    assert res["stream_inline"] in (0, 1) and res["persistent_ms"] > 0
    assert res["bounded_window_ops"] == res["window_ops"] and res["bounded_windowed_ms"] == res["windowed_ms"]
    best = min(res["launches_ms"], res["graph_ms"], res["persistent_ms"], res["windowed_ms"])
    assert [res["launches_ms"], res["graph_ms"], res["persistent_ms"], res["windowed_ms"]][res["best_mode"]] == best
    assert np.array_equal(gpu_raw(t, O).live_kv(), before)


@pytest.mark.parametrize("policy,layout", [("Default2Hash", "bucket8"), ("TripleHash", "bucket8"),
                                           ("Default2Hash", "bucket16"), ("TripleHash", "inline8")])
def test_windowed_stream_orders_each_key_by_batch(H, O, policy, layout):
    """Windows of many small batches in which the same keys recur in every
    batch (a hot set in all of them, cold keys from a small space): each
    key's ops must see exactly the batches before theirs (chains 63 batches
    long), against the oracle run batch by batch.  No key repeats inside a
    batch, so every op has one admissible outcome.  A few ops are NONE /
    ATOMIC_INC or carry key 0, which a window must answer like a batch."""
    width = SPB[layout]
    print(f"[case] windowed stream: recurring keys, {policy}, layout {layout}", flush=True)
    slots = 1 << 14
    rng = np.random.default_rng(7)
    pre = W.distinct_keys(slots // 4, 71, bits=13)
    hot = W.distinct_keys(48, 72, bits=13)
    parts, sizes = [], []
    for b in range(64):
        k = np.unique(np.concatenate([hot, rng.integers(1, 1 << 13, size=220, dtype=np.uint64).astype(np.uint32)]))
        k = k[rng.permutation(k.size)]
        parts.append(k)
        sizes.append(k.size)
    keys = np.concatenate(parts).astype(np.uint32)
    types = rng.choice(3, size=keys.size, p=[0.4, 0.4, 0.2]).astype(np.uint8)
    vals = rng.integers(0, 2**32, size=keys.size, dtype=np.uint64).astype(np.uint32)
    odd = rng.random(keys.size)
    types[odd < 0.02] = rng.choice([3, 4], size=int((odd < 0.02).sum())).astype(np.uint8)
    keys[(odd >= 0.02) & (odd < 0.03)] = 0
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    o = O.Oracle(slots // width, policy, slots_per_bucket=width)
    o.insert(pre)
    ov = np.empty(keys.size, np.uint32)
    os_ = np.empty(keys.size, np.uint8)
    for b in range(len(sizes)):
        lo, hi = int(off[b]), int(off[b + 1])
        ov[lo:hi], os_[lo:hi] = o.mixed(types[lo:hi], keys[lo:hi], vals[lo:hi])
    norm = lambda s: np.where(s == 3, 1, s)
    keyset = np.sort(o.dump() & 0xFFFFFFFF)
    for win, split in ((1 << 10, 1), (1 << 12, 0), (1 << 14, 1)):
        print(f"[case] windowed stream window={win} split={split}", flush=True)
        t = H.HiveTable(slots, policy, layout=layout)
        t.set_launch(0, 256, split_by_type=split)
        t.set_stream_window(win, 1 << 10)
        t.insert(pre)
        gv, gs, done, _ = t.mixed_stream(types, keys, vals, off, 3)
        gv, gs = np_(gv, np.uint32), np_(gs)
        assert (np.diff(done) >= 0).all() and done[-1] > 0
        assert np.array_equal(norm(gs), norm(os_)), win
        assert np.array_equal(gv, ov), win
        raw, _ = check_gpu_state(t, O, policy)
        assert np.array_equal(np.sort(raw.live_kv() & 0xFFFFFFFF), keyset), win


def test_bulk_autotune_leaves_table_unchanged(H, O):
    """The bulk autotuner times insert shapes on a scratch copy and lookup
    layouts on the table itself; the table's state is unchanged and the
    chosen shapes are stored and give the same lookup results."""
    print("[case] bulk autotune", flush=True)
    slots = 1 << 18
    keys = W.distinct_keys(int(slots * 0.9), 81)
    t = H.HiveTable(slots)
    t.insert(keys[: keys.size // 2])
    before = gpu_raw(t, O).live_kv()
    lk, _ = W.lookup_set(keys, keys.size // 2, 0.5, 82)
    ref_v, ref_s = (np_(x) for x in t.lookup(lk))
    res = t.autotune_bulk(keys[keys.size // 2:], None, lk)
    assert res["candidates"] == 9 + 15
    assert res["insert_block_threads"] in (256, 512, 1024) and res["insert_inline_slow_path"] in (0, 1, 2)
    assert res["lookup_block_threads"] in (256, 512, 1024) and res["lookup_lanes_per_key"] in (1, 4, 8)
    assert res["lookup_keys_per_lane"] in (1, 2, 4)
    assert res["lookup_lanes_per_key"] == 1 or res["lookup_keys_per_lane"] == 1
    assert res["load_factor"] == pytest.approx(0.45, abs=0.005)
    best = min(min(res["lookup_ms_by_lanes"].values()), min(res["lookup_ms_by_keys"].values()))
    assert best == pytest.approx(res["lookup_ms"])
    assert res["lookup_ms_by_keys"][1] == pytest.approx(res["lookup_ms_by_lanes"][1])
    assert np.array_equal(gpu_raw(t, O).live_kv(), before)
    lp = t.get_launch(1)
    assert (lp[0], lp[1], lp[5]) == (res["lookup_block_threads"], res["lookup_keys_per_lane"],
                                     res["lookup_lanes_per_key"])
    assert t.get_launch(0)[0] == res["insert_block_threads"]
    v, st = (np_(x) for x in t.lookup(lk))
    assert np.array_equal(v, ref_v) and np.array_equal(st, ref_s)


@pytest.mark.parametrize("start", ["bucket8", "inline8"])
def test_layout_autotune_keeps_content(H, O, start):
    """The layout tuner times the samples in every layout and applies the
    fastest: every layout is timed, the chosen one is the table's layout
    afterwards, its total time is the least, and the live entries and lookup
    results are those of before."""
    print(f"[case] layout autotune from {start}", flush=True)
    slots = 1 << 18
    keys = W.distinct_keys(int(slots * 0.9), 83)
    t = H.HiveTable(slots, layout=start)
    t.insert(keys[: keys.size // 2])
    before = gpu_raw(t, O).live_kv()
    lk, _ = W.lookup_set(keys, keys.size // 2, 0.5, 84)
    ref_v, ref_s = (np_(x) for x in t.lookup(lk))
    res = t.autotune_layout(keys[keys.size // 2:], None, lk)
    assert res["candidates"] == 3
    assert t.layout == H.LAYOUTS[res["layout"]] and t.slots_per_bucket == SPB[res["layout"]]
    assert t.stats()["capacity"] == slots
    tot = {n: res["insert_ms"][n] + res["lookup_ms"][n] for n in LAYOUT_NAMES}
    assert all(v > 0 for v in tot.values())
    assert tot[res["layout"]] == min(tot.values())
    raw, _ = check_gpu_state(t, O, "Default2Hash")
    assert np.array_equal(raw.live_kv(), before)
    v, st = (np_(x) for x in t.lookup(lk))
    assert np.array_equal(v, ref_v) and np.array_equal(st, ref_s)


def test_windowed_stream_thousands_of_tiny_batches(H, O):
    """5000 batches of 2 ops over 4 keys: windows reach their 4096-batch cap,
    and each key has a deferred op in most batches of its window, so its
    chain has about a thousand written batches (resolved by walking back);
    results equal the oracle run batch by batch (no key repeats inside a
    batch)."""
    print("[case] windowed stream: 5000 two-op batches over 4 keys", flush=True)
    rng = np.random.default_rng(17)
    nb = 5000
    keys = np.stack([rng.choice(np.arange(1, 5, dtype=np.uint32), 2, replace=False) for _ in range(nb)]).reshape(-1)
    types = rng.choice(3, size=keys.size, p=[0.4, 0.4, 0.2]).astype(np.uint8)
    vals = rng.integers(0, 2**32, size=keys.size, dtype=np.uint64).astype(np.uint32)
    off = (np.arange(nb + 1) * 2).astype(np.uint64)
    o = O.Oracle(1 << 7)
    ov = np.empty(keys.size, np.uint32)
    os_ = np.empty(keys.size, np.uint8)
    for b in range(nb):
        ov[2 * b:2 * b + 2], os_[2 * b:2 * b + 2] = o.mixed(types[2 * b:2 * b + 2], keys[2 * b:2 * b + 2],
                                                            vals[2 * b:2 * b + 2])
    t = H.HiveTable(1 << 10)
    t.set_stream_window(1 << 16, 1 << 10)
    gv, gs, done, _ = t.mixed_stream(types, keys, vals, off, 3)
    gv, gs = np_(gv, np.uint32), np_(gs)
    norm = lambda s: np.where(s == 3, 1, s)
    assert np.array_equal(norm(gs), norm(os_))
    assert np.array_equal(gv, ov)
    assert np.unique(done).size == 2  # two windows: 4096 batches, then 904
    assert np.array_equal(np.sort(gpu_raw(t, O).live_kv() & 0xFFFFFFFF), np.sort(o.dump() & 0xFFFFFFFF))
