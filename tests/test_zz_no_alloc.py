"""hive_mixed_noalloc: the mixed batch that allocates nothing (type lists in
the library's static device array) against hive_mixed, through the C-ABI.

Two calls in a row on one table pair -- 2^20 ops (the maximum), then
300,000 ops (a smaller batch, still above the 2^18-op type-list threshold,
so the array's last contents must not leak into it) -- then a small batch
that needs no lists, and a batch above the maximum (HIVE_EINVAL, nothing
run).  Keys are distinct inside each batch, so every op has one outcome:
status, value and the table content (with values) must equal hive_mixed's
and the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available
from hivehashtable_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _batch(seed, n, pool, pool_seed, call):
    """n distinct keys: half from the preload, half new (the preload's key
    stream continued, a fresh stretch per call)."""
    rng = np.random.default_rng(seed)
    old = rng.choice(pool, n // 2, replace=False)
    new = W.distinct_keys(n - old.size, pool_seed, start=pool.size + 2 + call * (1 << 21))
    keys = np.concatenate([old, new])[rng.permutation(n)].astype(np.uint32)
    types = rng.choice(np.array([0, 1, 2, 4], np.uint8), size=n, p=[0.4, 0.35, 0.2, 0.05])
    vals = W.values(n, seed + 1)
    return types, keys, vals


def test_mixed_noalloc_matches_mixed(oracle_mod):
    import torch
    import hivehashtable_amd as H
    O = oracle_mod
    L = H.lib()
    slots = 1 << 23
    pre = W.distinct_keys(1 << 21, 7001)
    ta, tb = H.HiveTable(slots), H.HiveTable(slots)
    for t in (ta, tb):
        t.set_launch(0, 256, split_by_type=1)
        t.insert(pre)
    o = O.Oracle(slots // 8)
    o.insert(pre)
    sptr = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    norm = lambda s: np.where(s == 3, 1, s)
    for call, n in enumerate((1 << 20, 300000, 5000)):
        print(f"[case] hive_mixed_noalloc call {call}: {n} ops", flush=True)
        types, keys, vals = _batch(7100 + call, n, pre, 7001, call)
        d_ty = torch.from_numpy(types).cuda()
        d_k = torch.from_numpy(keys.view(np.int32)).cuda()
        d_v = torch.from_numpy(vals.view(np.int32)).cuda()
        outs = []
        for t, fn in ((ta, L.hive_mixed), (tb, L.hive_mixed_noalloc)):
            ov = torch.full((n,), -1, dtype=torch.int32, device="cuda")
            st = torch.full((n,), 0xEE, dtype=torch.uint8, device="cuda")
            rc = fn(t.h, C.c_void_p(d_ty.data_ptr()), C.c_void_p(d_k.data_ptr()), C.c_void_p(d_v.data_ptr()), n,
                    C.c_void_p(ov.data_ptr()), C.c_void_p(st.data_ptr()), sptr)
            assert rc == 0, L.hive_last_error()
            torch.cuda.synchronize()
            outs.append((ov.cpu().numpy().view(np.uint32), st.cpu().numpy()))
        ref_v, ref_s = o.mixed(types, keys, vals)
        (av, as_), (bv, bs) = outs
        assert np.array_equal(norm(bs), norm(as_)) and np.array_equal(bv, av)
        assert np.array_equal(norm(bs), norm(ref_s)) and np.array_equal(bv, ref_v)
        ca = np.sort(ta.dump().cpu().numpy().view(np.uint64))
        cb = np.sort(tb.dump().cpu().numpy().view(np.uint64))
        assert np.array_equal(cb, ca) and np.array_equal(cb, np.sort(o.dump()))
    # above the maximum: rejected before anything runs
    n = (1 << 20) + 1
    d_ty = torch.ones(n, dtype=torch.uint8, device="cuda")
    d_k = torch.ones(n, dtype=torch.int32, device="cuda")
    st = torch.full((n,), 0xEE, dtype=torch.uint8, device="cuda")
    rc = L.hive_mixed_noalloc(tb.h, C.c_void_p(d_ty.data_ptr()), C.c_void_p(d_k.data_ptr()), None, n, None,
                              C.c_void_p(st.data_ptr()), sptr)
    torch.cuda.synchronize()
    assert rc == 1 and b"2^20" in L.hive_last_error()
    assert (st.cpu().numpy() == 0xEE).all()
    ta.close()
    tb.close()
