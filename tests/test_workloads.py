"""Workload generators: the properties the benchmark configurations state."""
import numpy as np

from hivehashtable_amd import workloads as W


def test_distinct_keys_streams_are_disjoint_and_in_range():
    a = W.distinct_keys(200_000, 9)
    b = W.distinct_keys(50_000, 9, start=a.size + 2)
    assert np.unique(a).size == a.size and np.unique(b).size == b.size
    assert np.intersect1d(a, b).size == 0
    both = np.concatenate([a, b])
    assert both.min() >= 1 and both.max() <= 0xFFFFFFFE
    c = W.distinct_keys(1000, 9, bits=27)
    assert c.min() >= 1 and c.max() <= 1 << 27 and np.unique(c).size == c.size


def test_log_uniform_sizes_bounds_and_spread():
    s = W.log_uniform_sizes(20_000, 1 << 10, 1 << 20, 3)
    assert s.min() >= 1 << 10 and s.max() <= 1 << 20
    # log-uniform: the median sits near the geometric mean 2^15
    assert 2**14 < np.median(s) < 2**16


def test_mixed_stream_deletes_draw_from_live_keys():
    live = W.distinct_keys(5000, 4)
    sizes = W.log_uniform_sizes(30, 64, 2048, 5)
    types, keys, vals, off = W.mixed_stream(live, 1 << 20, sizes, seed=6)
    assert types.size == keys.size == vals.size == int(off[-1])
    for b in range(sizes.size):
        lo, hi = int(off[b]), int(off[b + 1])
        before = np.concatenate([live, keys[:lo][types[:lo] == 0]])
        d = keys[lo:hi][types[lo:hi] == 2]
        assert np.isin(d, before).all()
    frac = np.bincount(types, minlength=3) / types.size
    assert abs(frac[0] - 0.5) < 0.03 and abs(frac[1] - 0.4) < 0.03 and abs(frac[2] - 0.1) < 0.03


def test_serving_trace_mixes():
    pre = W.distinct_keys(1 << 16, 11)
    fresh = W.distinct_keys(1 << 17, 11, start=pre.size + 2)
    for mix, (pr, pu, pn) in W.SERVE_MIXES.items():
        ty, k, v, off, arr = W.serving_trace(mix, pre, 120, 1000.0, 12, size_lo=64, size_hi=4096, fresh=fresh)
        sizes = np.diff(off.astype(np.int64))
        assert sizes.min() >= 64 and sizes.max() <= 4096 and ty.size == int(off[-1])
        assert arr[0] == 0.0 and (np.diff(arr) >= 0).all()
        # Poisson rate: offered 1000 Mops/s = 1000 ops per microsecond on average
        assert 0.5 < (ty.size / max(arr[-1], 1.0)) / 1000.0 < 2.0
        reads = (ty == 1).mean()
        assert abs(reads - pr) < 0.02
        if mix != "D":
            assert np.isin(k, pre).all()
            # Zipfian: rank-0 key is the most frequent
            vals_, counts = np.unique(k, return_counts=True)
            assert vals_[np.argmax(counts)] == pre[0]
        else:
            ins = k[ty == 0]
            assert np.unique(ins).size == ins.size and np.isin(ins, fresh).all()
            # reads skew to the latest keys: most reads hit keys inserted during the trace
            assert np.isin(k[ty == 1], fresh).mean() > 0.5
    ty, k, v, off, arr = W.serving_trace("C", pre, 10, 0, 13, size_lo=64, size_hi=128)
    assert (arr == 0).all()


def test_zipf_sampler_ranks():
    z = W.zipf_sampler(1 << 20, 0.99, 5)
    r = z(200_000)
    assert r.min() >= 0 and r.max() < 1 << 20
    c = np.bincount(r[r < 8], minlength=8)
    # frequencies fall roughly as 1/rank^theta
    assert c[0] > c[1] > c[3]
    assert 1.6 < c[0] / c[1] < 2.4
