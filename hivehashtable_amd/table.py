"""HiveTable: Python handle over the libhive.so C-ABI.

Mirrors the reference's table operations (bulk insert / lookup / delete,
mixed batches, resize, stash drain).  Inputs are device tensors (torch is
only the carrier of device memory and streams); numpy arrays are copied to
the current device first.  Every call is stream-ordered on the given torch
stream (default: torch's current stream).
"""
import ctypes as C

import numpy as np

from ._lib import (BulkTune, LayoutTune, Launch, Operation, SequenceReport, ServeConfig, ServeReport, Stats, Tune,
                   TuneOpts, LAYOUTS, POLICIES,
                   check, lib,
                   OP_INSERT,
                   OP_LOOKUP, OP_DELETE)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("HiveTable needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _stream_ptr(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _dev(x, dtype):
    """Contiguous device tensor of the given torch dtype."""
    torch = _torch()
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        x = np.ascontiguousarray(x)
        if x.dtype == np.uint32:  # keys and values are u32 bit patterns carried in int32 tensors
            x = x.view(np.int32)
        x = torch.from_numpy(x)
    if x.dtype != dtype:
        x = x.to(dtype)
    if not x.is_cuda:
        x = x.cuda()
    return x.contiguous()


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def layout_id(layout):
    return LAYOUTS[layout] if isinstance(layout, str) else int(layout)


class HiveTable:
    """A device-resident Hive hash table with an 8-slot stash per bucket.

    ``num_slots`` is the reference's ``table_size``: num_buckets = num_slots
    // slots per bucket.  ``layout``: "bucket8" (8-slot 64-byte buckets, the
    reference's bucket), "bucket16" (16-slot 128-byte buckets) or "inline8"
    (8 slots and an in-line header per 128-byte line, pooled stash).
    """

    def __init__(self, num_slots, policy="Default2Hash", layout="bucket8"):
        self.L = lib()
        self.policy = POLICIES[policy] if isinstance(policy, str) else int(policy)
        h = C.c_void_p()
        check(self.L.hive_create_ex(int(num_slots), self.policy, layout_id(layout), C.byref(h)), "hive_create_ex")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.hive_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_buckets(self):
        return int(self.L.hive_num_buckets(self.h))

    @property
    def slots_per_bucket(self):
        return int(self.stats()["slots_per_bucket"])

    @property
    def layout(self):
        return int(self.stats()["layout"])

    def set_layout(self, layout, stream=None):
        """Rehash into another layout at the same slot capacity."""
        check(self.L.hive_set_layout(self.h, layout_id(layout), _stream_ptr(stream)), "hive_set_layout")

    # ---------------------------------------------------------- bulk ops
    def clear(self, stream=None):
        check(self.L.hive_clear(self.h, _stream_ptr(stream)), "hive_clear")

    def insert(self, keys, vals=None, status=None, stream=None):
        torch = _torch()
        k = _dev(keys, torch.int32)
        v = _dev(vals, torch.int32)
        if status is None:
            status = torch.empty(k.numel(), dtype=torch.uint8, device=k.device)
        check(self.L.hive_insert(self.h, _ptr(k), _ptr(v), k.numel(), _ptr(status), _stream_ptr(stream)),
              "hive_insert")
        return status

    def lookup(self, keys, vals=None, status=None, stream=None):
        torch = _torch()
        k = _dev(keys, torch.int32)
        if vals is None:
            vals = torch.empty(k.numel(), dtype=torch.int32, device=k.device)
        if status is None:
            status = torch.empty(k.numel(), dtype=torch.uint8, device=k.device)
        check(self.L.hive_lookup(self.h, _ptr(k), k.numel(), _ptr(vals), _ptr(status), _stream_ptr(stream)),
              "hive_lookup")
        return vals, status

    def erase(self, keys, status=None, stream=None):
        torch = _torch()
        k = _dev(keys, torch.int32)
        if status is None:
            status = torch.empty(k.numel(), dtype=torch.uint8, device=k.device)
        check(self.L.hive_erase(self.h, _ptr(k), k.numel(), _ptr(status), _stream_ptr(stream)), "hive_erase")
        return status

    def mixed(self, types, keys, vals=None, out_vals=None, status=None, stream=None):
        torch = _torch()
        ty = _dev(types, torch.uint8)
        k = _dev(keys, torch.int32)
        v = _dev(vals, torch.int32)
        if out_vals is None:
            out_vals = torch.empty(k.numel(), dtype=torch.int32, device=k.device)
        if status is None:
            status = torch.empty(k.numel(), dtype=torch.uint8, device=k.device)
        check(self.L.hive_mixed(self.h, _ptr(ty), _ptr(k), _ptr(v), k.numel(), _ptr(out_vals), _ptr(status),
                                _stream_ptr(stream)), "hive_mixed")
        return out_vals, status

    def mixed_stream(self, types, keys, vals, offsets, mode=2, out_vals=None, status=None, stream=None):
        """Run consecutive mixed batches (batch b = ops[offsets[b]:offsets[b+1]]).

        mode: 0 per-batch launches, 1 captured hipGraph, 2 persistent kernel,
        3 windows of consecutive batches (set_stream_window).
        Returns (out_vals, status, batch_done_ns numpy array, setup_ms)."""
        torch = _torch()
        ty = _dev(types, torch.uint8)
        k = _dev(keys, torch.int32)
        v = _dev(vals, torch.int32)
        off = np.ascontiguousarray(offsets, np.uint64)
        nb = off.size - 1
        if out_vals is None:
            out_vals = torch.empty(k.numel(), dtype=torch.int32, device=k.device)
        if status is None:
            status = torch.empty(k.numel(), dtype=torch.uint8, device=k.device)
        done = np.zeros(max(nb, 1), np.float64)
        setup = C.c_double(0.0)
        check(self.L.hive_mixed_stream(self.h, _ptr(ty), _ptr(k), _ptr(v), off.ctypes.data_as(C.c_void_p), nb,
                                       _ptr(out_vals), _ptr(status), int(mode), done.ctypes.data_as(C.c_void_p),
                                       C.byref(setup), _stream_ptr(stream)), "hive_mixed_stream")
        return out_vals, status, done[:nb], setup.value

    def autotune_stream(self, types, keys, vals, offsets, stream=None, try_persistent=False, max_p99_us=-1.0,
                        p99_slack=0.1):
        """Time the stream forms on a scratch copy of this table; keep the
        fastest whose p99 batch service time stays within max_p99_us (< 0: the
        per-batch launches' p99 times 1 + p99_slack; 0: no bound).
        try_persistent: also time the persistent kernel."""
        torch = _torch()
        ty = _dev(types, torch.uint8)
        k = _dev(keys, torch.int32)
        v = _dev(vals, torch.int32)
        off = np.ascontiguousarray(offsets, np.uint64)
        res = Tune()
        opts = TuneOpts(int(bool(try_persistent)), float(max_p99_us), float(p99_slack))
        check(self.L.hive_autotune_stream_ex(self.h, _ptr(ty), _ptr(k), _ptr(v), off.ctypes.data_as(C.c_void_p),
                                             off.size - 1, C.byref(opts), C.byref(res), _stream_ptr(stream)),
              "hive_autotune_stream_ex")
        out = {f: getattr(res, f) for f, _ in Tune._fields_}
        out["windowed_ms_by_size"] = {1 << (15 + i): res.windowed_ms_by_size[i] for i in range(8)}
        out["windowed_p99_us_by_size"] = {1 << (15 + i): res.windowed_p99_us_by_size[i] for i in range(8)}
        return out

    def serve(self, types, keys, vals, req_offsets, arrival_us, max_batch_ops=0, max_inflight=0, max_load_factor=0.0,
              growth=0.0, hbm_reserve_bytes=0, stash_headroom=0.0, ordered=0, out_vals=None, status=None,
              stream=None, min_load_factor=0.0, shrink_load_factor=0.0):
        """Replay a request trace with continuous batching (hive_serve_trace).

        Request r = ops[req_offsets[r]:req_offsets[r+1]], arriving arrival_us[r]
        after the call starts; ordered=1: every request sees the writes of the
        requests before it, also inside one batch; min_load_factor > 0: shrink
        in place below it (to shrink_load_factor, default 0.75).  Returns
        (out_vals, status, latency_us, report)."""
        torch = _torch()
        ty = _dev(types, torch.uint8)
        k = _dev(keys, torch.int32)
        v = _dev(vals, torch.int32)
        off = np.ascontiguousarray(req_offsets, np.uint64)
        arr = np.ascontiguousarray(arrival_us, np.float64)
        nreq = off.size - 1
        if arr.size != nreq:
            raise ValueError("one arrival time per request")
        if out_vals is None:
            out_vals = torch.empty(k.numel(), dtype=torch.int32, device=k.device)
        if status is None:
            status = torch.empty(k.numel(), dtype=torch.uint8, device=k.device)
        cfg = ServeConfig(int(max_batch_ops), int(max_inflight), float(max_load_factor), float(growth),
                          int(hbm_reserve_bytes), float(stash_headroom), int(ordered), float(min_load_factor),
                          float(shrink_load_factor))
        lat = np.zeros(max(nreq, 1), np.float64)
        rep = ServeReport()
        check(self.L.hive_serve_trace(self.h, _ptr(ty), _ptr(k), _ptr(v), off.ctypes.data_as(C.c_void_p),
                                      arr.ctypes.data_as(C.c_void_p), nreq, _ptr(out_vals), _ptr(status),
                                      C.byref(cfg), lat.ctypes.data_as(C.c_void_p), C.byref(rep),
                                      _stream_ptr(stream)), "hive_serve_trace")
        return out_vals, status, lat[:nreq], {f: getattr(rep, f) for f, _ in ServeReport._fields_}

    def insert_sequence(self, keys, batch_keys, vals=None, lookup_keys=None, lookups_per_batch=0,
                        max_load_factor=0.9, growth=2.0, mode=0, status=None, lookup_status=None, stream=None):
        """Insert keys in batches with in-place growth (hive_insert_sequence);
        mode 0 plain per-batch calls, 1 one captured hipGraph.  Returns the report."""
        torch = _torch()
        k = _dev(keys, torch.int32)
        v = _dev(vals, torch.int32)
        lk = _dev(lookup_keys, torch.int32)
        rep = SequenceReport()
        check(self.L.hive_insert_sequence(self.h, _ptr(k), _ptr(v), k.numel(), int(batch_keys), _ptr(lk),
                                          int(lookups_per_batch), float(max_load_factor), float(growth), int(mode),
                                          _ptr(status), _ptr(lookup_status), C.byref(rep), _stream_ptr(stream)),
              "hive_insert_sequence")
        return {f: getattr(rep, f) for f, _ in SequenceReport._fields_}

    def set_stream_shape(self, tile_ops=0, blocks_per_cu=0, inline_slow_path=0):
        check(self.L.hive_set_stream_shape(self.h, int(tile_ops), int(blocks_per_cu), int(inline_slow_path)),
              "hive_set_stream_shape")

    def set_stream_window(self, window_ops=0, batch_ops=0):
        """Windowed stream form: batches below batch_ops ops are coalesced into
        windows of at most window_ops ops (0 = defaults)."""
        check(self.L.hive_set_stream_window(self.h, int(window_ops), int(batch_ops)), "hive_set_stream_window")

    # --------------------------------------------- reference record ops
    def _ops(self, fn, ops, results, stream):
        torch = _torch()
        if isinstance(ops, np.ndarray):
            ops = torch.from_numpy(ops.view(np.uint8)).cuda()
        n = ops.numel() * ops.element_size() // 16
        if results is None:
            results = torch.empty(n, dtype=torch.int64, device=ops.device)
        check(fn(self.h, _ptr(ops), n, _ptr(results), _stream_ptr(stream)), fn.__name__)
        return results

    def insert_ops(self, ops, results=None, stream=None):
        return self._ops(self.L.hive_insert_ops, ops, results, stream)

    def lookup_ops(self, ops, results=None, stream=None):
        return self._ops(self.L.hive_lookup_ops, ops, results, stream)

    def mixed_ops(self, ops, results=None, stream=None):
        return self._ops(self.L.hive_mixed_ops, ops, results, stream)

    # -------------------------------------------------------- maintenance
    def resize(self, new_num_slots, stream=None):
        check(self.L.hive_resize(self.h, int(new_num_slots), _stream_ptr(stream)), "hive_resize")

    def stash_drain(self, stream=None):
        torch = _torch()
        cap = self.num_buckets * 8
        out = torch.empty(max(cap, 1), dtype=torch.int64, device="cuda")
        n = C.c_size_t(0)
        check(self.L.hive_stash_drain(self.h, _ptr(out), cap, C.byref(n), _stream_ptr(stream)), "hive_stash_drain")
        return out[: n.value]

    def dump(self, cap=None, stream=None):
        """Every live entry as packed kv (int64 device tensor, unordered)."""
        torch = _torch()
        cap = self.num_buckets * (self.slots_per_bucket + 8) if cap is None else int(cap)
        out = torch.empty(max(cap, 1), dtype=torch.int64, device="cuda")
        n = C.c_size_t(0)
        check(self.L.hive_dump(self.h, _ptr(out), cap, C.byref(n), _stream_ptr(stream)), "hive_dump")
        return out[: n.value]

    def stash_reinsert(self, stream=None):
        n = C.c_size_t(0)
        check(self.L.hive_stash_reinsert(self.h, C.byref(n), _stream_ptr(stream)), "hive_stash_reinsert")
        return n.value

    def stats(self, stream=None):
        s = Stats()
        check(self.L.hive_stats(self.h, C.byref(s), _stream_ptr(stream)), "hive_stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def count_exact(self, stream=None):
        a, b = C.c_uint64(0), C.c_uint64(0)
        check(self.L.hive_count_exact(self.h, C.byref(a), C.byref(b), _stream_ptr(stream)), "hive_count_exact")
        return a.value, b.value

    def probe_stats(self, keys, stream=None):
        """(insert candidate buckets, lookup bucket reads, lookup stash reads) for a key set."""
        torch = _torch()
        k = _dev(keys, torch.int32)
        a, b, c = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        check(self.L.hive_probe_stats(self.h, _ptr(k), k.numel(), C.byref(a), C.byref(b), C.byref(c),
                                      _stream_ptr(stream)), "hive_probe_stats")
        return a.value, b.value, c.value

    def export_raw(self, stream=None):
        nb = self.num_buckets
        slots = np.empty(nb * self.slots_per_bucket, np.uint64)
        stash = np.empty(nb * 8, np.uint64)
        cnt = np.empty(nb, np.uint32)
        check(self.L.hive_export_raw(self.h, slots.ctypes.data_as(C.c_void_p), stash.ctypes.data_as(C.c_void_p),
                                     cnt.ctypes.data_as(C.c_void_p), _stream_ptr(stream)), "hive_export_raw")
        return slots, stash, cnt

    def import_raw(self, slots, stash, stash_count, stream=None):
        slots = np.ascontiguousarray(slots, np.uint64)
        stash = np.ascontiguousarray(stash, np.uint64)
        cnt = np.ascontiguousarray(stash_count, np.uint32)
        nb = self.num_buckets
        if slots.size != nb * self.slots_per_bucket or stash.size != nb * 8 or cnt.size != nb:
            raise ValueError("raw arrays do not match the table size")
        check(self.L.hive_import_raw(self.h, slots.ctypes.data_as(C.c_void_p), stash.ctypes.data_as(C.c_void_p),
                                     cnt.ctypes.data_as(C.c_void_p), _stream_ptr(stream)), "hive_import_raw")

    def set_launch(self, op_type, block_threads, keys_per_lane=1, inline_slow_path=0, combine_duplicates=0,
                   split_by_type=1, lanes_per_key=1):
        cfg = Launch(block_threads, keys_per_lane, inline_slow_path, combine_duplicates, split_by_type, lanes_per_key)
        check(self.L.hive_set_launch(self.h, op_type, C.byref(cfg)), "hive_set_launch")

    def get_launch(self, op_type):
        cfg = Launch()
        check(self.L.hive_get_launch(self.h, op_type, C.byref(cfg)), "hive_get_launch")
        return (cfg.block_threads, cfg.keys_per_lane, cfg.inline_slow_path, cfg.combine_duplicates, cfg.split_by_type,
                cfg.lanes_per_key)

    def autotune_bulk(self, insert_keys=None, insert_vals=None, lookup_keys=None, stream=None):
        """Time bulk insert and lookup shapes at the table's current load
        factor (inserts on a scratch copy); keep the fastest."""
        torch = _torch()
        ik = _dev(insert_keys, torch.int32) if insert_keys is not None else None
        iv = _dev(insert_vals, torch.int32) if insert_vals is not None else None
        lk = _dev(lookup_keys, torch.int32) if lookup_keys is not None else None
        res = BulkTune()
        check(self.L.hive_autotune_bulk(self.h, _ptr(ik) if ik is not None else None,
                                        _ptr(iv) if iv is not None else None, ik.numel() if ik is not None else 0,
                                        _ptr(lk) if lk is not None else None, lk.numel() if lk is not None else 0,
                                        C.byref(res), _stream_ptr(stream)), "hive_autotune_bulk")
        out = {f: getattr(res, f) for f, _ in BulkTune._fields_}
        out["lookup_ms_by_lanes"] = {1: res.lookup_ms_by_lanes[0], 4: res.lookup_ms_by_lanes[1],
                                     8: res.lookup_ms_by_lanes[2]}
        out["lookup_ms_by_keys"] = {1: res.lookup_ms_by_keys[0], 2: res.lookup_ms_by_keys[1],
                                    4: res.lookup_ms_by_keys[2]}
        out["insert_ms_by_slow_path"] = {"staged": res.insert_ms_by_slow_path[0],
                                         "inline": res.insert_ms_by_slow_path[1],
                                         "second launch": res.insert_ms_by_slow_path[2]}
        return out

    def autotune_layout(self, insert_keys=None, insert_vals=None, lookup_keys=None, stream=None):
        """Time the insert sample then the lookup sample on a copy of the table
        in each layout and switch the table to the layout with the least total
        time (content kept).  Returns the chosen layout's name and the times
        by layout name (None where the content did not fit)."""
        torch = _torch()
        ik = _dev(insert_keys, torch.int32) if insert_keys is not None else None
        iv = _dev(insert_vals, torch.int32) if insert_vals is not None else None
        lk = _dev(lookup_keys, torch.int32) if lookup_keys is not None else None
        res = LayoutTune()
        check(self.L.hive_autotune_layout(self.h, _ptr(ik) if ik is not None else None,
                                          _ptr(iv) if iv is not None else None, ik.numel() if ik is not None else 0,
                                          _ptr(lk) if lk is not None else None, lk.numel() if lk is not None else 0,
                                          C.byref(res), _stream_ptr(stream)), "hive_autotune_layout")
        names = {v: k for k, v in LAYOUTS.items()}
        return {"layout": names[res.layout], "candidates": res.candidates,
                "insert_ms": {names[i]: (res.insert_ms[i] if res.insert_ms[i] >= 0 else None) for i in range(3)},
                "lookup_ms": {names[i]: (res.lookup_ms[i] if res.lookup_ms[i] >= 0 else None) for i in range(3)}}


def make_ops(types, keys):
    """numpy array of the reference's 16-byte Operation records."""
    n = len(keys)
    a = np.zeros(n, dtype=np.dtype([("type", "<i4"), ("pad", "<u4"), ("key", "<u8")]))
    a["type"] = types
    a["key"] = np.asarray(keys, dtype=np.uint64)
    return a


def dispatch(ops, table_size, threads_per_block=256, num_iterations=1, policy="Default2Hash", layout=0):
    """The reference's hash_table_kernel_dispatch over host buffers.

    Returns (results uint64[n], elapsed_ms per iteration).
    """
    L = lib()
    ops = np.ascontiguousarray(ops)
    n = len(ops)
    res = np.zeros(n, np.uint64)
    ms = C.c_double(0.0)
    nblocks = (n + threads_per_block - 1) // threads_per_block
    check(L.hive_hash_table_kernel_dispatch(ops.ctypes.data_as(C.c_void_p), n, table_size, threads_per_block,
                                            nblocks, num_iterations, C.byref(ms), res.ctypes.data_as(C.c_void_p),
                                            1, layout, policy.encode()), "hive_hash_table_kernel_dispatch")
    return res, ms.value


def dispatch_ycsb(prefill_ops, workload_ops, table_size, threads_per_block=256, num_iterations=1, layout=0):
    """The reference's hash_table_kernel_dispatch_YCSB over host buffers."""
    L = lib()
    pre = np.ascontiguousarray(prefill_ops)
    work = np.ascontiguousarray(workload_ops)
    res = np.zeros(len(work), np.uint64)
    ms = C.c_double(0.0)
    check(L.hive_hash_table_kernel_dispatch_ycsb(pre.ctypes.data_as(C.c_void_p), work.ctypes.data_as(C.c_void_p),
                                                 len(pre), len(work), table_size, threads_per_block, num_iterations,
                                                 C.byref(ms), res.ctypes.data_as(C.c_void_p), 1, layout),
          "hive_hash_table_kernel_dispatch_ycsb")
    return res, ms.value


__all__ = ["HiveTable", "make_ops", "dispatch", "dispatch_ycsb", "Operation", "OP_INSERT", "OP_LOOKUP", "OP_DELETE"]
