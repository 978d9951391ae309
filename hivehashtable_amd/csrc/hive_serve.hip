// hive_serve.hip -- continuous batching of client requests over one resident
// table: admission (in-place resize by free HBM, stash headroom), coalescing
// of arrived requests into mixed batches, completion tracking per request.
//
// Reference analogue: the YCSB driver (src/hashTableBenchmark.cc:434-570,
// src/hive_dispatch.cu:317-549) runs one pre-built workload per call on a
// table created for that call; here the table stays resident and requests
// are served as they arrive.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <vector>

#include "hive_internal.hpp"

using namespace hive;

namespace {

// inserts of each request: the admission test needs an upper bound of new keys
// per request: its inserts, and its inserts plus deletes (writes); the
// request's types are read 16 at a time between its unaligned ends
__global__ void k_req_inserts(const uint8_t* types, const unsigned long long* off, uint32_t* counts,
                              uint32_t* writes) {
  __shared__ uint32_t total, wtotal;
  if (threadIdx.x == 0) total = wtotal = 0;
  __syncthreads();
  const unsigned long long lo = off[blockIdx.x], hi = off[blockIdx.x + 1];
  uint32_t c = 0, w = 0;
  auto count = [&](uint32_t ty) {
    c += ty == OP_INSERT;
    w += ty == OP_INSERT || ty == OP_DELETE;
  };
  const uintptr_t base = reinterpret_cast<uintptr_t>(types);
  unsigned long long a = lo + ((16u - ((base + lo) & 15u)) & 15u);  // first 16-byte aligned op
  if (a > hi) a = hi;
  const unsigned long long z = a + ((hi - a) & ~15ull);
  for (unsigned long long i = lo + threadIdx.x; i < a; i += blockDim.x) count(types[i]);
  for (unsigned long long i = a + 16ull * threadIdx.x; i < z; i += 16ull * blockDim.x) {
    const uint4 v = *reinterpret_cast<const uint4*>(types + i);
    for (uint32_t word : {v.x, v.y, v.z, v.w})
      for (int k = 0; k < 4; ++k) count((word >> (8 * k)) & 0xFFu);
  }
  for (unsigned long long i = z + threadIdx.x; i < hi; i += blockDim.x) count(types[i]);
  if (c) atomicAdd(&total, c);
  if (w) atomicAdd(&wtotal, w);
  __syncthreads();
  if (threadIdx.x == 0) {
    counts[blockIdx.x] = total;
    writes[blockIdx.x] = wtotal;
  }
}

// a rejected request's ops: status FAILED, value 0, table untouched
__global__ void k_reject(uint8_t* status, uint32_t* vals, size_t lo, size_t hi) {
  const size_t i = lo + static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  if (status) status[i] = R_FAILED;
  if (vals) vals[i] = 0;
}

using Clock = std::chrono::steady_clock;

double us_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::micro>(Clock::now() - t0).count();
}

struct Batch {
  size_t first, last;  // requests [first, last)
  hipEvent_t ev;
};

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = p / 100.0 * static_cast<double>(v.size() - 1);
  const size_t lo = static_cast<size_t>(pos);
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - static_cast<double>(lo));
}

struct EventPool {
  std::vector<hipEvent_t> all, free;
  ~EventPool() {
    for (hipEvent_t e : all) (void)hipEventDestroy(e);
  }
};

// Ordered windows alternate between the two window buffers; a window's
// key-set passes run on a side stream as soon as it is launched, overlapping
// the previous window's table work.  A window's table passes wait for its
// key-set passes (ready), and a buffer's next key-set passes wait for its
// last window's chain passes (done).
struct SidePipe {
  hipStream_t ms = nullptr;
  hipEvent_t start = nullptr, ready[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  uint32_t k = 0;
  ~SidePipe() {
    if (ms) (void)hipStreamSynchronize(ms);
    for (hipEvent_t e : {start, ready[0], ready[1], done[0], done[1]})
      if (e) (void)hipEventDestroy(e);
    if (ms) (void)hipStreamDestroy(ms);
  }
};

}  // namespace

extern "C" hive_status_t hive_serve_trace(hive_table_t* t, const uint8_t* types, const uint32_t* keys,
                                          const uint32_t* vals, const uint64_t* h_req_offsets,
                                          const double* h_arrival_us, size_t nreq, uint32_t* out_vals,
                                          uint8_t* status, const hive_serve_config_t* cfg, double* h_latency_us,
                                          hive_serve_report_t* report, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  if (!h_req_offsets || (nreq && !h_arrival_us)) return fail(HIVE_EINVAL, "null request offsets or arrivals");
  if (h_req_offsets[0] != 0) return fail(HIVE_EINVAL, "request offsets[0] must be 0");
  for (size_t i = 0; i < nreq; ++i) {
    if (h_req_offsets[i + 1] < h_req_offsets[i]) return fail(HIVE_EINVAL, "request offsets must be non-decreasing");
    if (!(h_arrival_us[i] >= 0.0) || !std::isfinite(h_arrival_us[i]) || (i && h_arrival_us[i] < h_arrival_us[i - 1]))
      return fail(HIVE_EINVAL, "arrivals must be finite, >= 0 and non-decreasing");
  }
  const uint64_t nops = h_req_offsets[nreq];
  if (nops && (!keys || !types)) return fail(HIVE_EINVAL, "null keys or types");
  hive_serve_config_t c = cfg ? *cfg : hive_serve_config_t{};
  if (c.max_batch_ops == 0) c.max_batch_ops = 1ull << 22;
  if (c.max_inflight == 0) c.max_inflight = 2;
  if (c.max_load_factor == 0.0) c.max_load_factor = 0.9;
  if (c.growth == 0.0) c.growth = 2.0;
  if (c.hbm_reserve_bytes == 0) c.hbm_reserve_bytes = 1ull << 30;
  if (c.stash_headroom == 0.0) c.stash_headroom = 0.5;
  if (c.min_load_factor > 0.0 && c.shrink_load_factor == 0.0) c.shrink_load_factor = 0.75;
  if (c.max_inflight > 8 || !(c.max_load_factor > 0.0 && c.max_load_factor <= 1.0) || !(c.growth > 1.0) ||
      !(c.stash_headroom > 0.0 && c.stash_headroom <= 1.0) || c.ordered > 1)
    return fail(HIVE_EINVAL, "serve config out of range");
  const bool shrinking = c.min_load_factor > 0.0;
  if (!(c.min_load_factor >= 0.0) ||
      (shrinking && !(c.shrink_load_factor > c.min_load_factor && c.shrink_load_factor <= c.max_load_factor)))
    return fail(HIVE_EINVAL, "shrink config out of range: need 0 < min_load_factor < shrink_load_factor <= "
                             "max_load_factor");
  hive_serve_report_t rep{};
  rep.requests = nreq;
  if (nreq == 0) {
    rep.final_slots = static_cast<uint64_t>(t->v.nb) * t->v.S;
    if (report) *report = rep;
    return HIVE_OK;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);

  // ---- set-up, before the trace clock starts
  std::vector<uint32_t> h_ins(nreq), h_wr(nreq);
  {
    DevArena ar;
    unsigned long long* d_off = nullptr;
    uint32_t* d_cnt = nullptr;
    HIP_TRY(ar.alloc(&d_off, nreq + 1));
    HIP_TRY(ar.alloc(&d_cnt, 2 * nreq));
    HIP_TRY(hipMemcpyAsync(d_off, h_req_offsets, (nreq + 1) * 8, hipMemcpyHostToDevice, s));
    for (size_t lo = 0; lo < nreq; lo += 0x7FFFFFFF) {  // grid.x limit
      const size_t cnt = std::min<size_t>(nreq - lo, 0x7FFFFFFF);
      if (nops) {
        k_req_inserts<<<static_cast<unsigned>(cnt), 256, 0, s>>>(types, d_off + lo, d_cnt + lo, d_cnt + nreq + lo);
      } else {
        HIP_TRY(hipMemsetAsync(d_cnt + lo, 0, cnt * 4, s));
        HIP_TRY(hipMemsetAsync(d_cnt + nreq + lo, 0, cnt * 4, s));
      }
    }
    HIP_CHECK_LAUNCH();
    HIP_TRY(hipMemcpyAsync(h_ins.data(), d_cnt, nreq * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_wr.data(), d_cnt + nreq, nreq * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  uint64_t max_req = 0;
  for (size_t i = 0; i < nreq; ++i) max_req = std::max<uint64_t>(max_req, h_req_offsets[i + 1] - h_req_offsets[i]);
  if ((r = ensure_partition(t, std::min<uint64_t>(nops, std::max<uint64_t>(c.max_batch_ops, max_req)))) != HIVE_OK)
    return r;
  // ordered: tiles of 256 ops of one request each (numbered by request) for
  // every request of the trace, and window scratch for the largest batch
  DevArena ord_ar;
  SidePipe pipe;
  std::vector<uint32_t> tile0;  // first tile of request r
  uint4* d_desc = nullptr;
  int ncu = 0;
  if (c.ordered) {
    if (nops > 0xFFFFFFFFull || nreq > 0xFFFFFFFFull) return fail(HIVE_EINVAL, "ordered serving: more than 2^32-1 ops");
    std::vector<uint4> desc;
    tile0.resize(nreq + 1);
    for (size_t q = 0; q < nreq; ++q) {
      tile0[q] = static_cast<uint32_t>(desc.size());
      for (uint64_t i = h_req_offsets[q]; i < h_req_offsets[q + 1]; i += 256) {
        if (desc.size() >= 0x7FFFFFFFu) return fail(HIVE_EINVAL, "ordered serving: too many tiles");
        desc.push_back(make_uint4(static_cast<uint32_t>(i),
                                  static_cast<uint32_t>(std::min<uint64_t>(i + 256, h_req_offsets[q + 1])),
                                  static_cast<uint32_t>(q), 0u));
      }
    }
    tile0[nreq] = static_cast<uint32_t>(desc.size());
    HIP_TRY(ord_ar.alloc(&d_desc, desc.size()));
    if (!desc.empty())
      HIP_TRY(hipMemcpyAsync(d_desc, desc.data(), desc.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
    if ((r = ensure_window(t, std::min<uint64_t>(nops, std::max<uint64_t>(c.max_batch_ops, max_req)))) != HIVE_OK)
      return r;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, t->device));
    HIP_TRY(hipStreamCreateWithFlags(&pipe.ms, hipStreamNonBlocking));
    for (hipEvent_t* e : {&pipe.start, &pipe.ready[0], &pipe.ready[1], &pipe.done[0], &pipe.done[1]})
      HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(pipe.start, s));  // inputs and tile descriptors written on s
    HIP_TRY(hipStreamWaitEvent(pipe.ms, pipe.start, 0));
    HIP_TRY(hipStreamSynchronize(s));
  }
  hive_stats_t st{};
  if ((r = hive_stats(t, &st, stream)) != HIVE_OK) return r;
  uint64_t capacity = st.capacity;
  uint64_t table_ub = st.table_live;  // upper bound of table entries
  uint64_t stash_ub = st.stash_live;  // upper bound of stash entries
  uint64_t del_since = 0;             // deletes launched since the last exact count
  EventPool pool;
  for (uint32_t i = 0; i < c.max_inflight; ++i) {
    hipEvent_t e = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    pool.all.push_back(e);
    pool.free.push_back(e);
  }
  std::vector<double> lat(nreq, 0.0);
  std::vector<uint8_t> rejected(nreq, 0);
  std::deque<Batch> q;
  size_t arrived = 0, launched = 0, completed = 0;
  double last_done = 0.0;
  const Clock::time_point t0 = Clock::now();

  // retire finished batches in order; block = wait for every queued batch
  auto retire = [&](bool block) -> hive_status_t {
    while (!q.empty()) {
      hipError_t e = block ? hipEventSynchronize(q.front().ev) : hipEventQuery(q.front().ev);
      if (e == hipErrorNotReady) return HIVE_OK;
      if (e != hipSuccess) return fail(HIVE_EDEVICE, "batch completion");
      const double done = us_since(t0);
      for (size_t i = q.front().first; i < q.front().last; ++i) lat[i] = done - h_arrival_us[i];
      completed += q.front().last - q.front().first;
      last_done = done;
      pool.free.push_back(q.front().ev);
      q.pop_front();
    }
    return HIVE_OK;
  };
  auto refresh = [&]() -> hive_status_t {  // exact counts; the queue is empty
    hive_stats_t x{};
    hive_status_t rr = hive_stats(t, &x, stream);
    if (rr != HIVE_OK) return rr;
    capacity = x.capacity;
    table_ub = x.table_live;
    stash_ub = x.stash_live;
    del_since = 0;
    return HIVE_OK;
  };
  // Shrink in place when the exact load factor (table entries / capacity, as
  // count_occupancy_kernel counts it, include/hive_kernels.cuh:303-325) is
  // below min_load_factor: to ceil(entries / shrink_load_factor / slots)
  // buckets, entries counting the stash (src/hive_dispatch.cu:189-206 sizes
  // the table for 75 % occupancy).  `drained`: the queue is empty and the
  // counts exact.
  auto maybe_shrink = [&](bool drained) -> hive_status_t {
    const Clock::time_point tr = Clock::now();
    hive_status_t rr;
    if (!drained && ((rr = retire(true)) != HIVE_OK || (rr = refresh()) != HIVE_OK)) return rr;
    if (static_cast<double>(table_ub) >= c.min_load_factor * static_cast<double>(capacity)) return HIVE_OK;
    const uint64_t live = table_ub + stash_ub;
    uint64_t want = static_cast<uint64_t>(std::ceil(static_cast<double>(live) / c.shrink_load_factor / t->v.S));
    if (want < 1) want = 1;
    if (want >= t->v.nb) return HIVE_OK;
    if ((rr = hive_resize(t, want * t->v.S, stream)) != HIVE_OK) {
      if (rr != HIVE_ECAPACITY) return rr;
      g_last_error.clear();  // the entries did not fit the smaller table: it stays as it is
      return HIVE_OK;
    }
    if ((rr = refresh()) != HIVE_OK) return rr;
    rep.shrink_ms += std::chrono::duration<double, std::milli>(Clock::now() - tr).count();
    ++rep.shrinks;
    return HIVE_OK;
  };
  auto fits = [&](uint64_t ins) {
    return static_cast<double>(table_ub + ins) <= c.max_load_factor * static_cast<double>(capacity);
  };

  while (completed < nreq) {
    const double now = us_since(t0);
    while (arrived < nreq && h_arrival_us[arrived] <= now) ++arrived;
    if ((r = retire(false)) != HIVE_OK) return r;
    if (launched < arrived && q.size() < c.max_inflight) {
      // the deletes launched since the last exact count may have taken the
      // load factor below the shrink threshold: count and maybe shrink
      if (shrinking && del_since &&
          static_cast<double>(table_ub - std::min(del_since, table_ub)) <
              c.min_load_factor * static_cast<double>(capacity) &&
          (r = maybe_shrink(false)) != HIVE_OK)
        return r;
      // coalesce every arrived request, up to max_batch_ops (at least one request)
      size_t end = launched + 1;
      uint64_t ins = h_ins[launched];
      const size_t max_req_per_batch = c.ordered ? kMaxWindowBatches : ~size_t(0);
      while (end < arrived && end - launched < max_req_per_batch &&
             h_req_offsets[end + 1] - h_req_offsets[launched] <= c.max_batch_ops)
        ins += h_ins[end++];
      if (ins && !fits(ins)) {
        if ((r = retire(true)) != HIVE_OK || (r = refresh()) != HIVE_OK) return r;
        if (!fits(ins)) {
          // grow in place if free HBM allows it
          double want = static_cast<double>(capacity) * c.growth;
          while (static_cast<double>(table_ub + ins) > c.max_load_factor * want) want *= c.growth;
          const uint64_t new_nb = static_cast<uint64_t>(std::ceil(want / t->v.S));
          size_t free_b = 0, total_b = 0;
          HIP_TRY(hipMemGetInfo(&free_b, &total_b));
          if (new_nb <= 0xFFFFFFFFull &&
              free_b >= view_bytes(static_cast<uint32_t>(new_nb), layout_of(t->v)) + c.hbm_reserve_bytes) {
            const Clock::time_point tr = Clock::now();
            if ((r = hive_resize(t, new_nb * t->v.S, stream)) != HIVE_OK) return r;
            if ((r = refresh()) != HIVE_OK) return r;
            rep.resize_ms += std::chrono::duration<double, std::milli>(Clock::now() - tr).count();
            ++rep.resizes;
          }
        }
        if (!fits(ins)) {
          // no room to grow: keep the longest prefix whose inserts still fit,
          // or that the stash can absorb
          // stash entries the table can hold: 8 per bucket, or, in the
          // in-line layout, 8 per bucket of its stash pool
          const uint64_t stash_cap =
              static_cast<uint64_t>(t->v.layout == kLayoutInline ? std::min(t->v.nb, t->v.pool_cap) : t->v.nb) *
              kStashSlots;
          const double stash_room = c.stash_headroom * static_cast<double>(stash_cap);
          size_t keep = launched;
          uint64_t kins = 0;
          while (keep < end && (fits(kins + h_ins[keep]) ||
                                static_cast<double>(stash_ub + kins + h_ins[keep]) < stash_room)) {
            kins += h_ins[keep++];
          }
          if (keep == launched) {  // the first request cannot be admitted: reject it
            const size_t lo = h_req_offsets[launched], hi = h_req_offsets[launched + 1];
            if (hi > lo) {
              k_reject<<<static_cast<unsigned>((hi - lo + 255) / 256), 256, 0, s>>>(status, out_vals, lo, hi);
              HIP_CHECK_LAUNCH();
            }
            rejected[launched] = 1;
            lat[launched] = us_since(t0) - h_arrival_us[launched];
            ++rep.rejected_requests;
            ++completed;
            ++launched;
            continue;
          }
          if (!fits(kins)) stash_ub += kins;  // these inserts may all land in the stash
          end = keep;
          ins = kins;
        }
      }
      const size_t o = h_req_offsets[launched], n = h_req_offsets[end] - o;
      // the phases the batch has ops for; ordered: each request a batch of
      // one window -- unless the requests write nothing, when every order
      // gives the same results
      uint64_t wr = 0;
      for (size_t q = launched; q < end; ++q) wr += h_wr[q];
      const bool writes = wr != 0;
      const uint32_t phases = (ins ? kPhaseInsert : 0u) | (wr > ins ? kPhaseDelete : 0u);
      if (c.ordered && writes && end - launched > 1 && n) {
        const int p = static_cast<int>(pipe.k++ & 1u);
        if (pipe.used[p]) HIP_TRY(hipStreamWaitEvent(pipe.ms, pipe.done[p], 0));
        if ((r = enqueue_window(t, types, keys, vals, o, n, out_vals, status, d_desc + tile0[launched],
                                tile0[end] - tile0[launched], static_cast<uint32_t>(launched), nullptr, 0, p, ncu,
                                pipe.ms, s, pipe.ready[p], phases)) != HIVE_OK)
          return r;
        HIP_TRY(hipEventRecord(pipe.done[p], s));
        pipe.used[p] = true;
      } else {
        SrcNative src{keys + o, vals ? vals + o : nullptr, types + o};
        OutNative out{status ? status + o : nullptr, out_vals ? out_vals + o : nullptr};
        if ((r = launch_mixed_native(t, src, n, out, s, BatchClock{nullptr, nullptr}, phases)) != HIVE_OK) return r;
      }
      hipEvent_t ev = pool.free.back();
      pool.free.pop_back();
      HIP_TRY(hipEventRecord(ev, s));
      q.push_back(Batch{launched, end, ev});
      table_ub += ins;
      del_since += wr - ins;
      ++rep.batches;
      rep.ops += n;
      launched = end;
    }
  }
  HIP_TRY(hipStreamSynchronize(s));
  if ((r = refresh()) != HIVE_OK) return r;
  if (shrinking && (r = maybe_shrink(true)) != HIVE_OK) return r;
  std::vector<double> served;
  served.reserve(nreq);
  for (size_t i = 0; i < nreq; ++i)
    if (!rejected[i]) served.push_back(lat[i]);
  rep.final_slots = capacity;
  rep.span_ms = (last_done - h_arrival_us[0]) * 1e-3;
  rep.mops = rep.span_ms > 0.0 ? static_cast<double>(rep.ops) / (rep.span_ms * 1e3) : 0.0;
  rep.p50_us = percentile(served, 50.0);
  rep.p99_us = percentile(served, 99.0);
  rep.max_us = served.empty() ? 0.0 : *std::max_element(served.begin(), served.end());
  rep.mean_batch_ops = rep.batches ? static_cast<double>(rep.ops) / static_cast<double>(rep.batches) : 0.0;
  if (h_latency_us) std::copy(lat.begin(), lat.end(), h_latency_us);
  if (report) *report = rep;
  return HIVE_OK;
}
