// hive_sequence.hip -- a stream of insert batches (each optionally followed by
// lookups) with in-place growth: run as plain per-batch calls that decide each
// resize on the host, or as one hipGraph that holds the whole sequence,
// resizes included.
//
// Reference analogue: the DYNAMIC_RESIZE path of run_hash_table_benchmark
// (src/hive_dispatch.cu:181-211), which checks the load factor on the host
// after each batch and calls hive_hash_grow.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "hive_internal.hpp"

using namespace hive;

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

struct Step {
  size_t lo, n;    // insert keys [lo, lo + n)
  int grow = -1;   // index of the target table to rehash into before the batch
};

}  // namespace

extern "C" hive_status_t hive_insert_sequence(hive_table_t* t, const uint32_t* keys, const uint32_t* vals, size_t n,
                                              size_t batch_keys, const uint32_t* lookup_keys,
                                              size_t lookups_per_batch, double max_load_factor, double growth,
                                              int mode, uint8_t* status, uint8_t* lookup_status,
                                              hive_sequence_report_t* report, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  if (batch_keys == 0 || (n && !keys) || (lookups_per_batch && !lookup_keys))
    return fail(HIVE_EINVAL, "bad batch size or null keys");
  if (!(max_load_factor > 0.0 && max_load_factor <= 1.0) || !(growth > 1.0))
    return fail(HIVE_EINVAL, "max_load_factor must be in (0, 1] and growth > 1");
  if (mode != HIVE_SEQUENCE_PLAIN && mode != HIVE_SEQUENCE_GRAPH) return fail(HIVE_EINVAL, "unknown sequence mode");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hive_sequence_report_t rep{};
  const size_t nbatch = (n + batch_keys - 1) / batch_keys;

  auto grow_target = [&](uint64_t live, uint64_t add, uint64_t cap) {
    double want = static_cast<double>(cap) * growth;
    while (static_cast<double>(live + add) > max_load_factor * want) want *= growth;
    return static_cast<uint64_t>(std::ceil(want / t->v.S));  // buckets
  };
  auto launch_batch = [&](size_t b, size_t lo, size_t m, hipStream_t q) -> hive_status_t {
    hive_status_t rr = launch_insert(t, SrcNative{keys + lo, vals ? vals + lo : nullptr, nullptr}, m,
                                     OutNative{status ? status + lo : nullptr, nullptr}, -1, q);
    if (rr != HIVE_OK || !lookups_per_batch) return rr;
    const size_t lo2 = b * lookups_per_batch;
    return launch_lookup(t, SrcNative{lookup_keys + lo2, nullptr, nullptr}, lookups_per_batch,
                         OutNative{lookup_status ? lookup_status + lo2 : nullptr, nullptr}, -1, q);
  };

  if (mode == HIVE_SEQUENCE_PLAIN) {
    // per batch: read the exact count on the host, grow with hive_resize, launch
    HIP_TRY(hipStreamSynchronize(s));
    const Clock::time_point t0 = Clock::now();
    for (size_t b = 0; b < nbatch; ++b) {
      const size_t lo = b * batch_keys, m = std::min(batch_keys, n - lo);
      hive_stats_t st{};
      if ((r = hive_stats(t, &st, stream)) != HIVE_OK) return r;
      if (static_cast<double>(st.table_live + st.stash_live + m) > max_load_factor * static_cast<double>(st.capacity)) {
        const uint64_t nb = grow_target(st.table_live + st.stash_live, m, st.capacity);
        if (nb > 0xFFFFFFFFull) return fail(HIVE_EINVAL, "growth beyond 2^32-1 buckets");
        if ((r = hive_resize(t, nb * t->v.S, stream)) != HIVE_OK) return r;
        ++rep.resizes;
      }
      if ((r = launch_batch(b, lo, m, s)) != HIVE_OK) return r;
    }
    HIP_TRY(hipStreamSynchronize(s));
    rep.elapsed_ms = ms_since(t0);
  } else {
    // the schedule is fixed up front from an upper bound of the live count
    // (every insert adds an entry); target tables are allocated and cleared
    // before capture, so the graph holds only kernels
    const Clock::time_point t0 = Clock::now();
    hive_stats_t st{};
    if ((r = hive_stats(t, &st, stream)) != HIVE_OK) return r;
    uint64_t live = st.table_live + st.stash_live, cap = st.capacity;
    std::vector<Step> steps(nbatch);
    std::vector<uint32_t> target_nb;
    for (size_t b = 0; b < nbatch; ++b) {
      steps[b].lo = b * batch_keys;
      steps[b].n = std::min(batch_keys, n - steps[b].lo);
      if (static_cast<double>(live + steps[b].n) > max_load_factor * static_cast<double>(cap)) {
        const uint64_t nb = grow_target(live, steps[b].n, cap);
        if (nb > 0xFFFFFFFFull) return fail(HIVE_EINVAL, "growth beyond 2^32-1 buckets");
        steps[b].grow = static_cast<int>(target_nb.size());
        target_nb.push_back(static_cast<uint32_t>(nb));
        cap = nb * t->v.S;
      }
      live += steps[b].n;
    }
    std::vector<TableView> views(target_nb.size());
    struct ViewsGuard {
      std::vector<TableView>* v;
      bool keep_last = false;
      ~ViewsGuard() {
        for (size_t i = 0; i < v->size(); ++i)
          if (!(keep_last && i + 1 == v->size())) free_view((*v)[i]);
      }
    } guard{&views};
    for (size_t i = 0; i < views.size(); ++i) {
      if ((r = alloc_view(views[i], target_nb[i], layout_of(t->v))) != HIVE_OK) return r;
      if ((r = clear_view(views[i], s)) != HIVE_OK) return r;
    }
    DevArena ar;
    unsigned long long* d_fail = nullptr;
    HIP_TRY(ar.alloc(&d_fail, 1));
    HIP_TRY(hipMemsetAsync(d_fail, 0, 8, s));
    HIP_TRY(hipStreamSynchronize(s));
    const TableView orig = t->v;
    hipStream_t cs = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hive_status_t rc = HIVE_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      rc = fail(HIVE_EDEVICE, "graph capture setup");
    } else {
      for (size_t b = 0; b < nbatch && rc == HIVE_OK; ++b) {
        if (steps[b].grow >= 0) {
          const TableView nv = views[steps[b].grow];
          HIVE_POLICY_SWITCH(table_kind(t), (rc = launch_rehash<P>(nv, t->v, d_fail, cs)));
          t->v = nv;
        }
        rc = launch_batch(b, steps[b].lo, steps[b].n, cs);
      }
      hipError_t e = hipStreamEndCapture(cs, &g);
      if (rc == HIVE_OK && e != hipSuccess) rc = fail(HIVE_EDEVICE, "graph capture");
      if (rc == HIVE_OK && hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess)
        rc = fail(HIVE_EDEVICE, "graph instantiate");
    }
    rep.setup_ms = ms_since(t0);
    if (rc == HIVE_OK) {
      const Clock::time_point t1 = Clock::now();
      if (hipGraphLaunch(ge, cs) != hipSuccess || hipStreamSynchronize(cs) != hipSuccess)
        rc = fail(HIVE_EDEVICE, "graph launch");
      rep.elapsed_ms = ms_since(t1);
    }
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    if (rc != HIVE_OK) {
      t->v = orig;  // nothing ran (or the launch failed): keep the original table
      return rc;
    }
    unsigned long long failed = 0;
    HIP_TRY(hipMemcpy(&failed, d_fail, 8, hipMemcpyDeviceToHost));
    if (!views.empty()) {
      TableView old = orig;
      free_view(old);
      guard.keep_last = true;  // the handle now owns the last target table
      t->bytes = view_bytes(t->v.nb, layout_of(t->v)) + 16;
    }
    rep.resizes = static_cast<uint32_t>(views.size());
    rep.rehash_failed = failed;
    if (failed) return fail(HIVE_ECAPACITY, "a rehash inside the sequence could not place every entry");
  }
  rep.final_slots = static_cast<uint64_t>(t->v.nb) * t->v.S;
  if (report) *report = rep;
  return HIVE_OK;
}
