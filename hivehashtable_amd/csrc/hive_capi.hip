// hive_capi.hip -- host runtime of libhive.so: table lifecycle, launches of the
// bulk / mixed / maintenance kernels, and the drop-in dispatch pair with the
// reference's host semantics (src/hive_dispatch.cu:29-315,317-549,551-872).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "hive_internal.hpp"

using namespace hive;

namespace hive {
thread_local std::string g_last_error;
}

namespace {

hive_status_t read_counters(hive_table* t, hipStream_t s, uint64_t* tab, uint64_t* st, uint64_t* slow = nullptr) {
  std::vector<unsigned long long> h(counter_words());
  HIP_TRY(hipMemcpyAsync(h.data(), t->v.counters, counter_words() * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  unsigned long long a = 0, b = 0, c = 0;
  for (int l = 0; l < kCounterLanes; ++l) {
    a += h[(0 * kCounterLanes + l) * kCounterStride];
    b += h[(1 * kCounterLanes + l) * kCounterStride];
    c += h[(2 * kCounterLanes + l) * kCounterStride];
  }
  *tab = a;
  *st = b;
  if (slow) *slow = c;
  return HIVE_OK;
}

hive_status_t count_exact_view(hive_table* t, const TableView& v, hipStream_t s, uint64_t* tab, uint64_t* st) {
  HIP_TRY(hipMemsetAsync(t->scratch, 0, 16, s));
  const size_t ns = static_cast<size_t>(v.nb) * v.S;
  HIVE_POLICY_SWITCH(table_kind(t->policy, layout_of(v)), (k_count_exact<P><<<fill_grid(ns), 256, 0, s>>>(v, t->scratch)));
  HIP_CHECK_LAUNCH();
  unsigned long long h[2];
  HIP_TRY(hipMemcpyAsync(h, t->scratch, 16, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *tab = h[0];
  *st = h[1];
  return HIVE_OK;
}

// set the spread counters to exact totals (lane 0 holds everything)
hive_status_t reset_counters(hive_table* t, const TableView& v, hipStream_t s, uint64_t tab, uint64_t st) {
  std::vector<unsigned long long> h(counter_words(), 0ull);
  h[0] = tab;
  h[static_cast<size_t>(kCounterLanes) * kCounterStride] = st;
  HIP_TRY(hipMemcpyAsync(v.counters, h.data(), counter_words() * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  (void)t;
  return HIVE_OK;
}

// ---- the no-allocation mixed batch (hive_mixed_noalloc)
// The type lists of one mixed batch of up to kNoAllocMaxOps ops live in a
// file-scope device array, one copy per GPU, held until the process ends.
// In host code `scratch` names a host-side stand-in, never the device
// memory: its device address is read once per device with
// hipGetSymbolAddress (noalloc_setup, from hive_create) and only that
// address is passed to kernels.
constexpr size_t kNoAllocMaxOps = size_t(1) << 20;
constexpr size_t kNoAllocWords = (partition_words(kNoAllocMaxOps) + 63) / 64 * 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
uint32_t* g_noalloc_ptr[kMaxDevices] = {};  // scratch's device address on each device (once set up)
bool g_noalloc_done[kMaxDevices] = {};
hipError_t g_noalloc_err[kMaxDevices] = {};

// with `dev` current: look up scratch's address on it, once
hive_status_t noalloc_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(HIVE_EDEVICE, "device ordinal above 63");
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (!g_noalloc_done[dev]) {
    void* p = nullptr;
    g_noalloc_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_noalloc_err[dev] == hipSuccess) g_noalloc_ptr[dev] = static_cast<uint32_t*>(p);
    else (void)hipGetLastError();
    g_noalloc_done[dev] = true;
  }
  if (g_noalloc_err[dev] != hipSuccess)
    return fail(HIVE_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_noalloc_err[dev]));
  return HIVE_OK;
}

// Rehash every live entry (buckets, then stash) into a new table of
// num_slots / width buckets of `width` slots, with the insert protocol; swap
// it in, or keep the old table on HIVE_ECAPACITY.
hive_status_t relayout(hive_table* t, uint64_t num_slots, int layout, hipStream_t s) {
  hive_status_t r;
  if ((r = set_device(t)) != HIVE_OK) return r;
  const uint64_t nb = num_slots / layout_slots(layout);
  if (nb == 0 || nb > 0xFFFFFFFFull) return fail(HIVE_EINVAL, "new size must give 1 .. 2^32-1 buckets");
  TableView nv;
  if ((r = alloc_view(nv, static_cast<uint32_t>(nb), layout)) != HIVE_OK) return r;
  if ((r = clear_view(nv, s)) != HIVE_OK) {
    free_view(nv);
    return r;
  }
  HIP_TRY(hipMemsetAsync(t->scratch, 0, 16, s));
  hive_status_t rr = HIVE_OK;
  switch (table_kind(t->policy, layout)) {
#define HIVE_RELAYOUT_CASE(K, T) \
    case K: rr = launch_rehash<T>(nv, t->v, t->scratch, s); break;
    HIVE_RELAYOUT_CASE(0, PolicyDefault2) HIVE_RELAYOUT_CASE(1, PolicyMurmurCity) HIVE_RELAYOUT_CASE(2, PolicyTriple)
    HIVE_RELAYOUT_CASE(3, PolicyMurmurCityBit) HIVE_RELAYOUT_CASE(4, PolicyLookup2) HIVE_RELAYOUT_CASE(5, PolicyLookup3)
    HIVE_RELAYOUT_CASE(8, Wide16<PolicyDefault2>) HIVE_RELAYOUT_CASE(9, Wide16<PolicyMurmurCity>)
    HIVE_RELAYOUT_CASE(10, Wide16<PolicyTriple>) HIVE_RELAYOUT_CASE(11, Wide16<PolicyMurmurCityBit>)
    HIVE_RELAYOUT_CASE(12, Wide16<PolicyLookup2>) HIVE_RELAYOUT_CASE(13, Wide16<PolicyLookup3>)
    HIVE_RELAYOUT_CASE(16, Inline8<PolicyDefault2>) HIVE_RELAYOUT_CASE(17, Inline8<PolicyMurmurCity>)
    HIVE_RELAYOUT_CASE(18, Inline8<PolicyTriple>) HIVE_RELAYOUT_CASE(19, Inline8<PolicyMurmurCityBit>)
    HIVE_RELAYOUT_CASE(20, Inline8<PolicyLookup2>) HIVE_RELAYOUT_CASE(21, Inline8<PolicyLookup3>)
#undef HIVE_RELAYOUT_CASE
    default: rr = fail(HIVE_EINVAL, "unknown policy or layout");
  }
  if (rr != HIVE_OK) {
    free_view(nv);
    return rr;
  }
  unsigned long long failed = 0;
  if (hipMemcpyAsync(&failed, t->scratch, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    free_view(nv);
    return fail(HIVE_EDEVICE, "rehash sync");
  }
  if (failed) {
    free_view(nv);
    return fail(HIVE_ECAPACITY, "resize target cannot hold every live entry");
  }
  free_view(t->v);
  t->v = nv;
  t->bytes = view_bytes(static_cast<uint32_t>(nb), layout) + 16;
  return HIVE_OK;
}

}  // namespace

extern "C" {

const char* hive_status_string(hive_status_t s) {
  switch (s) {
    case HIVE_OK: return "ok";
    case HIVE_EINVAL: return "invalid argument";
    case HIVE_ENOMEM: return "out of memory";
    case HIVE_EDEVICE: return "device error";
    case HIVE_ECAPACITY: return "capacity exceeded";
    case HIVE_ERANGE: return "output buffer too small";
  }
  return "unknown status";
}

const char* hive_last_error(void) { return g_last_error.c_str(); }

hive_status_t hive_policy_from_name(const char* name, int* policy) {
  static const char* names[] = {"Default2Hash", "MurmurCityHash", "TripleHash",
                                "MurmurCityBitHash", "Lookup2Hash", "Lookup3Hash"};
  if (!name || !policy) return fail(HIVE_EINVAL, "null argument");
  for (int i = 0; i < 6; ++i)
    if (std::strcmp(name, names[i]) == 0) {
      *policy = i;
      return HIVE_OK;
    }
  return fail(HIVE_EINVAL, std::string("unknown hash policy: ") + name);
}

hive_status_t hive_create(uint64_t num_slots, int policy, hive_table_t** out) {
  return hive_create_ex(num_slots, policy, HIVE_LAYOUT_BUCKET8, out);
}

hive_status_t hive_create_ex(uint64_t num_slots, int policy, int layout, hive_table_t** out) {
  if (!out) return fail(HIVE_EINVAL, "null out");
  *out = nullptr;
  if (policy < 0 || policy > HIVE_POLICY_LOOKUP3) return fail(HIVE_EINVAL, "unknown policy");
  if (!valid_layout(layout)) return fail(HIVE_EINVAL, "unknown layout");
  const uint64_t nb = num_slots / layout_slots(layout);
  if (nb == 0 || nb > 0xFFFFFFFFull) return fail(HIVE_EINVAL, "num_slots must give 1 .. 2^32-1 buckets");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  auto* t = new (std::nothrow) hive_table();
  if (!t) return fail(HIVE_ENOMEM, "host allocation");
  if (hive_status_t rs = noalloc_setup(dev); rs != HIVE_OK) {
    delete t;
    return rs;
  }
  t->device = dev;
  t->policy = policy;
  hive_status_t r = alloc_view(t->v, static_cast<uint32_t>(nb), layout);
  if (r != HIVE_OK) {
    delete t;
    return r;
  }
  if (hipMalloc(&t->scratch, 64) != hipSuccess) {
    free_view(t->v);
    delete t;
    return fail(HIVE_ENOMEM, "hipMalloc scratch");
  }
  t->bytes = view_bytes(static_cast<uint32_t>(nb), layout) + 16;
  if ((r = clear_view(t->v, nullptr)) != HIVE_OK || hipStreamSynchronize(nullptr) != hipSuccess) {
    free_view(t->v);
    (void)hipFree(t->scratch);
    delete t;
    return r != HIVE_OK ? r : fail(HIVE_EDEVICE, "clear at create");
  }
  *out = t;
  return HIVE_OK;
}

hive_status_t hive_destroy(hive_table_t* t) {
  if (!t) return HIVE_OK;
  (void)hipSetDevice(t->device);
  (void)hipDeviceSynchronize();
  free_view(t->v);
  if (t->scratch) (void)hipFree(t->scratch);
  if (t->part) (void)hipFree(t->part);
  if (t->pend) (void)hipFree(t->pend);
  (void)free_window(t);
  delete t;
  return HIVE_OK;
}

hive_status_t hive_clear(hive_table_t* t, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  return clear_view(t->v, static_cast<hipStream_t>(stream));
}

uint64_t hive_num_buckets(const hive_table_t* t) { return t ? t->v.nb : 0; }

hive_status_t hive_insert(hive_table_t* t, const uint32_t* keys, const uint32_t* vals, size_t n, uint8_t* status,
                          void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (n && !keys) return fail(HIVE_EINVAL, "null keys");
  if ((r = ensure_pending(t, n, static_cast<hipStream_t>(stream))) != HIVE_OK) return r;
  return launch_insert(t, SrcNative{keys, vals, nullptr}, n, OutNative{status, nullptr}, -1,
                       static_cast<hipStream_t>(stream));
}

hive_status_t hive_lookup(hive_table_t* t, const uint32_t* keys, size_t n, uint32_t* vals_out, uint8_t* status,
                          void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (n && !keys) return fail(HIVE_EINVAL, "null keys");
  return launch_lookup(t, SrcNative{keys, nullptr, nullptr}, n, OutNative{status, vals_out}, -1,
                       static_cast<hipStream_t>(stream));
}

hive_status_t hive_erase(hive_table_t* t, const uint32_t* keys, size_t n, uint8_t* status, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (n && !keys) return fail(HIVE_EINVAL, "null keys");
  return launch_erase(t, SrcNative{keys, nullptr, nullptr}, n, OutNative{status, nullptr}, -1,
                      static_cast<hipStream_t>(stream));
}

hive_status_t hive_mixed(hive_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals, size_t n,
                         uint32_t* out_vals, uint8_t* status, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (n && (!keys || !types)) return fail(HIVE_EINVAL, "null keys or types");
  if ((r = set_device(t)) != HIVE_OK || (r = ensure_partition(t, n)) != HIVE_OK) return r;
  return launch_mixed_native(t, SrcNative{keys, vals, types}, n, OutNative{status, out_vals},
                             static_cast<hipStream_t>(stream));
}

hive_status_t hive_mixed_noalloc(hive_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                                 size_t n, uint32_t* out_vals, uint8_t* status, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (n && (!keys || !types)) return fail(HIVE_EINVAL, "null keys or types");
  if (n > kNoAllocMaxOps) return fail(HIVE_EINVAL, "hive_mixed_noalloc: more than 2^20 ops in one batch");
  if ((r = set_device(t)) != HIVE_OK || (r = noalloc_setup(t->device)) != HIVE_OK) return r;
  const SrcNative src{keys, vals, types};
  const OutNative out{status, out_vals};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!use_partition(t, n, types)) return launch_mixed(t, src, n, out, s);
  uint32_t* part = g_noalloc_ptr[t->device];
  if ((r = partition_batch(src, n, part, s)) != HIVE_OK) return r;  // clears the list counts first
  return launch_mixed_listed(t, src, n, out, part, s);
}

hive_status_t hive_insert_ops(hive_table_t* t, const hive_operation_t* ops, size_t n, uint64_t* results,
                              void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (n && !ops) return fail(HIVE_EINVAL, "null ops");
  return launch_insert(t, SrcOps{reinterpret_cast<const Op16*>(ops)}, n,
                       OutU64{reinterpret_cast<unsigned long long*>(results)}, -1, static_cast<hipStream_t>(stream));
}

hive_status_t hive_lookup_ops(hive_table_t* t, const hive_operation_t* ops, size_t n, uint64_t* results,
                              void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (n && !ops) return fail(HIVE_EINVAL, "null ops");
  return launch_lookup(t, SrcOps{reinterpret_cast<const Op16*>(ops)}, n,
                       OutU64{reinterpret_cast<unsigned long long*>(results)}, -1, static_cast<hipStream_t>(stream));
}

hive_status_t hive_mixed_ops(hive_table_t* t, const hive_operation_t* ops, size_t n, uint64_t* results,
                             void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (n && !ops) return fail(HIVE_EINVAL, "null ops");
  return launch_mixed(t, SrcOps{reinterpret_cast<const Op16*>(ops)}, n,
                      OutU64{reinterpret_cast<unsigned long long*>(results)}, static_cast<hipStream_t>(stream));
}

hive_status_t hive_resize(hive_table_t* t, uint64_t new_num_slots, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  return relayout(t, new_num_slots, layout_of(t->v), static_cast<hipStream_t>(stream));
}

hive_status_t hive_set_layout(hive_table_t* t, int layout, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (!valid_layout(layout)) return fail(HIVE_EINVAL, "unknown layout");
  if (layout == layout_of(t->v)) return HIVE_OK;
  return relayout(t, static_cast<uint64_t>(t->v.nb) * t->v.S, layout, static_cast<hipStream_t>(stream));
}

hive_status_t hive_stash_drain(hive_table_t* t, uint64_t* out_kv, size_t cap, size_t* n_out, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipMemsetAsync(t->scratch, 0, 16, s));
  HIVE_POLICY_SWITCH(table_kind(t), (k_stash_count_live<P><<<grid_for(t->v.nb, 256), 256, 0, s>>>(t->v, t->scratch)));
  HIP_CHECK_LAUNCH();
  unsigned long long total = 0;
  HIP_TRY(hipMemcpyAsync(&total, t->scratch, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (n_out) *n_out = static_cast<size_t>(total);
  if (total > cap) return fail(HIVE_ERANGE, "stash drain: output buffer too small");
  if (total) {
    if (!out_kv) return fail(HIVE_EINVAL, "null out_kv");
    HIP_TRY(hipMemsetAsync(t->scratch, 0, 8, s));
    auto* o = reinterpret_cast<unsigned long long*>(out_kv);
    HIVE_POLICY_SWITCH(table_kind(t), (k_stash_compact<P><<<grid_for(t->v.nb, 256), 256, 0, s>>>(t->v, o, t->scratch)));
    HIP_CHECK_LAUNCH();
  }
  {
    const size_t nr = t->v.layout == kLayoutInline ? t->v.nb : meta_words(t->v.nb);
    HIVE_POLICY_SWITCH(table_kind(t), (k_reset_stash<P><<<grid_for(nr, 256), 256, 0, s>>>(t->v)));
  }
  HIP_CHECK_LAUNCH();
  uint64_t tab = 0, st = 0;
  if ((r = read_counters(t, s, &tab, &st)) != HIVE_OK) return r;
  return reset_counters(t, t->v, s, tab, 0);
}

hive_status_t hive_dump(hive_table_t* t, uint64_t* out_kv, size_t cap, size_t* n_out, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint64_t tab = 0, st = 0;
  if ((r = count_exact_view(t, t->v, s, &tab, &st)) != HIVE_OK) return r;
  const uint64_t total = tab + st;
  if (n_out) *n_out = static_cast<size_t>(total);
  if (total > cap) return fail(HIVE_ERANGE, "dump: output buffer too small");
  if (total == 0) return HIVE_OK;
  if (!out_kv) return fail(HIVE_EINVAL, "null out_kv");
  auto* out = reinterpret_cast<unsigned long long*>(out_kv);
  HIP_TRY(hipMemsetAsync(t->scratch, 0, 8, s));
  HIVE_POLICY_SWITCH(table_kind(t),
                     (k_dump_slots<P><<<grid_for(static_cast<size_t>(t->v.nb) * t->v.S, 256), 256, 0, s>>>(t->v, out, t->scratch),
                      k_stash_compact<P><<<grid_for(t->v.nb, 256), 256, 0, s>>>(t->v, out, t->scratch)));
  HIP_CHECK_LAUNCH();
  unsigned long long written = 0;
  HIP_TRY(hipMemcpyAsync(&written, t->scratch, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (written != total) return fail(HIVE_EDEVICE, "dump: the table changed during the export");
  return HIVE_OK;
}

hive_status_t hive_stash_reinsert(hive_table_t* t, size_t* n_reinserted, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t cap = static_cast<size_t>(t->v.nb) * kStashSlots;
  unsigned long long* buf = nullptr;
  HIP_TRY(hipMalloc(&buf, cap * 8));
  size_t n = 0;
  r = hive_stash_drain(t, reinterpret_cast<uint64_t*>(buf), cap, &n, stream);
  if (r == HIVE_OK && n) {
    unsigned long long* failed = t->scratch + 1;
    if (hipMemsetAsync(failed, 0, 8, s) != hipSuccess) r = fail(HIVE_EDEVICE, "memset");
    if (r == HIVE_OK) r = launch_insert(t, SrcKV{buf}, n, OutCount{failed}, -1, s);
    unsigned long long nf = 0;
    if (r == HIVE_OK && (hipMemcpyAsync(&nf, failed, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
                         hipStreamSynchronize(s) != hipSuccess))
      r = fail(HIVE_EDEVICE, "reinsert sync");
    if (r == HIVE_OK && nf) r = fail(HIVE_ECAPACITY, "stash re-insert: entries could not be placed");
  }
  (void)hipFree(buf);
  if (n_reinserted) *n_reinserted = n;
  return r;
}

hive_status_t hive_stats(hive_table_t* t, hive_stats_t* out, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  if (!out) return fail(HIVE_EINVAL, "null out");
  uint64_t tab = 0, st = 0, slow = 0;
  if ((r = read_counters(t, static_cast<hipStream_t>(stream), &tab, &st, &slow)) != HIVE_OK) return r;
  out->slow_path_ops = slow;
  out->num_buckets = t->v.nb;
  out->slots_per_bucket = t->v.S;
  out->capacity = static_cast<uint64_t>(t->v.nb) * t->v.S;
  out->table_live = tab;
  out->stash_live = st;
  out->load_factor = static_cast<double>(tab) / static_cast<double>(out->capacity);
  out->device_bytes = t->bytes;
  out->policy = t->policy;
  out->layout = layout_of(t->v);
  return HIVE_OK;
}

hive_status_t hive_count_exact(hive_table_t* t, uint64_t* table_live, uint64_t* stash_live, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  uint64_t a = 0, b = 0;
  if ((r = count_exact_view(t, t->v, static_cast<hipStream_t>(stream), &a, &b)) != HIVE_OK) return r;
  if (table_live) *table_live = a;
  if (stash_live) *stash_live = b;
  return HIVE_OK;
}

hive_status_t hive_probe_stats(hive_table_t* t, const uint32_t* keys, size_t n, uint64_t* candidate_buckets,
                               uint64_t* lookup_bucket_reads, uint64_t* lookup_stash_reads, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  if (n && !keys) return fail(HIVE_EINVAL, "null keys");
  hipStream_t s = static_cast<hipStream_t>(stream);
  constexpr size_t kWords = 64 * 8;
  unsigned long long* acc = nullptr;
  HIP_TRY(hipMalloc(&acc, kWords * 8));
  std::vector<unsigned long long> h(kWords, 0ull);
  hipError_t e = hipMemsetAsync(acc, 0, kWords * 8, s);
  if (e == hipSuccess && n) {
    HIVE_POLICY_SWITCH(table_kind(t), (k_probe_stats<P><<<grid_for(n, 256), 256, 0, s>>>(t->v, keys, n, acc)));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), acc, kWords * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(acc);
  if (e != hipSuccess) return fail(HIVE_EDEVICE, std::string("probe stats: ") + hipGetErrorString(e));
  unsigned long long tot[4] = {0, 0, 0, 0};
  for (int l = 0; l < 64; ++l)
    for (int f = 0; f < 4; ++f) tot[f] += h[l * 8 + f];
  if (candidate_buckets) *candidate_buckets = tot[0];
  if (lookup_bucket_reads) *lookup_bucket_reads = tot[1];
  if (lookup_stash_reads) *lookup_stash_reads = tot[2];
  return HIVE_OK;
}

hive_status_t hive_export_raw(hive_table_t* t, uint64_t* h_slots, uint64_t* h_stash, uint32_t* h_stash_count,
                              void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t nb = t->v.nb;
  if (t->v.layout == kLayoutInline) {  // converted on the device into the separate layout's arrays
    DevArena ar;
    unsigned long long *ds = nullptr, *dst = nullptr;
    uint32_t* dc = nullptr;
    HIP_TRY(ar.alloc(&ds, nb * kRefSlots));
    HIP_TRY(ar.alloc(&dst, nb * kStashSlots));
    HIP_TRY(ar.alloc(&dc, nb));
    HIVE_POLICY_SWITCH(table_kind(t), (k_inline_to_raw<P><<<grid_for(nb, 256), 256, 0, s>>>(t->v, ds, dst, dc)));
    HIP_CHECK_LAUNCH();
    if (h_slots) HIP_TRY(hipMemcpyAsync(h_slots, ds, nb * kRefSlots * 8, hipMemcpyDeviceToHost, s));
    if (h_stash) HIP_TRY(hipMemcpyAsync(h_stash, dst, nb * kStashSlots * 8, hipMemcpyDeviceToHost, s));
    if (h_stash_count) HIP_TRY(hipMemcpyAsync(h_stash_count, dc, nb * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return HIVE_OK;
  }
  if (h_slots) HIP_TRY(hipMemcpyAsync(h_slots, t->v.slots, nb * t->v.S * 8, hipMemcpyDeviceToHost, s));
  if (h_stash) HIP_TRY(hipMemcpyAsync(h_stash, t->v.stash, nb * kStashSlots * 8, hipMemcpyDeviceToHost, s));
  if (h_stash_count) HIP_TRY(hipMemcpyAsync(h_stash_count, t->v.stash_count, nb * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return HIVE_OK;
}

hive_status_t hive_import_raw(hive_table_t* t, const uint64_t* h_slots, const uint64_t* h_stash,
                              const uint32_t* h_stash_count, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  if (!h_slots || !h_stash || !h_stash_count) return fail(HIVE_EINVAL, "null raw array");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t nb = t->v.nb;
  for (size_t b = 0; b < nb; ++b)
    if (h_stash_count[b] > static_cast<uint32_t>(kStashSlots)) return fail(HIVE_EINVAL, "stash count above 8");
  if (t->v.layout == kLayoutInline) {  // converted on the device from the separate layout's arrays
    DevArena ar;
    unsigned long long *ds = nullptr, *dst = nullptr;
    uint32_t* dc = nullptr;
    HIP_TRY(ar.alloc(&ds, nb * kRefSlots));
    HIP_TRY(ar.alloc(&dst, nb * kStashSlots));
    HIP_TRY(ar.alloc(&dc, nb));
    HIP_TRY(hipMemcpyAsync(ds, h_slots, nb * kRefSlots * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dst, h_stash, nb * kStashSlots * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dc, h_stash_count, nb * 4, hipMemcpyHostToDevice, s));
    k_fill_u64<<<1, 64, 0, s>>>(t->scratch, 2, 0ull);
    k_fill_u32<<<kPoolCursors * kPoolCursorStride / 64, 64, 0, s>>>(t->v.pool_next, kPoolCursors * kPoolCursorStride,
                                                                     0u);
    HIVE_POLICY_SWITCH(table_kind(t), (k_raw_to_inline<P><<<grid_for(nb, 256), 256, 0, s>>>(t->v, ds, dst, dc, t->scratch),
                                       k_rebuild_meta<P><<<grid_for(nb * t->v.S, 256), 256, 0, s>>>(t->v)));
    HIP_CHECK_LAUNCH();
    unsigned long long failed = 0;
    HIP_TRY(hipMemcpyAsync(&failed, t->scratch, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (failed) return fail(HIVE_ECAPACITY, "import: more stash buckets than the stash pool holds");
    k_fill_u32<<<1, 64, 0, s>>>(t->v.holes, 1, 1u);  // an imported state may hold hint bits on non-full buckets
    HIP_CHECK_LAUNCH();
    uint64_t a = 0, b = 0;
    if ((r = count_exact_view(t, t->v, s, &a, &b)) != HIVE_OK) return r;
    return reset_counters(t, t->v, s, a, b);
  }
  HIP_TRY(hipMemcpyAsync(t->v.slots, h_slots, nb * t->v.S * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(t->v.stash, h_stash, nb * kStashSlots * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(t->v.stash_count, h_stash_count, nb * 4, hipMemcpyHostToDevice, s));
  k_fill_u32<<<fill_grid(meta_words(t->v.nb)), 256, 0, s>>>(t->v.meta, meta_words(t->v.nb), 0u);
  HIVE_POLICY_SWITCH(table_kind(t), (k_rebuild_meta<P><<<grid_for(nb * t->v.S, 256), 256, 0, s>>>(t->v)));
  k_fill_u32<<<fill_grid(t->v.nb), 256, 0, s>>>(t->v.locks, t->v.nb, 0u);
  k_fill_u32<<<1, 64, 0, s>>>(t->v.holes, 1, 1u);  // an imported state may hold hint bits on non-full buckets
  HIP_CHECK_LAUNCH();
  uint64_t a = 0, b = 0;
  if ((r = count_exact_view(t, t->v, s, &a, &b)) != HIVE_OK) return r;
  return reset_counters(t, t->v, s, a, b);
}

hive_status_t hive_set_launch(hive_table_t* t, int op_type, const hive_launch_t* cfg) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (!cfg || op_type < 0 || op_type > 2) return fail(HIVE_EINVAL, "bad op type or null config");
  hive_launch_t c = *cfg;
  if (c.block_threads == 0) c.block_threads = kDefaultBlock;
  if (c.keys_per_lane == 0) c.keys_per_lane = 1;
  if (!valid_block(c.block_threads)) return fail(HIVE_EINVAL, "block_threads must be a multiple of 64 in 64..1024");
  if (c.keys_per_lane != 1 && c.keys_per_lane != 2 && c.keys_per_lane != 4)
    return fail(HIVE_EINVAL, "keys_per_lane must be 1, 2 or 4");
  if (c.keys_per_lane != 1 && op_type != OP_LOOKUP) return fail(HIVE_EINVAL, "keys_per_lane > 1 is for lookups");
  if (c.inline_slow_path > 2) return fail(HIVE_EINVAL, "inline_slow_path must be 0, 1 or 2");
  if (c.inline_slow_path == 2 && op_type != OP_INSERT) return fail(HIVE_EINVAL, "inline_slow_path 2 is for inserts");
  if (c.combine_duplicates > 1) return fail(HIVE_EINVAL, "combine_duplicates must be 0 or 1");
  if (c.split_by_type > 1) return fail(HIVE_EINVAL, "split_by_type must be 0 or 1");
  if (c.lanes_per_key == 0) c.lanes_per_key = 1;
  if (c.lanes_per_key != 1 && c.lanes_per_key != 4 && c.lanes_per_key != 8)
    return fail(HIVE_EINVAL, "lanes_per_key must be 1, 4 or 8");
  if (c.lanes_per_key != 1 && op_type != OP_LOOKUP) return fail(HIVE_EINVAL, "lanes_per_key > 1 is for lookups");
  if (c.lanes_per_key != 1 && c.keys_per_lane != 1)
    return fail(HIVE_EINVAL, "lanes_per_key > 1 and keys_per_lane > 1 exclude each other");
  t->launch[op_type] = c;
  return HIVE_OK;
}

hive_status_t hive_get_launch(const hive_table_t* t, int op_type, hive_launch_t* cfg) {
  if (!t || !cfg || op_type < 0 || op_type > 2) return fail(HIVE_EINVAL, "bad argument");
  *cfg = t->launch[op_type];
  return HIVE_OK;
}

hive_status_t hive_ceiling_gather(const uint64_t* buf, uint64_t footprint_bytes, uint64_t n, uint32_t seed,
                                  uint32_t* sink, void* stream) {
  const uint64_t nbk = footprint_bytes / 64;
  if (!buf || nbk == 0 || nbk > 0xFFFFFFFFull) return fail(HIVE_EINVAL, "bad ceiling buffer");
  if (n == 0) return HIVE_OK;
  k_ceiling_gather<kRefSlots><<<grid_for(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      reinterpret_cast<const unsigned long long*>(buf), static_cast<uint32_t>(nbk), n, seed, sink);
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}

hive_status_t hive_ceiling_gather_bytes(const uint64_t* buf, uint64_t footprint_bytes, uint64_t n,
                                       uint32_t bucket_bytes, uint32_t seed, uint32_t* sink, void* stream) {
  if (bucket_bytes == 64) return hive_ceiling_gather(buf, footprint_bytes, n, seed, sink, stream);
  if (bucket_bytes != 128) return fail(HIVE_EINVAL, "bucket_bytes must be 64 or 128");
  const uint64_t nbk = footprint_bytes / 128;
  if (!buf || nbk == 0 || nbk > 0xFFFFFFFFull) return fail(HIVE_EINVAL, "bad ceiling buffer");
  if (n == 0) return HIVE_OK;
  k_ceiling_gather<16><<<grid_for(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      reinterpret_cast<const unsigned long long*>(buf), static_cast<uint32_t>(nbk), n, seed, sink);
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}

hive_status_t hive_ceiling_gather_lanes(const uint64_t* buf, uint64_t footprint_bytes, uint64_t n,
                                        uint32_t lanes_per_bucket, uint32_t seed, uint32_t* sink, void* stream) {
  if (lanes_per_bucket == 1) return hive_ceiling_gather(buf, footprint_bytes, n, seed, sink, stream);
  const uint64_t nbk = footprint_bytes / 64;
  if (!buf || nbk == 0 || nbk > 0xFFFFFFFFull) return fail(HIVE_EINVAL, "bad ceiling buffer");
  if (n == 0) return HIVE_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const auto* b = reinterpret_cast<const unsigned long long*>(buf);
  const uint32_t nb32 = static_cast<uint32_t>(nbk);
  switch (lanes_per_bucket) {
    case 4: k_ceiling_gather_coop<4><<<grid_for(n * 4, 256), 256, 0, s>>>(b, nb32, n, seed, sink); break;
    case 8: k_ceiling_gather_coop<8><<<grid_for(n * 8, 256), 256, 0, s>>>(b, nb32, n, seed, sink); break;
    case 16: k_ceiling_gather_coop<16><<<grid_for(n * 16, 256), 256, 0, s>>>(b, nb32, n, seed, sink); break;
    default: return fail(HIVE_EINVAL, "lanes_per_bucket must be 1, 4, 8 or 16");
  }
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}

hive_status_t hive_ceiling_cas(uint64_t* buf, uint64_t footprint_bytes, uint64_t n, uint32_t seed, void* stream) {
  const uint64_t nw = footprint_bytes / 8;
  if (!buf || nw == 0) return fail(HIVE_EINVAL, "bad ceiling buffer");
  if (n == 0) return HIVE_OK;
  k_ceiling_cas<<<grid_for(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      reinterpret_cast<unsigned long long*>(buf), nw, n, seed);
  HIP_CHECK_LAUNCH();
  return HIVE_OK;
}

}  // extern "C"
