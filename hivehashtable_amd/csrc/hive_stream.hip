// hive_stream.hip -- streams of uneven mixed batches: per-batch launches, one
// captured hipGraph, or one persistent kernel fed by a device work queue.
//
// No reference counterpart executes a stream: its harness launches one
// hive_mixed_kernel per batch and synchronizes (src/hive_dispatch.cu:119-164).
// Every mode keeps the per-batch semantics of hive_mixed (lookups, then
// deletes, then inserts; batch b+1 starts after batch b), so all three give
// the same results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "hive_internal.hpp"

using namespace hive;

namespace {

__global__ void k_stamp(unsigned long long* ts) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    __hip_atomic_store(ts, static_cast<unsigned long long>(wall_clock64()), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct StreamArgs {
  const uint8_t* types;
  const uint32_t* keys;
  const uint32_t* vals;
  uint32_t* out_vals;
  uint8_t* status;
  const uint4* tile_desc;      // per ticket: {phase q, first op, end op, batch}
  uint32_t ntiles;             // tickets in the stream
  unsigned int* ticket;        // next tile to claim
  unsigned int* tiles_left;    // per phase, one 128-byte line each: tiles not yet completed
  unsigned int* phases_done;   // phases 0 .. phases_done-1 are complete
  unsigned long long* ts;      // nbatch + 1: ts[0] stream start, ts[b + 1] batch b done
  unsigned int* error;         // set when a dependency wait gives up
  unsigned long long wait_ticks;  // wall-clock ticks one dependency wait may take before it gives up
};

constexpr uint32_t kStreamBlock = 256;
constexpr uint32_t kDoneStride = 32;        // u32 words per phase counter (own 128-byte line)
constexpr double kWaitLimitSeconds = 10.0;  // bounded waits: a stuck dependency ends the kernel
constexpr uint32_t kMaxTileOps = 4096;

// One thread per ticket: its descriptor, and (first tile of a phase) the
// phase's count of tiles left.
__global__ void k_tile_desc(const uint32_t* phase_start, uint32_t nphase, const unsigned long long* batch_off,
                            const uint32_t* tile_ops, uint4* desc, uint32_t ntiles, unsigned int* tiles_left) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ntiles) return;
  uint32_t lo = 0, hi = nphase;  // phase_start[lo] <= k < phase_start[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (phase_start[mid] <= k) lo = mid; else hi = mid;
  }
  const uint32_t q = lo, b = q / 3, tops = tile_ops[b];
  const unsigned long long bend = batch_off[b + 1];
  const unsigned long long first = batch_off[b] + static_cast<unsigned long long>(k - phase_start[q]) * tops;
  const unsigned long long end = first + tops < bend ? first + tops : bend;
  desc[k] = make_uint4(q, static_cast<uint32_t>(first), static_cast<uint32_t>(end), b);
  if (k == phase_start[q]) tiles_left[static_cast<size_t>(q) * kDoneStride] = phase_start[q + 1] - phase_start[q];
}

// Wait until phase r is complete.  Phases complete in order, and the last
// tile of a phase raises phases_done, so a block waiting far ahead backs off
// in proportion to its distance on that one watermark; only a block waiting
// on the phase that is running polls that phase's own tile counter.
__device__ bool wait_phase(const StreamArgs& a, uint32_t r) {
  const unsigned long long t0 = wall_clock64();
  uint32_t spins = 0;
  for (;;) {
    const uint32_t pd = ld_sync32(a.phases_done);
    if (pd > r) return true;
    if (pd == r) {
      while (ld_sync32(a.tiles_left + static_cast<size_t>(r) * kDoneStride) != 0u) {
        if ((++spins & 255u) == 0 && (ld_sync32(a.error) != 0u || wall_clock64() - t0 > a.wait_ticks)) return false;
        __builtin_amdgcn_s_sleep(1);
      }
      return true;
    }
    if (ld_sync32(a.error) != 0u || wall_clock64() - t0 > a.wait_ticks) return false;
    for (uint32_t d = min(r - pd, 16u); d > 0; --d) __builtin_amdgcn_s_sleep(32);
  }
}

// Persistent kernel.  Phases are numbered q = 3*b + {0 lookup, 1 delete, 2
// insert}.  A block claims tile tickets one ahead of the tile it works on;
// tickets are handed out in phase order, so the lowest unfinished tile always
// has an owner whose earlier tile is done and whose dependencies (all lower
// tickets) are complete: every wait ends.  A completed phase q-1 implies
// every earlier phase completed, so a block waits only when a ticket needs a
// later phase than it has seen complete.  No kernel boundary separates phases,
// so every table read is a device-coherent load (CoherentReads): no L1/L2 line
// that another XCD's writes made stale is ever read, and no cache invalidation
// is needed between phases.  Every finished tile's writes are performed
// (s_waitcnt vmcnt(0)) before its phase counter drops.  The first chunk of a
// tile's inputs (read-only) is loaded before the wait.
template <class P0, bool Inline>
__global__ __launch_bounds__(kStreamBlock) void k_stream(TableView t, StreamArgs a) {
  using P = CoherentReads<P0>;
  __shared__ uint4 s_desc;
  __shared__ uint32_t s_abort, qn;
  extern __shared__ uint32_t queue[];  // staged slow path: one entry per op of the largest tile
  int dtab = 0, dstash = 0, dslow = 0;
  int known = -1;  // thread 0: highest phase this block has seen complete
  uint32_t next_k = 0;
  uint4 next_d = make_uint4(0, 0, 0, 0);
  if (threadIdx.x == 0) {
    next_k = atomicAdd(a.ticket, 1u);
    if (next_k < a.ntiles) next_d = a.tile_desc[next_k];
  }
  for (;;) {
    if (threadIdx.x == 0) {
      s_desc = next_k < a.ntiles ? next_d : make_uint4(0xFFFFFFFFu, 0, 0, 0);
      qn = 0;
    }
    __syncthreads();
    const uint4 d = s_desc;
    if (d.x == 0xFFFFFFFFu) break;
    const uint32_t q = d.x, b = d.w, p = q % 3;
    const size_t lo = d.y, hi = d.z;
    // inputs of the first chunk: immutable, so loaded ahead of the dependency wait
    size_t i = lo + threadIdx.x;
    int ty = i < hi ? static_cast<int>(a.types[i]) : -1;
    uint32_t key = i < hi ? a.keys[i] : 0u;
    uint32_t val = (i < hi && a.vals) ? a.vals[i] : key + 1u;
    if (threadIdx.x == 0) {
      next_k = atomicAdd(a.ticket, 1u);
      uint32_t abort = 0;
      if (static_cast<int>(q) - 1 > known) {
        if (wait_phase(a, q - 1)) {
          known = static_cast<int>(q) - 1;
        } else {
          atomicOr(a.error, 1u);
          abort = 1;
        }
      }
      if (next_k < a.ntiles) next_d = a.tile_desc[next_k];
      s_abort = abort;
    }
    __syncthreads();
    if (s_abort) break;
    for (size_t base = lo;;) {
      if (p == 0) {
        if (ty == OP_LOOKUP) {
          uint32_t v;
          const uint8_t r = lookup_one<P>(t, key, v);
          if (a.status) a.status[i] = r;
          if (a.out_vals) a.out_vals[i] = v;
        } else if (ty > OP_DELETE || ty < -1) {
          if (a.status) a.status[i] = R_NOOP;
          if (a.out_vals) a.out_vals[i] = 0;
        }
      } else if (p == 1) {
        if (ty == OP_DELETE) {
          int rt, rs;
          const uint8_t r = erase_one<P>(t, key, rt, rs);
          if (a.status) a.status[i] = r;
          if (a.out_vals) a.out_vals[i] = 0;
          dtab -= rt;
          dstash -= rs;
        }
      } else if (ty == OP_INSERT) {
        uint8_t r = fast_insert<P>(t, key, val);  // no full-hint rule: this kernel's own delete phases may open holes
        if (r == R_PENDING) {
          if constexpr (Inline) {
            r = slow_insert<P>(t, key, val);
            ++dslow;
            dstash += (r == R_STASHED);
          } else {
            queue[atomicAdd(&qn, 1u)] = static_cast<uint32_t>(i - lo);
          }
        }
        if (r != R_PENDING) {
          if (a.status) a.status[i] = r;
          if (a.out_vals) a.out_vals[i] = 0;
          dtab += (r == R_INSERTED);
        }
      }
      base += kStreamBlock;
      if (base >= hi) break;
      i = base + threadIdx.x;
      ty = i < hi ? static_cast<int>(a.types[i]) : -1;
      key = i < hi ? a.keys[i] : 0u;
      val = (i < hi && a.vals) ? a.vals[i] : key + 1u;
    }
    if constexpr (!Inline) {
      if (p == 2) {  // LDS-staged slow path of the whole tile on packed lanes
        __syncthreads();
        const uint32_t pend = qn;
        for (uint32_t j = threadIdx.x; j < pend; j += kStreamBlock) {
          const size_t idx = lo + queue[j];
          const uint32_t k2 = a.keys[idx];
          const uint8_t r = slow_insert<P>(t, k2, a.vals ? a.vals[idx] : k2 + 1u);
          if (a.status) a.status[idx] = r;
          if (a.out_vals) a.out_vals[idx] = 0;
          dtab += (r == R_INSERTED);
          dstash += (r == R_STASHED);
          ++dslow;
        }
      }
    }
    wait_vmem();  // this thread's table writes are performed before the tile is published
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t prev = atomicSub(a.tiles_left + static_cast<size_t>(q) * kDoneStride, 1u);
      if (prev == 1u) {
        atomicMax(a.phases_done, q + 1);
        if (p == 2)
          __hip_atomic_store(a.ts + b + 1, static_cast<unsigned long long>(wall_clock64()), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  wave_count(t, dtab, dstash, dslow);
}

hive_status_t enqueue_launches(hive_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                               const uint64_t* off, size_t nb, uint32_t* out_vals, uint8_t* status,
                               unsigned int* ctrs, unsigned long long* ts, hipStream_t s) {
  k_stamp<<<1, 64, 0, s>>>(ts);
  HIP_CHECK_LAUNCH();
  for (size_t b = 0; b < nb; ++b) {
    const size_t o = off[b], n = off[b + 1] - off[b];
    if (n == 0) {
      k_stamp<<<1, 64, 0, s>>>(ts + b + 1);
      HIP_CHECK_LAUNCH();
      continue;
    }
    SrcNative src{keys + o, vals ? vals + o : nullptr, types + o};
    OutNative out{status ? status + o : nullptr, out_vals ? out_vals + o : nullptr};
    hive_status_t r = launch_mixed_native(t, src, n, out, s, BatchClock{ctrs + b, ts + b + 1});
    if (r != HIVE_OK) return r;
  }
  return HIVE_OK;
}

}  // namespace

extern "C" {

hive_status_t hive_set_stream_shape(hive_table_t* t, uint32_t tile_ops, uint32_t blocks_per_cu,
                                    uint32_t inline_slow_path) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (tile_ops && (tile_ops % kStreamBlock != 0 || tile_ops > kMaxTileOps))
    return fail(HIVE_EINVAL, "tile_ops must be 0 or a multiple of 256 up to 4096");
  if (blocks_per_cu > 16) return fail(HIVE_EINVAL, "blocks_per_cu must be 0..16");
  if (inline_slow_path > 1) return fail(HIVE_EINVAL, "inline_slow_path must be 0 or 1");
  t->stream_tile = tile_ops;
  t->stream_blocks_per_cu = blocks_per_cu;
  t->stream_inline = inline_slow_path;
  return HIVE_OK;
}

hive_status_t hive_set_stream_window(hive_table_t* t, uint64_t window_ops, uint64_t batch_ops) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (window_ops && (window_ops < (1u << 10) || window_ops > (1u << 24)))
    return fail(HIVE_EINVAL, "window_ops must be 0 or 2^10..2^24");
  if (batch_ops > (1u << 24)) return fail(HIVE_EINVAL, "batch_ops must be 0..2^24");
  t->stream_window_ops = window_ops;
  t->stream_window_batch = batch_ops;
  return HIVE_OK;
}

hive_status_t hive_mixed_stream(hive_table_t* t, const uint8_t* types, const uint32_t* keys, const uint32_t* vals,
                                const uint64_t* h_offsets, size_t nbatches, uint32_t* out_vals, uint8_t* status,
                                int mode, double* batch_done_ns, double* setup_ms, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  if (!h_offsets) return fail(HIVE_EINVAL, "null offsets");
  if (setup_ms) *setup_ms = 0.0;
  if (nbatches == 0) return HIVE_OK;
  for (size_t b = 0; b < nbatches; ++b)
    if (h_offsets[b + 1] < h_offsets[b]) return fail(HIVE_EINVAL, "offsets must be non-decreasing");
  if (h_offsets[0] != 0) return fail(HIVE_EINVAL, "offsets[0] must be 0");
  if (h_offsets[nbatches] && (!keys || !types)) return fail(HIVE_EINVAL, "null keys or types");
  if (mode < HIVE_STREAM_LAUNCHES || mode > HIVE_STREAM_WINDOWED) return fail(HIVE_EINVAL, "unknown stream mode");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int dev = t->device, khz = 0, ncu = 0;
  HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  DevArena ar;
  unsigned long long* ts = nullptr;
  HIP_TRY(ar.alloc(&ts, nbatches + 1));
  HIP_TRY(hipMemsetAsync(ts, 0, (nbatches + 1) * 8, s));
  const auto t0 = std::chrono::steady_clock::now();
  if (mode == HIVE_STREAM_WINDOWED) {
    if (h_offsets[nbatches] > 0xFFFFFFFFull) return fail(HIVE_EINVAL, "windowed stream: more than 2^32-1 ops");
    const uint64_t wops = t->stream_window_ops ? t->stream_window_ops : kDefaultWindowOps;
    const uint64_t wbatch = t->stream_window_batch ? t->stream_window_batch : kDefaultWindowBatchOps;
    uint64_t maxn = wops;
    for (size_t b = 0; b < nbatches; ++b) maxn = std::max<uint64_t>(maxn, h_offsets[b + 1] - h_offsets[b]);
    if ((r = ensure_partition(t, maxn)) != HIVE_OK) return r;
    if ((r = ensure_window(t, wops)) != HIVE_OK) return r;
    std::vector<WindowPlan> plan;
    std::vector<uint4> desc;
    if ((r = plan_windows(h_offsets, nbatches, wops, wbatch, plan, desc)) != HIVE_OK) return r;
    uint4* d_desc = nullptr;
    unsigned int* ctrs = nullptr;
    HIP_TRY(ar.alloc(&d_desc, desc.size()));
    HIP_TRY(ar.alloc(&ctrs, nbatches));
    if (!desc.empty())
      HIP_TRY(hipMemcpyAsync(d_desc, desc.data(), desc.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(ctrs, 0, nbatches * 4, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (setup_ms)
      *setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    k_stamp<<<1, 64, 0, s>>>(ts);
    HIP_CHECK_LAUNCH();
    if ((r = enqueue_windows(t, types, keys, vals, h_offsets, out_vals, status, ctrs, ts, plan, d_desc, ncu, s)) !=
        HIVE_OK)
      return r;
  } else if (mode == HIVE_STREAM_LAUNCHES || mode == HIVE_STREAM_GRAPH) {
    uint64_t maxn = 0;
    for (size_t b = 0; b < nbatches; ++b) maxn = std::max<uint64_t>(maxn, h_offsets[b + 1] - h_offsets[b]);
    if ((r = ensure_partition(t, maxn)) != HIVE_OK) return r;  // before any capture
    unsigned int* ctrs = nullptr;
    HIP_TRY(ar.alloc(&ctrs, nbatches));
    HIP_TRY(hipMemsetAsync(ctrs, 0, nbatches * 4, s));
    if (mode == HIVE_STREAM_LAUNCHES) {
      if ((r = enqueue_launches(t, types, keys, vals, h_offsets, nbatches, out_vals, status, ctrs, ts, s)) != HIVE_OK)
        return r;
    } else {
      // capture on a private stream (the caller's may be the legacy default stream)
      hipStream_t cs = nullptr;
      hipEvent_t ready = nullptr;
      hipGraph_t g = nullptr;
      hipGraphExec_t ge = nullptr;
      HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      hive_status_t rc = HIVE_OK;
      if (hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess ||
          hipEventRecord(ready, s) != hipSuccess || hipStreamWaitEvent(cs, ready, 0) != hipSuccess ||
          hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = fail(HIVE_EDEVICE, "graph capture setup");
      } else {
        rc = enqueue_launches(t, types, keys, vals, h_offsets, nbatches, out_vals, status, ctrs, ts, cs);
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (rc == HIVE_OK && e != hipSuccess) rc = fail(HIVE_EDEVICE, "graph capture");
        if (rc == HIVE_OK && hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess)
          rc = fail(HIVE_EDEVICE, "graph instantiate");
      }
      if (rc == HIVE_OK && setup_ms)
        *setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (rc == HIVE_OK && (hipGraphLaunch(ge, cs) != hipSuccess || hipEventRecord(ready, cs) != hipSuccess ||
                            hipStreamWaitEvent(s, ready, 0) != hipSuccess))
        rc = fail(HIVE_EDEVICE, "graph launch");
      if (cs) (void)hipStreamSynchronize(cs);
      if (ge) (void)hipGraphExecDestroy(ge);
      if (g) (void)hipGraphDestroy(g);
      if (ready) (void)hipEventDestroy(ready);
      if (cs) (void)hipStreamDestroy(cs);
      if (rc != HIVE_OK) return rc;
    }
  } else {
    // persistent kernel: per-batch tile size, phase prefix, per-ticket descriptors
    if (h_offsets[nbatches] > 0xFFFFFFFFull) return fail(HIVE_EINVAL, "persistent stream: more than 2^32-1 ops");
    const void* kfn = nullptr;
    if (t->stream_inline) {
      HIVE_POLICY_SWITCH(table_kind(t), (kfn = reinterpret_cast<const void*>(k_stream<P, true>)));
    } else {
      HIVE_POLICY_SWITCH(table_kind(t), (kfn = reinterpret_cast<const void*>(k_stream<P, false>)));
    }
    int occ = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kfn, kStreamBlock, t->stream_inline ? 0 : kMaxTileOps * 4));
    unsigned bpc = t->stream_blocks_per_cu ? t->stream_blocks_per_cu : 4;
    if (occ > 0 && bpc > static_cast<unsigned>(occ)) bpc = static_cast<unsigned>(occ);  // every block resident
    const unsigned grid = static_cast<unsigned>(ncu) * bpc;
    std::vector<uint32_t> tile_ops(nbatches), phase_start(3 * nbatches + 1);
    uint64_t acc = 0;
    for (size_t b = 0; b < nbatches; ++b) {
      const uint64_t n = h_offsets[b + 1] - h_offsets[b];
      uint32_t tops = t->stream_tile;
      if (!tops) {  // about one tile per resident block, 256..4096 ops
        uint64_t want = (n + grid - 1) / grid;
        want = (want + kStreamBlock - 1) / kStreamBlock * kStreamBlock;
        tops = static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(want, kStreamBlock), kMaxTileOps));
      }
      tile_ops[b] = tops;
      const uint64_t tiles = n ? (n + tops - 1) / tops : 1;  // an empty batch still completes through one tile
      for (int p = 0; p < 3; ++p) {
        phase_start[3 * b + p] = static_cast<uint32_t>(acc);
        acc += tiles;
      }
    }
    if (acc >= 0xFFFFFFFFull - grid) return fail(HIVE_EINVAL, "stream too long for one persistent launch");
    phase_start[3 * nbatches] = static_cast<uint32_t>(acc);
    const uint32_t ntiles = static_cast<uint32_t>(acc), nphase = static_cast<uint32_t>(3 * nbatches);
    unsigned long long* d_off = nullptr;
    uint32_t *d_tile = nullptr, *d_phase = nullptr;
    uint4* d_desc = nullptr;
    unsigned int *d_ctl = nullptr, *d_left = nullptr;
    HIP_TRY(ar.alloc(&d_off, nbatches + 1));
    HIP_TRY(ar.alloc(&d_tile, nbatches));
    HIP_TRY(ar.alloc(&d_phase, nphase + 1));
    HIP_TRY(ar.alloc(&d_desc, ntiles));
    HIP_TRY(ar.alloc(&d_ctl, 3 * kDoneStride));  // ticket, phases_done, error: one line each
    HIP_TRY(ar.alloc(&d_left, static_cast<size_t>(nphase) * kDoneStride));
    HIP_TRY(hipMemcpyAsync(d_off, h_offsets, (nbatches + 1) * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_tile, tile_ops.data(), nbatches * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_phase, phase_start.data(), (nphase + 1) * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(d_ctl, 0, 3 * kDoneStride * 4, s));
    k_tile_desc<<<(ntiles + 255) / 256, 256, 0, s>>>(d_phase, nphase, d_off, d_tile, d_desc, ntiles, d_left);
    HIP_CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(s));
    if (setup_ms)
      *setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const uint32_t max_tile = *std::max_element(tile_ops.begin(), tile_ops.end());
    StreamArgs args{types, keys, vals, out_vals, status, d_desc, ntiles, d_ctl, d_left,
                    d_ctl + kDoneStride, ts, d_ctl + 2 * kDoneStride,
                    static_cast<unsigned long long>(kWaitLimitSeconds * 1e3 * khz)};
    k_stamp<<<1, 64, 0, s>>>(ts);
    if (t->stream_inline) {
      HIVE_POLICY_SWITCH(table_kind(t), (k_stream<P, true><<<grid, kStreamBlock, 0, s>>>(t->v, args)));
    } else {
      HIVE_POLICY_SWITCH(table_kind(t), (k_stream<P, false><<<grid, kStreamBlock, max_tile * 4, s>>>(t->v, args)));
    }
    HIP_CHECK_LAUNCH();
    unsigned int err = 0;
    HIP_TRY(hipMemcpyAsync(&err, d_ctl + 2 * kDoneStride, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (err) return fail(HIVE_EDEVICE, "persistent stream: a phase dependency did not complete");
  }
  std::vector<unsigned long long> hts(nbatches + 1);
  HIP_TRY(hipMemcpyAsync(hts.data(), ts, (nbatches + 1) * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (batch_done_ns) {
    const double ns_per_tick = 1e6 / static_cast<double>(khz);
    for (size_t b = 0; b < nbatches; ++b) batch_done_ns[b] = static_cast<double>(hts[b + 1] - hts[0]) * ns_per_tick;
  }
  return HIVE_OK;
}

}  // extern "C"

namespace {

hive_status_t copy_view(const TableView& dst, const TableView& src, hipStream_t s) {
  const size_t nb = src.nb;
  HIP_TRY(hipMemcpyAsync(dst.counters, src.counters, counter_bytes(), hipMemcpyDeviceToDevice, s));
  if (src.layout == kLayoutInline) {
    HIP_TRY(hipMemcpyAsync(dst.slots, src.slots, nb * kInlineStride * 8, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(dst.stash, src.stash, static_cast<size_t>(src.pool_cap) * kStashSlots * 8,
                           hipMemcpyDeviceToDevice, s));
    return HIVE_OK;
  }
  HIP_TRY(hipMemcpyAsync(dst.slots, src.slots, nb * src.S * 8, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(dst.stash, src.stash, nb * kStashSlots * 8, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(dst.stash_count, src.stash_count, nb * 4, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(dst.meta, src.meta, meta_words(src.nb) * 4, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemsetAsync(dst.locks, 0, nb * 4, s));
  return HIVE_OK;
}

}  // namespace

namespace {

// p99 of the batches' service times: a batch's time from the previous
// distinct completion to its own (batches of one window complete together,
// and completions are in stream order), percentile by linear interpolation
double p99_service_us(const std::vector<double>& done_ns) {
  if (done_ns.empty()) return 0.0;
  std::vector<double> lat(done_ns.size());
  double prev = 0.0, cur = 0.0;
  for (size_t b = 0; b < done_ns.size(); ++b) {
    if (b == 0 || done_ns[b] != cur) {
      prev = b == 0 ? 0.0 : cur;
      cur = done_ns[b];
    }
    lat[b] = (done_ns[b] - prev) * 1e-3;
  }
  std::sort(lat.begin(), lat.end());
  const double pos = 0.99 * static_cast<double>(lat.size() - 1);
  const size_t lo = static_cast<size_t>(pos), hi = std::min(lo + 1, lat.size() - 1);
  return lat[lo] + (lat[hi] - lat[lo]) * (pos - static_cast<double>(lo));
}

}  // namespace

extern "C" hive_status_t hive_autotune_stream_ex(hive_table_t* t, const uint8_t* types, const uint32_t* keys,
                                                 const uint32_t* vals, const uint64_t* h_offsets, size_t nbatches,
                                                 const hive_tune_opts_t* opts, hive_tune_t* out, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  if (!out || !h_offsets || nbatches == 0) return fail(HIVE_EINVAL, "autotune needs a sample stream and an output");
  const hive_tune_opts_t o = opts ? *opts : hive_tune_opts_t{0u, -1.0, 0.1};
  if (o.try_persistent > 1 || !(o.max_p99_us == o.max_p99_us) || !(o.p99_slack >= 0.0))
    return fail(HIVE_EINVAL, "bad autotune options");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t nops = h_offsets[nbatches];
  uint64_t maxn = 0;
  for (size_t b = 0; b < nbatches; ++b) maxn = std::max<uint64_t>(maxn, h_offsets[b + 1] - h_offsets[b]);
  // windows of 2^15 .. 2^22 ops, each coalescing batches below 1/32 .. 1/4
  // of its size (2^10 .. 2^19): small windows finish within a big batch's
  // time, big ones amortize the window's fixed kernels over more ops
  constexpr int kWindowSizes = 8, kWindowLog0 = 15;
  const uint64_t wmax = uint64_t(1) << (kWindowLog0 + kWindowSizes - 1);
  // scratch shared with the copy below, so the copy never reallocates it
  if ((r = ensure_partition(t, std::max<uint64_t>(maxn, wmax))) != HIVE_OK) return r;
  if ((r = ensure_window(t, wmax)) != HIVE_OK) return r;
  hive_table tmp = *t;  // same policy and device, scratch copy of the table state
  if ((r = alloc_view(tmp.v, t->v.nb, layout_of(t->v))) != HIVE_OK) return r;
  struct ViewGuard {
    TableView* v;
    ~ViewGuard() { free_view(*v); }
  } guard{&tmp.v};
  DevArena ar;
  uint32_t* ov = nullptr;
  uint8_t* st = nullptr;
  HIP_TRY(ar.alloc(&ov, nops));
  HIP_TRY(ar.alloc(&st, nops));
  std::vector<double> done(nbatches);
  // one timed run of the sample: total ms and p99 batch service time
  auto run = [&](int mode, double* ms, double* p99) -> hive_status_t {
    hive_status_t rr = copy_view(tmp.v, t->v, s);
    if (rr != HIVE_OK) return rr;
    double setup = 0.0;
    rr = hive_mixed_stream(&tmp, types, keys, vals, h_offsets, nbatches, ov, st, mode, done.data(), &setup, stream);
    if (rr != HIVE_OK) return rr;
    *ms = done[nbatches - 1] * 1e-6;
    *p99 = p99_service_us(done);
    return HIVE_OK;
  };
  hive_tune_t best{};
  best.launches_ms = best.graph_ms = best.windowed_ms = 1e300;
  best.bounded_windowed_ms = 1e300;
  uint32_t cands = 0;
  const uint32_t blocks[] = {128, 256, 512, 1024};
  const bool any_big = maxn >= kCompactMinOps;  // the split only applies to big batches
  for (uint32_t split = any_big ? 0u : 1u; split < 2; ++split)
    for (uint32_t inl = 0; inl < 2; ++inl)
      for (uint32_t blk : blocks) {
        for (int op = 0; op < 3; ++op)
          tmp.launch[op] = hive_launch_t{blk, 1, inl, t->launch[op].combine_duplicates, split,
                                         t->launch[op].lanes_per_key};
        double ms = 0, p99 = 0;
        if ((r = run(HIVE_STREAM_LAUNCHES, &ms, &p99)) != HIVE_OK) return r;
        ++cands;
        if (ms < best.launches_ms) {
          best.launches_ms = ms;
          best.launches_p99_us = p99;
          best.block_threads = blk;
          best.inline_slow_path = inl;
          best.split_by_type = split;
        }
      }
  const double bound = o.max_p99_us < 0.0 ? best.launches_p99_us * (1.0 + o.p99_slack) : o.max_p99_us;  // 0: none
  auto within = [&](double p99) { return bound <= 0.0 || p99 <= bound; };
  for (int op = 0; op < 3; ++op)
    tmp.launch[op] = hive_launch_t{best.block_threads, 1, best.inline_slow_path, t->launch[op].combine_duplicates,
                                   best.split_by_type, t->launch[op].lanes_per_key};
  if ((r = run(HIVE_STREAM_GRAPH, &best.graph_ms, &best.graph_p99_us)) != HIVE_OK) return r;
  ++cands;
  if (o.try_persistent) {
    best.persistent_ms = 1e300;
    const uint32_t tiles[] = {0, 256, 1024, 4096};
    const uint32_t bpcs[] = {2, 4, 8};
    for (uint32_t inl = 0; inl < 2; ++inl)
      for (uint32_t tile : tiles)
        for (uint32_t bpc : bpcs) {
          tmp.stream_tile = tile;
          tmp.stream_blocks_per_cu = bpc;
          tmp.stream_inline = inl;
          double ms = 0, p99 = 0;
          if ((r = run(HIVE_STREAM_PERSISTENT, &ms, &p99)) != HIVE_OK) return r;
          ++cands;
          if (ms < best.persistent_ms) {
            best.persistent_ms = ms;
            best.persistent_p99_us = p99;
            best.tile_ops = tile;
            best.blocks_per_cu = bpc;
            best.stream_inline = inl;
          }
        }
  }
  uint32_t win_split = best.split_by_type;
  {  // first touch of the window scratch, untimed
    double ms = 0, p99 = 0;
    tmp.stream_window_ops = wmax;
    tmp.stream_window_batch = wmax / 8;
    r = run(HIVE_STREAM_WINDOWED, &ms, &p99);
    t->win_epoch = tmp.win_epoch;
    if (r != HIVE_OK) return r;
  }
  for (uint32_t split = 0; split < 2; ++split)
    for (int wl = kWindowLog0; wl < kWindowLog0 + kWindowSizes; ++wl)
      for (int cl = wl - 5; cl <= wl - 2; ++cl) {
        const uint64_t w = uint64_t(1) << wl, cb = uint64_t(1) << cl;
        if (cl > 19) continue;
        // type compaction applies to windows of at least kCompactMinOps ops
        if (split && w < kCompactMinOps) continue;
        for (int op = 0; op < 3; ++op)
          tmp.launch[op] = hive_launch_t{best.block_threads, 1, best.inline_slow_path,
                                         t->launch[op].combine_duplicates, split, t->launch[op].lanes_per_key};
        tmp.stream_window_ops = w;
        tmp.stream_window_batch = cb;
        double ms = 0, p99 = 0;
        r = run(HIVE_STREAM_WINDOWED, &ms, &p99);
        t->win_epoch = tmp.win_epoch;  // the key set is shared: later windows continue its epochs
        if (r != HIVE_OK) return r;
        ++cands;
        double& by_size = best.windowed_ms_by_size[wl - kWindowLog0];
        if (by_size == 0.0 || ms < by_size) {
          by_size = ms;
          best.windowed_p99_us_by_size[wl - kWindowLog0] = p99;
        }
        if (ms < best.windowed_ms) {
          best.windowed_ms = ms;
          best.windowed_p99_us = p99;
          best.window_ops = w;
          best.window_batch_ops = cb;
          win_split = split;
        }
        if (within(p99) && ms < best.bounded_windowed_ms) {
          best.bounded_windowed_ms = ms;
          best.bounded_windowed_p99_us = p99;
          best.bounded_window_ops = w;
          best.bounded_window_batch_ops = cb;
          best.bounded_split_by_type = split;
        }
      }
  best.candidates = cands;
  best.p99_bound_us = bound > 0.0 ? bound : 0.0;
  const bool have_bounded = best.bounded_window_ops != 0;
  if (!have_bounded) best.bounded_windowed_ms = 0.0;
  // the fastest form within the bound (per-batch launches are always within
  // the default bound, which they define)
  double best_ms = 1e300;
  best.best_mode = HIVE_STREAM_LAUNCHES;
  auto consider = [&](int mode, double ms, double p99) {
    if (within(p99) && ms < best_ms) {
      best_ms = ms;
      best.best_mode = mode;
    }
  };
  consider(HIVE_STREAM_LAUNCHES, best.launches_ms, o.max_p99_us < 0.0 ? 0.0 : best.launches_p99_us);
  consider(HIVE_STREAM_GRAPH, best.graph_ms, best.graph_p99_us);
  if (o.try_persistent) consider(HIVE_STREAM_PERSISTENT, best.persistent_ms, best.persistent_p99_us);
  if (have_bounded) consider(HIVE_STREAM_WINDOWED, best.bounded_windowed_ms, best.bounded_windowed_p99_us);
  // the handle keeps the window within the bound (else the fastest window)
  const uint64_t keep_w = have_bounded ? best.bounded_window_ops : best.window_ops;
  const uint64_t keep_cb = have_bounded ? best.bounded_window_batch_ops : best.window_batch_ops;
  const uint32_t keep_split = have_bounded ? best.bounded_split_by_type : win_split;
  if (best.best_mode == HIVE_STREAM_WINDOWED) best.split_by_type = keep_split;
  for (int op = 0; op < 3; ++op)
    t->launch[op] = hive_launch_t{best.block_threads, 1, best.inline_slow_path, t->launch[op].combine_duplicates,
                                  best.split_by_type, t->launch[op].lanes_per_key};
  if (o.try_persistent) {
    t->stream_tile = best.tile_ops;
    t->stream_blocks_per_cu = best.blocks_per_cu;
    t->stream_inline = best.stream_inline;
  } else {
    best.persistent_ms = 0.0;
  }
  t->stream_window_ops = keep_w;
  t->stream_window_batch = keep_cb;
  *out = best;
  return HIVE_OK;
}

extern "C" hive_status_t hive_autotune_stream(hive_table_t* t, const uint8_t* types, const uint32_t* keys,
                                              const uint32_t* vals, const uint64_t* h_offsets, size_t nbatches,
                                              hive_tune_t* out, void* stream) {
  return hive_autotune_stream_ex(t, types, keys, vals, h_offsets, nbatches, nullptr, out, stream);
}

extern "C" hive_status_t hive_autotune_bulk(hive_table_t* t, const uint32_t* insert_keys, const uint32_t* insert_vals,
                                            size_t n_insert, const uint32_t* lookup_keys, size_t n_lookup,
                                            hive_bulk_tune_t* out, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  if (!out || (n_insert && !insert_keys) || (n_lookup && !lookup_keys)) return fail(HIVE_EINVAL, "bad sample or output");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hive_bulk_tune_t best{};
  best.insert_ms = best.lookup_ms = 1e300;
  for (double& v : best.lookup_ms_by_lanes) v = 1e300;
  for (double& v : best.lookup_ms_by_keys) v = 1e300;
  hive_stats_t st{};
  if ((r = hive_stats(t, &st, stream)) != HIVE_OK) return r;
  best.load_factor = st.load_factor;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  struct Events {
    hipEvent_t* a;
    hipEvent_t* b;
    ~Events() {
      if (*a) (void)hipEventDestroy(*a);
      if (*b) (void)hipEventDestroy(*b);
    }
  } ev_guard{&e0, &e1};
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  auto timed = [&](auto&& launch, double* ms) -> hive_status_t {
    HIP_TRY(hipEventRecord(e0, s));
    hive_status_t rr = launch();
    if (rr != HIVE_OK) return rr;
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    return HIVE_OK;
  };
  const uint32_t blocks[] = {256, 512, 1024};
  uint32_t cands = 0;
  DevArena ar;
  if (n_insert) {
    // the split form's pending list, shared with the scratch copy below
    if ((r = ensure_pending(t, n_insert, s, true)) != HIVE_OK) return r;
    hive_table tmp = *t;  // scratch copy of the table state, same policy and device
    if ((r = alloc_view(tmp.v, t->v.nb, layout_of(t->v))) != HIVE_OK) return r;
    struct ViewGuard {
      TableView* v;
      ~ViewGuard() { free_view(*v); }
    } guard{&tmp.v};
    uint8_t* st8 = nullptr;
    HIP_TRY(ar.alloc(&st8, n_insert));
    // slow path staged in LDS (0), inline (1), or in a second launch (2;
    // only when the pending list could be allocated)
    for (uint32_t inl = 0; inl < (t->pend_cap >= n_insert ? 3u : 2u); ++inl)
      for (uint32_t blk : blocks) {
        tmp.launch[OP_INSERT] = hive_launch_t{blk, 1, inl, 0, t->launch[OP_INSERT].split_by_type, 1};
        if ((r = copy_view(tmp.v, t->v, s)) != HIVE_OK) return r;
        double ms = 0;
        r = timed([&] { return launch_insert(&tmp, SrcNative{insert_keys, insert_vals, nullptr}, n_insert,
                                             OutNative{st8, nullptr}, -1, s); },
                  &ms);
        if (r != HIVE_OK) return r;
        ++cands;
        double& by_mode = best.insert_ms_by_slow_path[inl];
        if (by_mode == 0.0 || ms < by_mode) by_mode = ms;
        if (ms < best.insert_ms) {
          best.insert_ms = ms;
          best.insert_block_threads = blk;
          best.insert_inline_slow_path = inl;
        }
      }
  }
  if (n_lookup) {
    uint8_t* st8 = nullptr;
    uint32_t* v32 = nullptr;
    HIP_TRY(ar.alloc(&st8, n_lookup));
    HIP_TRY(ar.alloc(&v32, n_lookup));
    const hive_launch_t saved = t->launch[OP_LOOKUP];
    // (lanes per key, keys per lane): the lane-per-key form, the
    // bucket-cooperative forms, and lane-per-key with several keys per lane
    const uint32_t lanes[] = {1, 4, 8, 1, 1};
    const uint32_t keys[] = {1, 1, 1, 2, 4};
    for (int li = 0; li < 5; ++li)
      for (uint32_t blk : blocks) {
        t->launch[OP_LOOKUP] = hive_launch_t{blk, keys[li], 0, 0, saved.split_by_type, lanes[li]};
        double ms = 1e300;
        for (int rep = 0; rep < 3; ++rep) {  // read-only: repeat and keep the fastest
          double m = 0;
          r = timed([&] { return launch_lookup(t, SrcNative{lookup_keys, nullptr, nullptr}, n_lookup,
                                               OutNative{st8, v32}, -1, s); },
                    &m);
          if (r != HIVE_OK) {
            t->launch[OP_LOOKUP] = saved;
            return r;
          }
          ms = std::min(ms, m);
        }
        ++cands;
        if (keys[li] == 1) best.lookup_ms_by_lanes[li] = std::min(best.lookup_ms_by_lanes[li], ms);
        if (lanes[li] == 1) {
          const int ki = keys[li] == 1 ? 0 : (keys[li] == 2 ? 1 : 2);
          best.lookup_ms_by_keys[ki] = std::min(best.lookup_ms_by_keys[ki], ms);
        }
        if (ms < best.lookup_ms) {
          best.lookup_ms = ms;
          best.lookup_block_threads = blk;
          best.lookup_lanes_per_key = lanes[li];
          best.lookup_keys_per_lane = keys[li];
        }
      }
    t->launch[OP_LOOKUP] = saved;
  }
  best.candidates = cands;
  if (n_insert) {
    t->launch[OP_INSERT].block_threads = best.insert_block_threads;
    t->launch[OP_INSERT].inline_slow_path = best.insert_inline_slow_path;
  }
  if (n_lookup) {
    t->launch[OP_LOOKUP].block_threads = best.lookup_block_threads;
    t->launch[OP_LOOKUP].lanes_per_key = best.lookup_lanes_per_key;
    t->launch[OP_LOOKUP].keys_per_lane = best.lookup_keys_per_lane;
  }
  *out = best;
  return HIVE_OK;
}

// Layout choice: the sample insert, then the sample lookup, run on a copy of
// the table in each layout; the layout with the least insert + lookup time
// is applied to the table (its content kept, rehashed into that layout).
extern "C" hive_status_t hive_autotune_layout(hive_table_t* t, const uint32_t* insert_keys,
                                              const uint32_t* insert_vals, size_t n_insert,
                                              const uint32_t* lookup_keys, size_t n_lookup, hive_layout_tune_t* out,
                                              void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  if (!out || (n_insert && !insert_keys) || (n_lookup && !lookup_keys) || n_insert + n_lookup == 0)
    return fail(HIVE_EINVAL, "layout tuning needs a non-empty sample and an output");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hive_layout_tune_t best{};
  best.layout = layout_of(t->v);
  double best_ms = 1e300;
  for (int l = 0; l < 3; ++l) best.insert_ms[l] = best.lookup_ms[l] = -1.0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  struct Events {
    hipEvent_t* a;
    hipEvent_t* b;
    ~Events() {
      if (*a) (void)hipEventDestroy(*a);
      if (*b) (void)hipEventDestroy(*b);
    }
  } ev_guard{&e0, &e1};
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  auto timed = [&](auto&& launch, double* ms) -> hive_status_t {
    HIP_TRY(hipEventRecord(e0, s));
    hive_status_t rr = launch();
    if (rr != HIVE_OK) return rr;
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    return HIVE_OK;
  };
  DevArena ar;
  uint8_t* st8 = nullptr;
  uint32_t* v32 = nullptr;
  HIP_TRY(ar.alloc(&st8, std::max<size_t>(std::max(n_insert, n_lookup), 1)));
  if (n_lookup) HIP_TRY(ar.alloc(&v32, n_lookup));
  const int start = layout_of(t->v);
  const int order[3] = {start, (start + 1) % 3, (start + 2) % 3};  // the current layout first: no rehash
  for (int layout : order) {
    r = hive_set_layout(t, layout, stream);
    if (r == HIVE_ECAPACITY) continue;  // the content does not fit this layout: not a candidate
    if (r != HIVE_OK) return r;
    hive_table tmp = *t;  // scratch copy of the table state in this layout
    if ((r = alloc_view(tmp.v, t->v.nb, layout)) != HIVE_OK) return r;
    struct ViewGuard {
      TableView* v;
      ~ViewGuard() { free_view(*v); }
    } guard{&tmp.v};
    double ins = 1e300, look = 1e300;
    for (int rep = 0; rep < 2; ++rep) {  // keep the faster of two runs from the same state
      if ((r = copy_view(tmp.v, t->v, s)) != HIVE_OK) return r;
      double m = 0.0;
      if (n_insert) {
        r = timed([&] { return launch_insert(&tmp, SrcNative{insert_keys, insert_vals, nullptr}, n_insert,
                                             OutNative{st8, nullptr}, -1, s); },
                  &m);
        if (r != HIVE_OK) return r;
        ins = std::min(ins, m);
      }
      if (n_lookup) {
        r = timed([&] { return launch_lookup(&tmp, SrcNative{lookup_keys, nullptr, nullptr}, n_lookup,
                                             OutNative{st8, v32}, -1, s); },
                  &m);
        if (r != HIVE_OK) return r;
        look = std::min(look, m);
      }
    }
    ++best.candidates;
    best.insert_ms[layout] = n_insert ? ins : 0.0;
    best.lookup_ms[layout] = n_lookup ? look : 0.0;
    const double total = best.insert_ms[layout] + best.lookup_ms[layout];
    if (total < best_ms) {
      best_ms = total;
      best.layout = layout;
    }
  }
  if (best.candidates == 0) return fail(HIVE_ECAPACITY, "the table's content fits no layout at this size");
  if ((r = hive_set_layout(t, best.layout, stream)) != HIVE_OK) return r;
  *out = best;
  return HIVE_OK;
}
