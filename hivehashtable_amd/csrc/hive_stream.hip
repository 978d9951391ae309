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
  const unsigned long long* batch_off;  // nbatch + 1 op offsets
  const uint32_t* tile_ops;             // ops per tile of each batch
  const uint32_t* phase_start;          // 3 * nbatch + 1 tile prefix over (batch, phase)
  uint32_t nphase;                      // 3 * nbatch
  unsigned int* ticket;                 // next tile to claim
  unsigned int* done_tiles;             // per phase: tiles completed
  unsigned long long* ts;               // nbatch + 1: ts[0] stream start, ts[b + 1] batch b done
  unsigned int* error;                  // set when a dependency wait gives up
  uint32_t max_tile;                    // largest tile_ops (LDS queue entries)
  uint32_t phases_in_lds;               // phase_start copied into LDS after the queue
};

constexpr uint32_t kStreamBlock = 256;
constexpr uint32_t kSpinLimit = 1u << 26;  // ~seconds of s_sleep: a stuck dependency ends the kernel

// Persistent kernel.  Phases are numbered q = 3*b + {0 lookup, 1 delete, 2
// insert}.  A block claims the next tile ticket; tickets are handed out in
// phase order, so every tile of the phases before the claimed one is owned by
// a running or finished block and the wait for phase q-1 always ends.  A
// completed phase q-1 implies every earlier phase completed, so each block
// keeps the highest phase it has seen complete and waits (with a dependency
// poll) and acquires (buffer_inv sc1, dropping L1/L2 lines that other XCDs'
// writes made stale -- no kernel boundary separates phases) only when a
// ticket needs a later one.  The tile-to-phase map is read from LDS.
constexpr uint32_t kLdsPhases = 8192;  // phase_start entries cached in LDS (32 KiB at most)
constexpr uint32_t kMaxTileOps = 4096;

template <class P>
__global__ __launch_bounds__(kStreamBlock) void k_stream(TableView t, StreamArgs a) {
  __shared__ uint32_t s_k, s_q, s_abort, qn;
  __shared__ int s_known;  // highest phase this block has seen complete (and acquired after)
  extern __shared__ uint32_t dyn_lds[];  // [max_tile] slow-path queue, then [nphase + 1] phase map
  uint32_t* queue = dyn_lds;
  uint32_t* lds_phase = dyn_lds + a.max_tile;
  const bool phase_in_lds = a.phases_in_lds != 0;
  if (phase_in_lds)
    for (uint32_t i = threadIdx.x; i <= a.nphase; i += kStreamBlock) lds_phase[i] = a.phase_start[i];
  if (threadIdx.x == 0) s_known = -1;
  __syncthreads();
  const uint32_t* ps = phase_in_lds ? lds_phase : a.phase_start;
  int dtab = 0, dstash = 0, dslow = 0;
  const uint32_t total = ps[a.nphase];
  for (;;) {
    if (threadIdx.x == 0) {
      const uint32_t k = atomicAdd(a.ticket, 1u);
      uint32_t q = 0, abort = 0;
      if (k < total) {
        uint32_t lo = 0, hi = a.nphase;  // ps[lo] <= k < ps[hi]
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (ps[mid] <= k) lo = mid; else hi = mid;
        }
        q = lo;
        if (static_cast<int>(q) - 1 > s_known) {
          const uint32_t need = ps[q] - ps[q - 1];
          uint32_t spins = 0;
          while (ld_sync32(a.done_tiles + q - 1) < need) {
            if (ld_sync32(a.error) != 0u || ++spins > kSpinLimit) {
              atomicOr(a.error, 1u);
              abort = 1;
              break;
            }
            __builtin_amdgcn_s_sleep(2);
          }
          if (!abort) s_known = -2 - static_cast<int>(q - 1);  // new: acquire below, then record
        }
      }
      s_k = k;
      s_q = q;
      s_abort = abort;
      qn = 0;
    }
    __syncthreads();
    const uint32_t k = s_k, q = s_q;
    if (k >= total || s_abort) break;
    if (s_known < -1) {  // this block just saw a later phase complete
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __syncthreads();
      if (threadIdx.x == 0) s_known = -2 - s_known;
    }
    const uint32_t b = q / 3, p = q % 3;
    const uint32_t tile = k - ps[q];
    const size_t bend = a.batch_off[b + 1];
    const uint32_t tops = a.tile_ops[b];
    const size_t lo = a.batch_off[b] + static_cast<size_t>(tile) * tops;
    const size_t hi = lo + tops < bend ? lo + tops : bend;
    for (size_t base = lo; base < hi; base += kStreamBlock) {
      const size_t i = base + threadIdx.x;
      const int ty = i < hi ? static_cast<int>(a.types[i]) : -1;
      if (p == 0) {
        if (ty == OP_LOOKUP) {
          uint32_t v;
          const uint8_t r = lookup_one<P>(t, a.keys[i], v);
          if (a.status) a.status[i] = r;
          if (a.out_vals) a.out_vals[i] = v;
        } else if (ty > OP_DELETE || ty < -1) {
          if (a.status) a.status[i] = R_NOOP;
          if (a.out_vals) a.out_vals[i] = 0;
        }
      } else if (p == 1) {
        if (ty == OP_DELETE) {
          int rt, rs;
          const uint8_t r = erase_one<P>(t, a.keys[i], rt, rs);
          if (a.status) a.status[i] = r;
          if (a.out_vals) a.out_vals[i] = 0;
          dtab -= rt;
          dstash -= rs;
        }
      } else if (ty == OP_INSERT) {
        const uint32_t key = a.keys[i];
        const uint8_t r = fast_insert<P>(t, key, a.vals ? a.vals[i] : key + 1u);
        if (r == R_PENDING) {
          queue[atomicAdd(&qn, 1u)] = static_cast<uint32_t>(i - lo);
        } else {
          if (a.status) a.status[i] = r;
          if (a.out_vals) a.out_vals[i] = 0;
          dtab += (r == R_INSERTED);
        }
      }
    }
    if (p == 2) {  // LDS-staged slow path of the whole tile on packed lanes
      __syncthreads();
      const uint32_t pend = qn;
      for (uint32_t j = threadIdx.x; j < pend; j += kStreamBlock) {
        const size_t idx = lo + queue[j];
        const uint32_t key = a.keys[idx];
        const uint8_t r = slow_insert<P>(t, key, a.vals ? a.vals[idx] : key + 1u);
        if (a.status) a.status[idx] = r;
        if (a.out_vals) a.out_vals[idx] = 0;
        dtab += (r == R_INSERTED);
        dstash += (r == R_STASHED);
        ++dslow;
      }
    }
    wait_vmem();  // this thread's table writes are performed before the tile is published
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t need = ps[q + 1] - ps[q];
      const uint32_t prev = atomicAdd(a.done_tiles + q, 1u);
      if (p == 2 && prev + 1 == need)
        __hip_atomic_store(a.ts + b + 1, static_cast<unsigned long long>(wall_clock64()), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  wave_count(t, dtab, dstash, dslow);
}

struct DevArena {
  std::vector<void*> ptrs;
  ~DevArena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(T));
    if (e == hipSuccess) {
      ptrs.push_back(q);
      *p = static_cast<T*>(q);
    }
    return e;
  }
};

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
    hive_status_t r = launch_mixed(t, src, n, out, s, BatchClock{ctrs + b, ts + b + 1});
    if (r != HIVE_OK) return r;
  }
  return HIVE_OK;
}

}  // namespace

extern "C" {

hive_status_t hive_set_stream_shape(hive_table_t* t, uint32_t tile_ops, uint32_t blocks_per_cu) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK) return r;
  if (tile_ops && (tile_ops % kStreamBlock != 0 || tile_ops > kMaxTileOps))
    return fail(HIVE_EINVAL, "tile_ops must be 0 or a multiple of 256 up to 4096");
  if (blocks_per_cu > 16) return fail(HIVE_EINVAL, "blocks_per_cu must be 0..16");
  t->stream_tile = tile_ops;
  t->stream_blocks_per_cu = blocks_per_cu;
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
  if (mode < HIVE_STREAM_LAUNCHES || mode > HIVE_STREAM_PERSISTENT) return fail(HIVE_EINVAL, "unknown stream mode");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int dev = t->device, khz = 0, ncu = 0;
  HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  DevArena ar;
  unsigned long long* ts = nullptr;
  HIP_TRY(ar.alloc(&ts, nbatches + 1));
  HIP_TRY(hipMemsetAsync(ts, 0, (nbatches + 1) * 8, s));
  const auto t0 = std::chrono::steady_clock::now();
  if (mode == HIVE_STREAM_LAUNCHES || mode == HIVE_STREAM_GRAPH) {
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
    // persistent kernel: per-batch tile size, phase prefix, device queue state
    const unsigned bpc = t->stream_blocks_per_cu ? t->stream_blocks_per_cu : 4;
    const unsigned grid = static_cast<unsigned>(ncu) * bpc;
    std::vector<uint32_t> tile_ops(nbatches), phase_start(3 * nbatches + 1);
    uint64_t acc = 0;
    for (size_t b = 0; b < nbatches; ++b) {
      const uint64_t n = h_offsets[b + 1] - h_offsets[b];
      uint32_t tops = t->stream_tile;
      if (!tops) {  // about one tile per resident block, 256..4096 ops
        uint64_t want = (n + grid - 1) / grid;
        want = (want + kStreamBlock - 1) / kStreamBlock * kStreamBlock;
        tops = static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(want, kStreamBlock), 4096));
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
    unsigned long long* d_off = nullptr;
    uint32_t *d_tile = nullptr, *d_phase = nullptr;
    unsigned int *d_ticket = nullptr, *d_done = nullptr;
    HIP_TRY(ar.alloc(&d_off, nbatches + 1));
    HIP_TRY(ar.alloc(&d_tile, nbatches));
    HIP_TRY(ar.alloc(&d_phase, 3 * nbatches + 1));
    HIP_TRY(ar.alloc(&d_ticket, 2));
    HIP_TRY(ar.alloc(&d_done, 3 * nbatches));
    HIP_TRY(hipMemcpyAsync(d_off, h_offsets, (nbatches + 1) * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_tile, tile_ops.data(), nbatches * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_phase, phase_start.data(), (3 * nbatches + 1) * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(d_ticket, 0, 8, s));
    HIP_TRY(hipMemsetAsync(d_done, 0, 3 * nbatches * 4, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (setup_ms)
      *setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const uint32_t max_tile = *std::max_element(tile_ops.begin(), tile_ops.end());
    const uint32_t in_lds = 3 * nbatches + 1 <= kLdsPhases ? 1u : 0u;
    const size_t lds = (static_cast<size_t>(max_tile) + (in_lds ? 3 * nbatches + 1 : 0)) * 4;
    StreamArgs args{types, keys, vals, out_vals, status, d_off, d_tile, d_phase,
                    static_cast<uint32_t>(3 * nbatches), d_ticket, d_done, ts, d_ticket + 1, max_tile, in_lds};
    k_stamp<<<1, 64, 0, s>>>(ts);
    HIVE_POLICY_SWITCH(t->policy, (k_stream<P><<<grid, kStreamBlock, lds, s>>>(t->v, args)));
    HIP_CHECK_LAUNCH();
    unsigned int err = 0;
    HIP_TRY(hipMemcpyAsync(&err, d_ticket + 1, 4, hipMemcpyDeviceToHost, s));
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
  HIP_TRY(hipMemcpyAsync(dst.slots, src.slots, nb * kSlots * 8, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(dst.stash, src.stash, nb * kStashSlots * 8, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(dst.stash_count, src.stash_count, nb * 4, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(dst.meta, src.meta, meta_words(src.nb) * 4, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(dst.counters, src.counters, counter_words() * 8, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemsetAsync(dst.locks, 0, nb * 4, s));
  return HIVE_OK;
}

}  // namespace

extern "C" hive_status_t hive_autotune_stream(hive_table_t* t, const uint8_t* types, const uint32_t* keys,
                                              const uint32_t* vals, const uint64_t* h_offsets, size_t nbatches,
                                              hive_tune_t* out, void* stream) {
  hive_status_t r;
  if ((r = check_table(t)) != HIVE_OK || (r = set_device(t)) != HIVE_OK) return r;
  if (!out || !h_offsets || nbatches == 0) return fail(HIVE_EINVAL, "autotune needs a sample stream and an output");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t nops = h_offsets[nbatches];
  hive_table tmp = *t;  // same policy and device, scratch copy of the table state
  if ((r = alloc_view(tmp.v, t->v.nb)) != HIVE_OK) return r;
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
  auto run = [&](int mode, double* ms) -> hive_status_t {
    hive_status_t rr = copy_view(tmp.v, t->v, s);
    if (rr != HIVE_OK) return rr;
    double setup = 0.0;
    rr = hive_mixed_stream(&tmp, types, keys, vals, h_offsets, nbatches, ov, st, mode, done.data(), &setup, stream);
    if (rr != HIVE_OK) return rr;
    *ms = done[nbatches - 1] * 1e-6;
    return HIVE_OK;
  };
  hive_tune_t best{};
  best.launches_ms = best.graph_ms = best.persistent_ms = 1e300;
  uint32_t cands = 0;
  const uint32_t blocks[] = {128, 256, 512, 1024};
  for (uint32_t blk : blocks) {
    for (int op = 0; op < 3; ++op) tmp.launch[op] = hive_launch_t{blk, 1};
    double ms = 0;
    if ((r = run(HIVE_STREAM_LAUNCHES, &ms)) != HIVE_OK) return r;
    ++cands;
    if (ms < best.launches_ms) {
      best.launches_ms = ms;
      best.block_threads = blk;
    }
  }
  for (int op = 0; op < 3; ++op) tmp.launch[op] = hive_launch_t{best.block_threads, 1};
  if ((r = run(HIVE_STREAM_GRAPH, &best.graph_ms)) != HIVE_OK) return r;
  ++cands;
  const uint32_t tiles[] = {0, 256, 1024, 4096};
  const uint32_t bpcs[] = {2, 4, 8};
  for (uint32_t tile : tiles)
    for (uint32_t bpc : bpcs) {
      tmp.stream_tile = tile;
      tmp.stream_blocks_per_cu = bpc;
      double ms = 0;
      if ((r = run(HIVE_STREAM_PERSISTENT, &ms)) != HIVE_OK) return r;
      ++cands;
      if (ms < best.persistent_ms) {
        best.persistent_ms = ms;
        best.tile_ops = tile;
        best.blocks_per_cu = bpc;
      }
    }
  best.candidates = cands;
  best.best_mode = HIVE_STREAM_PERSISTENT;
  if (best.launches_ms < best.persistent_ms && best.launches_ms <= best.graph_ms) best.best_mode = HIVE_STREAM_LAUNCHES;
  else if (best.graph_ms < best.persistent_ms) best.best_mode = HIVE_STREAM_GRAPH;
  for (int op = 0; op < 3; ++op) t->launch[op] = hive_launch_t{best.block_threads, 1};
  t->stream_tile = best.tile_ops;
  t->stream_blocks_per_cu = best.blocks_per_cu;
  *out = best;
  return HIVE_OK;
}
