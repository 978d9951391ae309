// tsan_driver.cpp -- TEST INFRASTRUCTURE ONLY: race test of the multithreaded
// CPU path (hive_cpu_mt.cpp) under ThreadSanitizer (`make -C oracle tsan`).
//
// Drives the concurrent protocol where it is hardest -- a small table, many
// threads, the same keys inserted by many threads at once (cuckoo kicks,
// stash pushes, update races), then mixed lookup/delete/insert batches --
// and checks after each step, with the sequential oracle's invariant checker,
// that the state the threads leave is a legal Hive table.  ThreadSanitizer
// reports any data race in the protocol; the exit code is non-zero on a race
// report (TSAN's own exit code) or on a failed check.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

extern "C" {
#include "hive_oracle.h"
}

extern "C" {
struct hcpu_table;
hcpu_table* hcpu_create(uint32_t nb, int policy);
void hcpu_destroy(hcpu_table* h);
double hcpu_insert(hcpu_table* h, const uint32_t* keys, const uint32_t* vals, size_t n, uint8_t* status, int threads);
double hcpu_lookup(hcpu_table* h, const uint32_t* keys, size_t n, uint32_t* vals, uint8_t* status, int threads);
double hcpu_mixed(hcpu_table* h, const uint8_t* types, const uint32_t* keys, const uint32_t* vals, size_t n,
                  uint32_t* out_vals, uint8_t* status, int threads);
void hcpu_export(const hcpu_table* h, uint64_t* slots, uint64_t* stash, uint32_t* stash_count);
}

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

// load the CPU path's state into the sequential oracle and run its checker
void check_state(hcpu_table* h, uint32_t nb, int policy, const char* where) {
  std::vector<uint64_t> slots(static_cast<size_t>(nb) * HO_SLOTS), stash(static_cast<size_t>(nb) * HO_STASH_SLOTS);
  std::vector<uint32_t> cnt(nb);
  hcpu_export(h, slots.data(), stash.data(), cnt.data());
  ho_table_t* o = ho_create(nb, policy);
  ho_import(o, slots.data(), stash.data(), cnt.data());
  char msg[256] = {0};
  const int r = ho_check_invariants(o, msg, sizeof msg);
  if (r != 0) std::fprintf(stderr, "%s: invariant %d: %s\n", where, r, msg);
  expect(r == 0, where);
  ho_destroy(o);
}

}  // namespace

int main() {
  const int threads = 8;
  std::mt19937 rng(7);
  for (int policy : {0, 1, 2}) {
    // 1. every key inserted by 8 threads at once into a table near LF 0.9
    const uint32_t nb = 64;
    hcpu_table* h = hcpu_create(nb, policy);
    std::vector<uint32_t> base(460);
    for (size_t i = 0; i < base.size(); ++i) base[i] = static_cast<uint32_t>(i * 2654435761u % 4000000000u) + 1u;
    std::vector<uint32_t> keys;
    for (int r = 0; r < 8; ++r) keys.insert(keys.end(), base.begin(), base.end());
    std::shuffle(keys.begin(), keys.end(), rng);
    std::vector<uint32_t> vals(keys.size());
    for (size_t i = 0; i < keys.size(); ++i) vals[i] = keys[i] ^ 0x5bd1e995u;
    std::vector<uint8_t> st(keys.size());
    hcpu_insert(h, keys.data(), vals.data(), keys.size(), st.data(), threads);
    check_state(h, nb, policy, "contended inserts");
    size_t placed = 0;
    for (uint8_t s : st) placed += (s == HO_R_INSERTED || s == HO_R_STASHED);
    std::vector<uint32_t> lv(base.size());
    std::vector<uint8_t> ls(base.size());
    hcpu_lookup(h, base.data(), base.size(), lv.data(), ls.data(), threads);
    size_t found = 0;
    for (size_t i = 0; i < base.size(); ++i) {
      if (ls[i] == HO_R_FOUND) {
        ++found;
        expect(lv[i] == (base[i] ^ 0x5bd1e995u), "value of a contended key");
      }
    }
    expect(found == placed, "one placement per stored key");
    // 2. mixed batches over a small key space (every op type races with the others)
    std::uniform_int_distribution<uint32_t> kd(1, 700), td(0, 2), vd;
    for (int b = 0; b < 20; ++b) {
      const size_t n = 1500;
      std::vector<uint8_t> types(n), status(n);
      std::vector<uint32_t> k(n), v(n), out(n);
      for (size_t i = 0; i < n; ++i) {
        types[i] = static_cast<uint8_t>(td(rng));
        k[i] = kd(rng);
        v[i] = vd(rng);
      }
      hcpu_mixed(h, types.data(), k.data(), v.data(), n, out.data(), status.data(), threads);
      check_state(h, nb, policy, "mixed batch");
    }
    hcpu_destroy(h);
  }
  if (failures) return 1;
  std::printf("tsan driver: all checks passed\n");
  return 0;
}
