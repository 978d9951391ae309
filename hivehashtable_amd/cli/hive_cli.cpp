// hive_cli.cpp -- the reference's three benchmark binaries, over libhive.so.
//
// Built three times by hivehashtable_amd/build.py (host code only, g++):
//   HIVE_CLI_MODE 1  hive_hash_table_benchmark            benchmark/hive_table_benchmark.cpp
//   HIVE_CLI_MODE 2  hive_hash_table_ycsb                 benchmark/ycsb_benchmark.cpp
//   HIVE_CLI_MODE 3  hive_hash_table_lookup_only_workload benchmark/hive_lookup_only_benchmark.cpp
// Same flags, defaults and printed throughput lines; the workloads restate
// HashTableBenchmark (src/hashTableBenchmark.cc:21-100, 156-272, 275-432,
// 434-570) and call the dispatch pair of include/hive.h.  One extension:
// --seed N makes a run reproducible (the reference seeds from
// std::random_device, include/hashTableBenchmark.h:57-59).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "hive.h"

#ifndef HIVE_CLI_MODE
#define HIVE_CLI_MODE 1
#endif

namespace {

struct Args {
  size_t table_log2;
  double load_factor;
  double insert_ratio = 0.0, lookup_ratio = 1.0, delete_ratio = 0.0;
  int iterations;
  int threads_per_block = 256;
  size_t n_operations = size_t(1) << 23;
  std::string distribution = "uniform";
  std::string layout;
  std::string policy = "Default2Hash";
  std::string ycsb = "C";
  uint64_t seed = 0;
  bool seeded = false;
};

void check(hive_status_t s, const char* what) {
  if (s != HIVE_OK) {  // the reference prints and exits on device errors (include/cuda_helper.cuh:8-16)
    std::cerr << what << ": " << hive_status_string(s) << " (" << hive_last_error() << ")\n";
    std::exit(EXIT_FAILURE);
  }
}

hive_operation_t op(int type, uint64_t key) {
  hive_operation_t o;
  o.type = type;
  o.pad = 0;
  o.key = key;
  return o;
}

// HashTableDataLayout (include/utils.h:23-28) from the CLI names
int layout_code(const std::string& s, bool allow_warpspeed) {
  if (s == "HybridSoA-AoS") return 0;
  if (s == "AaoS") return 1;
  if (s == "AaoS-LeadMetaData") return 2;
  if (allow_warpspeed && s == "WarpSpeed") return 3;
  return -1;
}

// generateTestKeys, 64-bit form (src/hashTableBenchmark.cc:21-57)
std::vector<uint64_t> test_keys(size_t n, const std::string& dist, std::mt19937_64& rng) {
  std::vector<uint64_t> keys(n);
  if (dist == "unique") {
    std::iota(keys.begin(), keys.end(), uint64_t(1));
    std::shuffle(keys.begin(), keys.end(), rng);
  } else if (dist == "uniform") {
    std::uniform_int_distribution<uint64_t> d(1, UINT64_MAX - 1);
    for (auto& k : keys) k = d(rng) % n + 1;  // keys in [1, n], repeats allowed
  } else if (dist == "clustered") {
    std::uniform_int_distribution<uint64_t> center(1, 1000), offset(0, 100);
    for (auto& k : keys) k = center(rng) * 10000 + offset(rng);
  } else if (dist == "skewed") {
    std::uniform_real_distribution<double> d(0.0, 1.0);
    for (auto& k : keys) {
      const double x = d(rng);
      k = static_cast<uint64_t>(std::pow(x, 3) * static_cast<double>(UINT64_MAX - 1)) + 1;
    }
  } else {
    throw std::invalid_argument("Unsupported distribution type");
  }
  return keys;
}

// Zipf over ranks 1..N with P(k) proportional to k^-s, sampled in O(1) with
// Vose's alias method (the distribution of ZipfAlias,
// include/zipf_distribution.h:82-194).
class Zipf {
 public:
  Zipf(size_t n, double s, std::mt19937_64& rng) : rng_(rng), prob_(n), alias_(n) {
    std::vector<double> p(n);
    double h = 0.0;
    for (size_t k = 0; k < n; ++k) h += (p[k] = std::pow(static_cast<double>(k + 1), -s));
    std::vector<uint32_t> small, large;
    for (size_t k = 0; k < n; ++k) {
      p[k] = p[k] / h * static_cast<double>(n);
      (p[k] < 1.0 ? small : large).push_back(static_cast<uint32_t>(k));
    }
    while (!small.empty() && !large.empty()) {
      const uint32_t a = small.back(), b = large.back();
      small.pop_back();
      prob_[a] = p[a];
      alias_[a] = b;
      p[b] = (p[b] + p[a]) - 1.0;
      if (p[b] < 1.0) {
        large.pop_back();
        small.push_back(b);
      }
    }
    for (uint32_t k : large) prob_[k] = 1.0;
    for (uint32_t k : small) prob_[k] = 1.0;
  }
  uint64_t sample() {  // in [1, N]
    std::uniform_int_distribution<size_t> col(0, prob_.size() - 1);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const size_t c = col(rng_);
    return (u(rng_) < prob_[c] ? c : alias_[c]) + 1;
  }

 private:
  std::mt19937_64& rng_;
  std::vector<double> prob_;
  std::vector<uint32_t> alias_;
};

double mops(size_t n, double ms) { return static_cast<double>(n) / (ms / 1000.0) / 1e6; }

// runBenchmarkHiveHashTableWithMixOps (src/hashTableBenchmark.cc:156-272)
int run_mixed(const Args& a, std::mt19937_64& rng) {
  const size_t table_size = size_t(1) << a.table_log2;
  const int layout = layout_code(a.layout, true);
  if (layout < 0) {
    std::cerr << "Unknown data layout type. Supported types: HybridSoA-AoS, AaoS, AaoS-LeadMetaData, WarpSpeed\n";
    return 1;
  }
  const size_t total = static_cast<size_t>(table_size * a.load_factor);
  const size_t n_ins = static_cast<size_t>(total * a.insert_ratio);
  const size_t n_look = static_cast<size_t>(total * a.lookup_ratio);
  const size_t n_del = static_cast<size_t>(total * a.delete_ratio);
  std::cout << "Table Size: " << table_size << ", Load Factor: " << a.load_factor << "\n"
            << "Distribution: " << a.distribution << "\nHash Policy: " << a.policy << "\n"
            << "Num Inserts: " << n_ins << ", Num Queries: " << n_look << ", Num Deletes: " << n_del << "\n";
  const size_t nkeys = n_ins ? n_ins : (n_look ? n_look : (n_del ? n_del : 1));
  const std::vector<uint64_t> keys = test_keys(nkeys, a.distribution, rng);
  std::vector<hive_operation_t> ops;
  ops.reserve(n_ins + n_look + n_del);
  if (n_ins)
    for (uint64_t k : keys) ops.push_back(op(HIVE_OP_INSERT, k));
  for (size_t i = 0; i < n_look; ++i) ops.push_back(op(HIVE_OP_LOOKUP, i < keys.size() ? keys[i] : 0));
  for (size_t i = 0; i < n_del; ++i) ops.push_back(op(HIVE_OP_DELETE, i < keys.size() ? keys[i] : 0));
  std::shuffle(ops.begin(), ops.end(), rng);
  const size_t tpb = static_cast<size_t>(a.threads_per_block);
  const size_t blocks = (ops.size() + tpb - 1) / tpb;
  std::cout << "Num Blocks: " << blocks << ", Threads per Block: " << tpb << "\n";
  std::vector<uint64_t> results(std::max(total, ops.size()), 0);
  double ms = 0.0;
  check(hive_hash_table_kernel_dispatch(ops.data(), ops.size(), table_size, tpb, blocks, a.iterations, &ms,
                                        results.data(), 1, layout, a.policy.c_str()),
        "hash_table_kernel_dispatch");
  std::cout << "Throughput: " << mops(ops.size(), ms) << std::endl;
  return 0;
}

// runYCSBBenchmarkWithHiveHashTable (src/hashTableBenchmark.cc:434-570)
int run_ycsb(const Args& a, std::mt19937_64& rng) {
  const size_t table_size = size_t(1) << a.table_log2;
  const int layout = layout_code(a.layout, false);
  if (layout < 0) {
    std::cerr << "Unknown data layout type. Supported types: HybridSoA-AoS, AaoS, AaoS-LeadMetaData\n";
    return 1;
  }
  double read = 0, update = 0, latest = 0;
  if (a.ycsb == "A") read = 0.5, update = 0.5;
  else if (a.ycsb == "B") read = 0.95, update = 0.05;
  else if (a.ycsb == "C") read = 1.0;
  else if (a.ycsb == "D") read = 0.95, latest = 0.05;
  else read = 0.5, update = 0.5;  // the reference keeps WORKLOAD_A for unknown names
  const size_t records = static_cast<size_t>(table_size * a.load_factor);
  const size_t n = a.n_operations;
  std::vector<hive_operation_t> prefill;
  prefill.reserve(records);
  for (size_t i = 1; i <= records; ++i) prefill.push_back(op(HIVE_OP_INSERT, i));
  const size_t n_read = std::min(static_cast<size_t>(n * read), records);
  const size_t n_upd = std::min(static_cast<size_t>(n * update), records);
  const size_t n_new = std::min(static_cast<size_t>(n * latest), records);
  std::cout << "Table Size: " << table_size << ", Load Factor: " << a.load_factor << "\n"
            << "YCSB Workload Ops - Reads: " << n_read << ", Updates: " << n_upd << ", RMWs: 0, Latest Inserts: "
            << n_new << "\n";
  Zipf zipf(std::max<size_t>(records, 2), 1.5, rng);
  std::vector<hive_operation_t> work;
  work.reserve(n);
  for (size_t i = 0; i < n_new; ++i) work.push_back(op(HIVE_OP_INSERT, records + i + 1));
  for (size_t i = 0; i < n_read; ++i)
    work.push_back(op(HIVE_OP_LOOKUP, (a.ycsb == "D" && n_new) ? records + (i % n_new) + 1 : zipf.sample()));
  for (size_t i = 0; i < n_upd; ++i) work.push_back(op(HIVE_OP_INSERT, zipf.sample()));
  // the reference passes n and reads past the generated ops when they are
  // fewer; here the remainder is NONE ops (no effect, result 0)
  while (work.size() < n) work.push_back(op(HIVE_OP_NONE, 0));
  std::shuffle(work.begin(), work.end(), rng);
  std::vector<uint64_t> results(n, 0);
  double ms = 0.0;
  check(hive_hash_table_kernel_dispatch_ycsb(prefill.data(), work.data(), records, n, table_size,
                                             static_cast<size_t>(a.threads_per_block), a.iterations, &ms,
                                             results.data(), 1, layout),
        "hash_table_kernel_dispatch_YCSB");
  std::cout << "Throughput: " << mops(n, ms) << std::endl;
  return 0;
}

// runAllLookupsBenchmark (src/hashTableBenchmark.cc:275-432): lookups of
// record_count keys with 100/75/50/25/0 % present
int run_lookup_only(const Args& a, std::mt19937_64& rng) {
  const size_t table_size = size_t(1) << a.table_log2;
  const int layout = layout_code(a.layout, false);
  if (layout < 0) {
    std::cerr << "Unknown data layout type. Supported types: HybridSoA-AoS, AaoS, AaoS-LeadMetaData\n";
    return 1;
  }
  const size_t records = static_cast<size_t>(table_size * a.load_factor);
  std::cout << "Table Size: " << table_size << ", Load Factor: " << a.load_factor << "\n";
  const std::vector<uint64_t> universe = test_keys(records, "unique", rng);
  std::vector<hive_operation_t> prefill;
  prefill.reserve(records);
  for (uint64_t k : universe) prefill.push_back(op(HIVE_OP_INSERT, k));
  std::shuffle(prefill.begin(), prefill.end(), rng);
  const int pct[] = {100, 75, 50, 25, 0};
  for (int p : pct) {
    std::vector<hive_operation_t> look;
    // loops `i < records * fraction` over a double bound: ceil of the product
    const size_t hit = static_cast<size_t>(std::ceil(static_cast<double>(records) * p / 100.0));
    const size_t miss = static_cast<size_t>(std::ceil(static_cast<double>(records) * (100 - p) / 100.0));
    for (size_t i = 0; i < hit; ++i) look.push_back(op(HIVE_OP_LOOKUP, universe[i]));
    for (size_t i = 0; i < miss; ++i) look.push_back(op(HIVE_OP_LOOKUP, universe[i] + records));  // absent keys
    if (p != 50) std::shuffle(look.begin(), look.end(), rng);  // the reference leaves its 50 % set unshuffled
    std::vector<uint64_t> results(look.size(), 0);
    double ms = 0.0;
    check(hive_hash_table_kernel_dispatch_ycsb(prefill.data(), look.data(), records, look.size(), table_size,
                                               static_cast<size_t>(a.threads_per_block), a.iterations, &ms,
                                               results.data(), 1, layout),
          "hash_table_kernel_dispatch_YCSB");
    size_t found = 0;
    for (uint64_t r : results) found += r != 0;
    std::cout << "Throughput (" << p << "% exist): " << mops(look.size(), ms) << "  (found " << found << " of "
              << look.size() << ")" << std::endl;
  }
  return 0;
}

const char* kName[] = {"", "hive_hash_table_benchmark", "hive_hash_table_ycsb", "hive_hash_table_lookup_only_workload"};

void help(int mode) {
  std::cout << "Usage: ./bin/" << kName[mode] << " [OPTIONS]\nOptions:\n"
            << "  --table_size <log2>          table size as a power of two\n"
            << "  --load_factor <factor>       load factor\n"
            << "  --num_iterations <count>     timed iterations\n"
            << "  --threads-per-block <count>  threads per block\n"
            << "  --data_layout <type>         HybridSoA-AoS, AaoS, AaoS-LeadMetaData"
            << (mode == 1 ? ", WarpSpeed" : "") << "\n";
  if (mode != 2)
    std::cout << "  --insert_ratio / --lookup_ratio / --delete_ratio <ratio>\n"
              << "  --distribution <type>        uniform, skewed, clustered, unique\n"
              << "  --hash_policy <type>         Default2Hash, TripleHash, MurmurCityHash, MurmurCityBitHash,"
                 " Lookup2Hash, Lookup3Hash\n";
  if (mode == 2)
    std::cout << "  --ycsb_workload_type <type>  A, B, C, D\n"
              << "  --num_operations_ycsb <n>    operations\n";
  if (mode == 3) std::cout << "  --n_operations <count>       operations\n";
  std::cout << "  --seed <n>                   fixed random seed (default: std::random_device)\n"
            << "  --help                       this message\n";
}

}  // namespace

int main(int argc, char** argv) {
  const int mode = HIVE_CLI_MODE;
  Args a;
  // defaults of each reference binary
  if (mode == 1) {  // benchmark/hive_table_benchmark.cpp:17-31
    a.table_log2 = 22, a.load_factor = 1.0, a.insert_ratio = 0.5, a.lookup_ratio = 0.4, a.delete_ratio = 0.1;
    a.iterations = 1, a.layout = "HybridSoA-AoS";
  } else if (mode == 2) {  // benchmark/ycsb_benchmark.cpp:18-27
    a.table_log2 = 25, a.load_factor = 0.5, a.iterations = 10, a.layout = "AaoS-LeadMetaData", a.ycsb = "C";
  } else {  // benchmark/hive_lookup_only_benchmark.cpp:18-32
    a.table_log2 = 24, a.load_factor = 0.7, a.iterations = 10, a.layout = "HybridSoA-AoS";
  }
  try {
    for (int i = 1; i < argc; ++i) {
      const std::string arg = argv[i];
      auto next = [&]() -> std::string {
        if (i + 1 >= argc) throw std::invalid_argument("missing value for " + arg);
        return argv[++i];
      };
      if (arg == "--help") {
        help(mode);
        return 0;
      } else if (arg == "--table_size") a.table_log2 = std::stoul(next());
      else if (arg == "--load_factor") a.load_factor = std::stod(next());
      else if (arg == "--num_iterations") a.iterations = std::stoi(next());
      else if (arg == "--threads-per-block") a.threads_per_block = std::stoi(next());
      else if (arg == "--data_layout") a.layout = next();
      else if (arg == "--seed") a.seed = std::stoull(next()), a.seeded = true;
      else if (mode != 2 && arg == "--insert_ratio") a.insert_ratio = std::stod(next());
      else if (mode != 2 && arg == "--lookup_ratio") a.lookup_ratio = std::stod(next());
      else if (mode != 2 && arg == "--delete_ratio") a.delete_ratio = std::stod(next());
      else if (mode != 2 && arg == "--distribution") a.distribution = next();
      else if (mode != 2 && arg == "--hash_policy") a.policy = next();
      else if (mode == 2 && arg == "--ycsb_workload_type") a.ycsb = next();
      else if (mode == 2 && arg == "--num_operations_ycsb") a.n_operations = std::stoul(next());
      else if (mode == 3 && arg == "--n_operations") a.n_operations = std::stoul(next());
      else {
        std::cerr << "Error: Unknown argument '" << arg << "'\n";
        help(mode);
        return 1;
      }
    }
    if (a.table_log2 >= 40 || a.iterations < 1 || a.threads_per_block < 1 || !(a.load_factor > 0.0))
      throw std::invalid_argument("table_size, num_iterations, threads-per-block and load_factor must be positive");
    std::mt19937_64 rng(a.seeded ? a.seed : std::random_device{}());
    std::cout << "Config: Table Size=" << (size_t(1) << a.table_log2) << ", Load Factor=" << std::fixed
              << std::setprecision(2) << a.load_factor * 100 << "%, Iterations=" << a.iterations << "\n"
              << std::defaultfloat;
    if (mode == 1) return run_mixed(a, rng);
    if (mode == 2) return run_ycsb(a, rng);
    return run_lookup_only(a, rng);
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << "\n";
    return 1;
  }
}
