// Test-infrastructure driver for the REFERENCE build under oracle/_ref/.
//
// This file is ours; it is compiled together with the reference's own,
// unmodified sources (jit/jitmain.cpp and synchrolib/**/*.cpp read in place
// from the reference tree) by oracle/build_ref.sh.  It plays the part of the
// reference's app/main.cpp + app/jit.hpp (which JIT-compile and dlopen the
// same sources): it reads automata in the reference input format
// (app/io.hpp:55-78, "K N" then N*K transitions), calls the library entry
// `run` (jit/jitmain.cpp:11-49) with a plan, and prints one result line per
// automaton in the reference output format (app/io.hpp:95-143) followed by
// machine-readable fields.
//
// Only tests, the smoke check and bench.py's cpu_baseline leg run the binary
// this file produces; nothing in the product path links or calls it.
//
// Usage: ref_N_K_T <input-file> [--bounds] [--plan Brute,Eppstein,Beam,Exact]
//   --bounds   additionally run the plan prefixes ending at Eppstein and at
//              Beam, to report each upper bound the pipeline produced.
//   --verbose  the library's VERBOSE log (per-step list sizes) on stderr.
#include <synchrolib/algorithm/algorithm.hpp>
#include <synchrolib/utils/logger.hpp>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

extern "C" void run(const std::string& aut_encoded,
    const std::vector<std::string>& algorithms, synchrolib::AlgoResult& result,
    synchrolib::Logger::LogLevel log_level);

namespace {

std::vector<std::string> split_plan(const std::string& s) {
  std::vector<std::string> out;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, ',')) {
    if (!item.empty()) out.push_back(item);
  }
  return out;
}

struct Outcome {
  bool non_synchro;
  unsigned long long lo, hi;
  double us;
  std::vector<std::pair<std::string, size_t>> times;
};

synchrolib::Logger::LogLevel g_level = synchrolib::Logger::LogLevel::ERROR;

Outcome run_plan(const std::string& enc, const std::vector<std::string>& plan) {
  synchrolib::AlgoResult result;
  auto t0 = std::chrono::steady_clock::now();
  run(enc, plan, result, g_level);
  auto t1 = std::chrono::steady_clock::now();
  Outcome o;
  o.non_synchro = result.non_synchro;
  o.lo = result.mlsw_lower_bound;
  o.hi = result.mlsw_upper_bound;
  o.us = std::chrono::duration<double, std::micro>(t1 - t0).count();
  for (const auto& r : result.algorithms_run) o.times.emplace_back(r.name, r.time);
  return o;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <input> [--bounds] [--plan a,b,c]\n", argv[0]);
    return 2;
  }
  bool bounds = false;
  std::vector<std::string> plan = {"Brute", "Eppstein", "Beam", "Exact"};
  for (int i = 2; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--bounds")) bounds = true;
    else if (!std::strcmp(argv[i], "--verbose")) g_level = synchrolib::Logger::LogLevel::VERBOSE;
    else if (!std::strcmp(argv[i], "--plan") && i + 1 < argc) plan = split_plan(argv[++i]);
  }
  synchrolib::Logger::set_log_stream(&std::cerr);

  std::ifstream in(argv[1]);
  if (!in) {
    std::fprintf(stderr, "cannot open %s\n", argv[1]);
    return 2;
  }
  unsigned k, n;
  size_t index = 0;
  while (in >> k >> n) {
    std::string enc;
    for (size_t i = 0; i < static_cast<size_t>(k) * n; ++i) {
      unsigned x;
      in >> x;
      if (!enc.empty()) enc.push_back(' ');
      enc += std::to_string(x);
    }
    if (n != AUT_N || k != AUT_K) {
      std::printf("%zu: SKIP (binary built for N=%u K=%u)\n", index++, (unsigned)AUT_N, (unsigned)AUT_K);
      continue;
    }
    unsigned long long epp = 0, beam = 0;
    if (bounds) {
      std::vector<std::string> p1, p2;
      for (const auto& a : plan) {
        p1.push_back(a);
        if (a == "Eppstein") break;
      }
      for (const auto& a : plan) {
        p2.push_back(a);
        if (a == "Beam") break;
      }
      auto o1 = run_plan(enc, p1);
      auto o2 = run_plan(enc, p2);
      epp = o1.hi;
      beam = o2.hi;
    }
    auto o = run_plan(enc, plan);
    std::printf("%zu: ", index);
    if (o.non_synchro) {
      std::printf("NON SYNCHRO");
    } else {
      std::printf("[%llu, %llu] (", o.lo, o.hi);
      for (size_t i = 0; i < o.times.size(); ++i) {
        std::printf("%s(%s, %zu)", i ? ", " : "", o.times[i].first.c_str(), o.times[i].second);
      }
      std::printf(")");
    }
    std::printf(" | non_synchro=%d lo=%llu hi=%llu", o.non_synchro ? 1 : 0, o.lo, o.hi);
    if (bounds) std::printf(" eppstein=%llu beam=%llu", epp, beam);
    std::printf(" total_us=%.1f\n", o.us);
    std::fflush(stdout);
    ++index;
  }
  return 0;
}
