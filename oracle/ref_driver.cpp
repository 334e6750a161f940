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
//   --reduce-out FILE  when a plan's Reduce reduced the automaton (the run
//              returned early with the reduction pending, jit/jitmain.cpp:
//              41-44), write the state the reference's app hands to its next
//              JIT build (app/main.cpp:57-66) to FILE and print
//              "reduce_pending=1 reduced_n=R" in the result line.
//   --reduce-in FILE   instead of reading automata, continue that state: the
//              reduced automaton with the reduction marked done, the plan's
//              remaining algorithms, the bounds and times so far (this
//              binary must be built for the reduced N).
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

Outcome run_plan(const std::string& enc, const std::vector<std::string>& plan,
    synchrolib::AlgoResult* state = nullptr) {
  synchrolib::AlgoResult local;
  synchrolib::AlgoResult& result = state ? *state : local;
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

// The pending reduction as text: "N K bfs_steps lower upper non_synchro",
// the reduced automaton's transitions, the number of algorithms run with
// their names and times, the number of list sets, then each set as 0/1.
void write_reduce(const char* path, const synchrolib::AlgoResult& r) {
  std::ofstream f(path);
  const auto& rd = *r.reduce;
  f << rd.aut.N << ' ' << rd.aut.K << ' ' << rd.bfs_steps << ' ' << r.mlsw_lower_bound << ' '
    << r.mlsw_upper_bound << ' ' << (r.non_synchro ? 1 : 0) << '\n';
  f << rd.aut.code() << '\n';
  f << r.algorithms_run.size() << '\n';
  for (const auto& a : r.algorithms_run) f << a.name << ' ' << a.time << '\n';
  f << rd.list_bfs.size() << '\n';
  for (const auto& v : rd.list_bfs) {
    for (size_t i = 0; i < v.size(); ++i) f << (v[i] ? '1' : '0');
    f << '\n';
  }
}

// Reads that state back; returns the reduced automaton's encoding.
std::string read_reduce(const char* path, synchrolib::AlgoResult& r, unsigned& n, unsigned& k) {
  std::ifstream f(path);
  unsigned long long steps = 0, lo = 0, hi = 0;
  int ns = 0;
  f >> n >> k >> steps >> lo >> hi >> ns;
  std::string enc;
  for (size_t i = 0; i < size_t(n) * k; ++i) {
    unsigned x;
    f >> x;
    if (!enc.empty()) enc.push_back(' ');
    enc += std::to_string(x);
  }
  size_t runs = 0;
  f >> runs;
  for (size_t i = 0; i < runs; ++i) {
    std::string name;
    size_t t;
    f >> name >> t;
    r.algorithms_run.emplace_back(name, t);
  }
  size_t sets = 0;
  f >> sets;
  synchrolib::ReduceData rd;
  rd.aut = synchrolib::VarAutomaton(n, k);
  std::stringstream ss(enc);
  for (size_t i = 0; i < size_t(n) * k; ++i) ss >> rd.aut.t[i];
  for (size_t i = 0; i < sets; ++i) {
    std::string bits;
    f >> bits;
    synchrolib::FastVector<bool> v(n);
    for (unsigned q = 0; q < n && q < bits.size(); ++q) v[q] = bits[q] == '1';
    rd.list_bfs.push_back(v);
  }
  rd.bfs_steps = unsigned(steps);
  rd.done = true;
  r.reduce = std::move(rd);
  r.non_synchro = ns != 0;
  r.mlsw_lower_bound = lo;
  r.mlsw_upper_bound = hi;
  return enc;
}

void print_outcome(size_t index, const Outcome& o) {
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
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <input> [--bounds] [--plan a,b,c]\n", argv[0]);
    return 2;
  }
  bool bounds = false;
  const char* reduce_out = nullptr;
  const char* reduce_in = nullptr;
  std::vector<std::string> plan = {"Brute", "Eppstein", "Beam", "Exact"};
  for (int i = 2; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--bounds")) bounds = true;
    else if (!std::strcmp(argv[i], "--verbose")) g_level = synchrolib::Logger::LogLevel::VERBOSE;
    else if (!std::strcmp(argv[i], "--plan") && i + 1 < argc) plan = split_plan(argv[++i]);
    else if (!std::strcmp(argv[i], "--reduce-out") && i + 1 < argc) reduce_out = argv[++i];
    else if (!std::strcmp(argv[i], "--reduce-in") && i + 1 < argc) reduce_in = argv[++i];
  }
  synchrolib::Logger::set_log_stream(&std::cerr);
  if (reduce_in) {
    synchrolib::AlgoResult state;
    unsigned rn = 0, rk = 0;
    const std::string enc = read_reduce(reduce_in, state, rn, rk);
    if (rn != AUT_N || rk != AUT_K) {
      std::printf("0: SKIP (binary built for N=%u K=%u)\n", (unsigned)AUT_N, (unsigned)AUT_K);
      return 0;
    }
    auto o = run_plan(enc, plan, &state);
    print_outcome(0, o);
    std::printf(" total_us=%.1f\n", o.us);
    return 0;
  }

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
    synchrolib::AlgoResult state;
    auto o = run_plan(enc, plan, &state);
    print_outcome(index, o);
    if (state.reduce && !state.reduce->done) {
      std::printf(" reduce_pending=1 reduced_n=%u", state.reduce->aut.N);
      if (reduce_out) write_reduce(reduce_out, state);
    }
    if (bounds) std::printf(" eppstein=%llu beam=%llu", epp, beam);
    std::printf(" total_us=%.1f\n", o.us);
    std::fflush(stdout);
    ++index;
  }
  return 0;
}
