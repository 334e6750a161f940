// synchro_amd — the reference's command-line driver (app/main.cpp:10-78,
// app/args.hpp, app/io.hpp) on the device library: the same input and output
// files, so that scripts written for the reference's `synchro` run unchanged.
//
//   synchro_amd -f automata.txt -c config.json [-o results.txt [--continue]]
//               [-v | -q | -d] [-b suffix] [-j jobs] [--reference-word] [-h]
//
// Input: automata one after another, each "K N" then K*N transitions,
// delta(q, a) at position q*K + a (app/io.hpp:50-77).  Output: one line per
// automaton, "ID: NON SYNCHRO" or "ID: [lower, upper] ((name, ms), ...)"
// followed by " {w_0 w_1 ...}" when a reset word of length `upper` is known
// (app/io.hpp:95-143).  With --continue the non-empty lines already in the
// output file are counted and that many automata are skipped
// (app/main.cpp:26-46).  -b is accepted for compatibility (there is no JIT
// build here).  -j N (not in the reference) solves N automata at a time, each
// with its own device context and stream, and still writes the lines in input
// order.  The word field holds a shortest reset word whenever Exact ran; with
// --reference-word it holds what the reference writes instead: Eppstein's
// word when the plan asks for it (find_word), else nothing.  Exit codes follow the reference: 1 arguments or output file,
// 2 config, 3 input, 4 a failed run.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cctype>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <mutex>
#include <optional>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "synchrowords_amd.h"

namespace {

int g_level = 1;  // 0 quiet, 1 info, 2 verbose, 3 debug

void log(int level, const char* tag, const std::string& msg) {
  if (level > g_level && level != 0) return;
  const auto now = std::chrono::system_clock::now();
  const std::time_t t = std::chrono::system_clock::to_time_t(now);
  const long ms = long(std::chrono::duration_cast<std::chrono::milliseconds>(now.time_since_epoch()).count() % 1000);
  std::tm tm{};
  localtime_r(&t, &tm);
  std::fprintf(stderr, "[%02d:%02d:%02d.%03ld] [%s] %s\n", tm.tm_hour, tm.tm_min, tm.tm_sec, ms, tag, msg.c_str());
}
void info(const std::string& m) { log(1, "INFO", m); }
[[noreturn]] void fail(int code, const std::string& m) {
  log(0, "ERROR", m);
  std::exit(code);
}

struct Args {
  std::string input, config, build_suffix;
  std::optional<std::string> output;
  bool verbose = false, quiet = false, debug = false, cont = false, reference_word = false;
  int jobs = 1;
};

void usage() {
  std::cout << "Find synchronizing word of an automaton\n"
               "Usage:\n  synchro_amd [OPTION...]\n\n"
               "  -f, --file arg          Path to the input file\n"
               "  -c, --config arg        Path to the config file\n"
               "  -o, --output arg        Path to the output file\n"
               "  -b, --build-suffix arg  Suffix of the build folder (default: \"\")\n"
               "      --continue          Do not overwrite the output file and run algorithms only for remaining\n"
               "                          automata\n"
               "  -v, --verbose           Verbose output\n"
               "  -q, --quiet             Quiet output (only warnings and errors)\n"
               "  -d, --debug             Debug output (all messages and timers)\n"
               "  -j, --jobs arg          Automata solved at a time, one device context each (default: 1)\n"
               "      --reference-word    Write the word the reference writes (Eppstein's, with find_word)\n"
               "                          instead of a shortest one\n"
               "  -h, --help              Print usage\n";
}

Args parse(int argc, char** argv) {
  Args a;
  bool have_f = false, have_c = false;
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    auto value = [&](const char* name) -> std::string {
      const size_t eq = s.find('=');
      if (s.rfind("--", 0) == 0 && eq != std::string::npos) return s.substr(eq + 1);
      if (i + 1 >= argc) fail(1, std::string("Option '") + name + "' is missing an argument");
      return argv[++i];
    };
    auto is = [&](const char* shortf, const char* longf) {
      const std::string l = longf;
      return (shortf && s == shortf) || s == l || s.rfind(l + "=", 0) == 0;
    };
    if (is("-h", "--help")) {
      usage();
      std::exit(0);
    } else if (is("-f", "--file")) {
      a.input = value("file");
      have_f = true;
    } else if (is("-c", "--config")) {
      a.config = value("config");
      have_c = true;
    } else if (is("-o", "--output")) {
      a.output = value("output");
    } else if (is("-b", "--build-suffix")) {
      a.build_suffix = value("build-suffix");
    } else if (s == "--continue") {
      a.cont = true;
    } else if (s == "--reference-word") {
      a.reference_word = true;
    } else if (is("-v", "--verbose")) {
      a.verbose = true;
    } else if (is("-q", "--quiet")) {
      a.quiet = true;
    } else if (is("-d", "--debug")) {
      a.debug = true;
    } else if (is("-j", "--jobs")) {
      const std::string v = value("jobs");
      char* end = nullptr;
      const long j = std::strtol(v.c_str(), &end, 10);
      if (!end || *end || j < 1 || j > 64) fail(1, "Option 'jobs' needs an integer in [1, 64]");
      a.jobs = int(j);
    } else {
      fail(1, "Option '" + s + "' does not exist");
    }
  }
  if (!have_f) fail(1, "Option 'file' is required");
  if (!have_c) fail(1, "Option 'config' is required");
  if (int(a.verbose) + int(a.quiet) + int(a.debug) > 1)
    fail(1, "Only one of [--verbose, --quiet, --debug] can be enabled at one time");
  if (a.cont && !a.output) fail(1, "--continue can be enabled only if an output file is specified");
  return a;
}

struct Automaton {
  uint32_t n, k;
  std::vector<uint32_t> delta;
};

std::vector<Automaton> read_automata(const std::string& path) {
  std::ifstream in(path);
  if (!in) fail(3, "Cannot open the input file " + path);
  std::vector<Automaton> out;
  long long K, N;
  while (in >> K >> N) {
    if (N <= 0 || K <= 0) fail(3, "N and K must be greater than 0");
    Automaton a{uint32_t(N), uint32_t(K), std::vector<uint32_t>(size_t(N) * size_t(K))};
    for (size_t i = 0; i < a.delta.size(); ++i) {
      long long x;
      if (!(in >> x)) fail(3, "Expected " + std::to_string(a.delta.size()) + " integers, found " + std::to_string(i));
      if (x < 0 || x >= N)
        fail(3, "Expected integer in range [0, " + std::to_string(N - 1) + "], found " + std::to_string(x));
      a.delta[i] = uint32_t(x);
    }
    out.push_back(std::move(a));
  }
  info("Read " + std::to_string(out.size()) + " automata");
  return out;
}

std::string read_config(const std::string& path) {
  std::ifstream in(path);
  if (!in) fail(2, "Cannot open the config file " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  std::string s = ss.str();
  size_t i = 0;
  while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
  if (i == s.size() || s[i] != '{') fail(2, "The config file is not a JSON object: " + path);
  return s;
}

size_t count_nonempty_lines(const std::string& path) {
  std::ifstream in(path);
  size_t c = 0;
  std::string line;
  while (std::getline(in, line)) {
    for (char ch : line)
      if (!std::isspace(static_cast<unsigned char>(ch))) {
        ++c;
        break;
      }
  }
  return c;
}

}  // namespace

int main(int argc, char** argv) {
  Args args = parse(argc, argv);
  if (args.verbose) g_level = 2;
  if (args.debug) g_level = 3;
  if (args.quiet) g_level = 0;

  const std::vector<Automaton> auts = read_automata(args.input);
  const std::string plan = read_config(args.config);

  size_t skip = 0;
  std::ofstream out;
  if (args.output) {
    if (args.cont) skip = count_nonempty_lines(*args.output);
    out.open(*args.output, args.cont ? std::ios_base::app : std::ios_base::trunc);
    if (!out) fail(1, "Error while opening the output file");
  }
  if (skip) info("Skipping " + std::to_string(skip) + " automata");

  // One solved automaton: its result, the word, or the error of a failed run.
  struct Solved {
    bool done = false, ok = false;
    sw_result r{};
    std::vector<uint32_t> word;
    std::string error;
  };
  const size_t total = auts.size();
  std::vector<Solved> solved(total);
  std::mutex mu;
  std::condition_variable cv;
  std::atomic<size_t> next{skip};
  std::atomic<bool> stop{false};
  auto worker = [&]() {
    sw_ctx* ctx = nullptr;
    std::string ctx_error;
    if (sw_ctx_create(0, &ctx) != SW_OK) ctx_error = std::string("device: ") + sw_last_error();
    std::vector<uint32_t> word, rword;
    for (size_t index; !stop && (index = next.fetch_add(1)) < total;) {
      Solved out;
      if (ctx) {
        const Automaton& a = auts[index];
        const uint64_t cap = uint64_t(a.n) * a.n * a.n / 6 + uint64_t(a.n) * a.n + 8;  // longest bound a plan may give
        // Grown, never cleared: the library writes the words it reports.
        const size_t want = size_t(std::min<uint64_t>(cap, 1ull << 26));
        if (word.size() < want) word.resize(want);
        int rc;
        uint64_t rwl = 0;
        if (args.reference_word) {
          if (rword.size() < word.size()) rword.resize(word.size());
          rc = sw_pipeline_solve_result_word(ctx, a.n, a.k, a.delta.data(), plan.c_str(), &out.r, word.data(),
              word.size(), rword.data(), rword.size(), &rwl);
        } else {
          rc = sw_pipeline_solve(ctx, a.n, a.k, a.delta.data(), plan.c_str(), &out.r, word.data(), word.size());
        }
        if (rc == SW_OK) {
          out.ok = true;
          if (args.reference_word) {
            out.r.word_len = std::min<uint64_t>(rwl, rword.size());
            out.word.assign(rword.begin(), rword.begin() + out.r.word_len);
          } else {
            out.word.assign(word.begin(), word.begin() + std::min<uint64_t>(out.r.word_len, word.size()));
          }
        } else {
          out.error = "automaton " + std::to_string(index) + ": " + sw_last_error();
        }
      } else {
        out.error = ctx_error;
      }
      out.done = true;
      if (!out.ok) stop = true;
      {
        std::lock_guard<std::mutex> lk(mu);
        solved[index] = std::move(out);
      }
      cv.notify_all();
    }
    if (ctx) sw_ctx_destroy(ctx);
  };
  std::vector<std::thread> pool;
  const size_t jobs = std::min<size_t>(size_t(args.jobs), total > skip ? total - skip : 0);
  for (size_t t = 0; t < jobs; ++t) pool.emplace_back(worker);

  // Lines in input order, each as soon as it and every one before it is done.
  std::string failure;
  for (size_t index = skip; index < total; ++index) {
    Solved s;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return solved[index].done; });
      s = std::move(solved[index]);
    }
    if (!s.ok) {
      failure = s.error;
      break;
    }
    const sw_result& r = s.r;
    if (r.non_synchro) {
      info("NON SYNCHRO");
    } else {
      info("Minimum synchronizing word length: [" + std::to_string(r.lower) + ", " + std::to_string(r.upper) + "]");
    }
    if (!args.output) {
      if (r.word_len) info("Found synchronizing word of length " + std::to_string(r.word_len) +
                           " (use the -o flag to save it)");
      continue;
    }
    out << index << ": ";
    if (r.non_synchro) {
      out << "NON SYNCHRO\n";
      out.flush();
      continue;
    }
    out << "[" << r.lower << ", " << r.upper << "] (";
    for (int i = 0; i < r.n_algorithms; ++i) {
      if (i) out << ", ";
      out << "(" << r.names[i] << ", " << static_cast<long long>(r.ms[i]) << ")";
    }
    out << ")";
    if (r.word_len) {
      info("Saving synchronizing word of length " + std::to_string(r.word_len));
      out << " {";
      for (uint64_t i = 0; i < r.word_len; ++i) out << (i ? " " : "") << s.word[i];
      out << "}";
    }
    out << "\n";
    out.flush();
  }
  stop = true;
  for (auto& th : pool) th.join();
  if (!failure.empty()) fail(4, failure);
  return 0;
}
