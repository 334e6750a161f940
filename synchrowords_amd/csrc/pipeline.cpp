// The plan pipeline around the device exact search: the reference's run()
// (jit/jitmain.cpp:11-49) over a config-JSON plan (configs/exact.json).
//
// Brute, Eppstein and Beam are the bound suppliers that sit before Exact in
// the reference's plans, Reduce may shrink the automaton Exact searches.
// All five run on the device: Brute as a power-set BFS (brute.cpp),
// Eppstein's pair tree and greedy word (eppstein.cpp), the Beam in the
// reference's order (mlp.cpp, sw_beam_device), Reduce's short forward search
// (reduce.cpp) and Exact (exact.cpp).  This file parses the plan, evaluates
// its C++ expressions, keeps [lower, upper] and the word, and does the
// Beam's presort and dynamic sizing on the host.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include "sw_host.hpp"
#include "sw_kernels.hpp"
#include "sw_reduce.hpp"

namespace sw {
namespace {

// ------------------------------------------------------------- tiny JSON
struct Json {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;
  const Json* get(const std::string& k) const {
    for (auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct JsonParser {
  const char* p;
  const char* end;
  void ws() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r')) ++p;
  }
  [[noreturn]] void fail(const char* what) { throw Error{SW_ERR_INVALID, std::string("plan JSON: ") + what}; }
  Json parse() {
    ws();
    if (p >= end) fail("unexpected end");
    Json j;
    if (*p == '{') {
      j.kind = Json::OBJ;
      ++p;
      ws();
      if (*p == '}') {
        ++p;
        return j;
      }
      while (true) {
        ws();
        Json k = parse();
        if (k.kind != Json::STR) fail("object key");
        ws();
        if (*p != ':') fail("expected ':'");
        ++p;
        j.obj.emplace_back(k.str, parse());
        ws();
        if (*p == ',') {
          ++p;
          continue;
        }
        if (*p == '}') {
          ++p;
          break;
        }
        fail("expected ',' or '}'");
      }
    } else if (*p == '[') {
      j.kind = Json::ARR;
      ++p;
      ws();
      if (*p == ']') {
        ++p;
        return j;
      }
      while (true) {
        j.arr.push_back(parse());
        ws();
        if (*p == ',') {
          ++p;
          continue;
        }
        if (*p == ']') {
          ++p;
          break;
        }
        fail("expected ',' or ']'");
      }
    } else if (*p == '"') {
      j.kind = Json::STR;
      ++p;
      while (p < end && *p != '"') {
        if (*p == '\\' && p + 1 < end) ++p;
        j.str.push_back(*p++);
      }
      if (p >= end) fail("unterminated string");
      ++p;
    } else if (!strncmp(p, "true", 4)) {
      j.kind = Json::BOOL;
      j.b = true;
      p += 4;
    } else if (!strncmp(p, "false", 5)) {
      j.kind = Json::BOOL;
      p += 5;
    } else if (!strncmp(p, "null", 4)) {
      p += 4;
    } else {
      j.kind = Json::NUM;
      char* e = nullptr;
      j.num = strtod(p, &e);
      if (e == p) fail("bad value");
      p = e;
    }
    return j;
  }
};

// ------------------------------------------------- expression evaluation
// Plan values may be C++ expressions over AUT_N and AUT_K
// (docs/config.md:9-12), e.g. "AUT_N * std::log2(AUT_N)".  Integer operands
// keep integer arithmetic (C++ division), anything else is long double.
struct Val {
  bool is_int;
  long long i;
  long double f;
  long double as_f() const { return is_int ? (long double)i : f; }
};

struct Expr {
  const char* p;
  long long N, K;
  [[noreturn]] void fail() { throw Error{SW_ERR_INVALID, std::string("plan expression: cannot parse near '") + p + "'"}; }
  void ws() {
    while (*p == ' ') ++p;
  }
  static Val I(long long v) { return {true, v, 0}; }
  static Val F(long double v) { return {false, 0, v}; }
  Val prim() {
    ws();
    if (*p == '(') {
      ++p;
      Val v = sum();
      ws();
      if (*p != ')') fail();
      ++p;
      return v;
    }
    if (*p == '-') {
      ++p;
      Val v = prim();
      return v.is_int ? I(-v.i) : F(-v.f);
    }
    if ((*p >= '0' && *p <= '9') || *p == '.') {
      const char* s = p;
      bool flt = false;
      while ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E') {
        if (*p == '.' || *p == 'e' || *p == 'E') flt = true;
        ++p;
      }
      std::string tok(s, p);
      while (*p == 'u' || *p == 'U' || *p == 'l' || *p == 'L' || *p == 'f' || *p == 'F') ++p;
      return flt ? F(strtold(tok.c_str(), nullptr)) : I(strtoll(tok.c_str(), nullptr, 10));
    }
    const char* s = p;
    while ((*p >= 'a' && *p <= 'z') || (*p >= 'A' && *p <= 'Z') || *p == '_' || *p == ':' || (*p >= '0' && *p <= '9'))
      ++p;
    std::string id(s, p);
    if (id == "AUT_N") return I(N);
    if (id == "AUT_K") return I(K);
    if (id == "true") return I(1);
    if (id == "false") return I(0);
    std::string fn = id.rfind("std::", 0) == 0 ? id.substr(5) : id;
    ws();
    if (*p != '(') fail();
    ++p;
    Val a = sum();
    ws();
    Val b = I(0);
    if (*p == ',') {
      ++p;
      b = sum();
      ws();
    }
    if (*p != ')') fail();
    ++p;
    long double x = a.as_f();
    if (fn == "log2") return F(std::log2(x));
    if (fn == "log") return F(std::log(x));
    if (fn == "sqrt") return F(std::sqrt(x));
    if (fn == "pow") return F(std::pow(x, b.as_f()));
    if (fn == "ceil") return F(std::ceil(x));
    if (fn == "floor") return F(std::floor(x));
    fail();
  }
  Val prod() {
    Val v = prim();
    while (true) {
      ws();
      char op = *p;
      if (op != '*' && op != '/' && op != '%') return v;
      ++p;
      Val r = prim();
      if (v.is_int && r.is_int) {
        if (op == '*') v.i *= r.i;
        else if (r.i == 0) fail();
        else if (op == '/') v.i /= r.i;
        else v.i %= r.i;
      } else {
        long double a = v.as_f(), b = r.as_f();
        v = F(op == '*' ? a * b : a / b);
      }
    }
  }
  Val sum() {
    Val v = prod();
    while (true) {
      ws();
      char op = *p;
      if (op != '+' && op != '-') return v;
      ++p;
      Val r = prod();
      if (v.is_int && r.is_int) v.i = op == '+' ? v.i + r.i : v.i - r.i;
      else v = F(op == '+' ? v.as_f() + r.as_f() : v.as_f() - r.as_f());
    }
  }
};

long double eval_num(const Json* j, const char* dflt, uint32_t n, uint32_t k) {
  if (j && j->kind == Json::NUM) return j->num;
  if (j && j->kind == Json::BOOL) return j->b ? 1 : 0;
  std::string s = (j && j->kind == Json::STR) ? j->str : std::string(dflt);
  Expr e{s.c_str(), n, k};
  Val v = e.sum();
  e.ws();
  if (*e.p) e.fail();
  return v.as_f();
}

bool eval_bool(const Json* j, bool dflt, uint32_t n, uint32_t k) {
  if (!j) return dflt;
  if (j->kind == Json::BOOL) return j->b;
  return eval_num(j, "0", n, k) != 0;
}

// ------------------------------------------------------------------ Beam
// Inverse BFS from the sink singletons keeping the `beam` largest preimages
// (ties by per-word numeric order); the first depth reaching all N states
// bounds the length (algorithm/beam/beam.hpp:98-188).  Each run is the device
// beam (sw_beam_device in its reference order), which also returns the word.
struct BeamSettings {
  bool dynamic = true;
  long double beam_size = 0, min_beam = 0, max_beam = 0, ratio = 0.01;
  bool presort_indeg = true;
  long long max_iter = -1;
};

struct BeamResult {
  uint64_t bound;
  std::vector<uint32_t> word;
};

struct DeviceBeam {
  sw_ctx* ctx;
  uint32_t n, k;
  std::vector<uint32_t> d;  // relabelled automaton
  BeamResult run(uint64_t width, uint64_t max_len, const BeamStop* stop = nullptr) const {
    BeamResult r{max_len, {}};
    if (max_len <= 1 || width == 0) return r;
    std::vector<uint32_t> w(std::min<uint64_t>(max_len, 1u << 22));
    uint64_t b = max_len, wl = 0;
    const uint32_t wd = uint32_t(std::min<uint64_t>(width, 0xffffffffu));
    if (beam_device(ctx, nullptr, n, k, d.data(), wd, max_len, &b, w.data(), w.size(), &wl, stop) != SW_OK)
      throw Error{SW_ERR_HIP, std::string("device beam: ") + sw_last_error()};
    r.bound = b;
    if (b < max_len && wl == b) r.word.assign(w.begin(), w.begin() + wl);
    return r;
  }
};

// A host thread with a device context of its own that runs one task at a
// time for a context (the Beam's first pass beside Eppstein), made on first
// use and ended with the context.  Both sides spin for a short while before
// they sleep: a pass can last a tenth of a millisecond, about what waking a
// sleeping thread can take.
class Helper {
 public:
  explicit Helper(int device) {
    if (sw_ctx_create(device, &ctx_) != SW_OK) throw Error{SW_ERR_HIP, sw_last_error()};
    mark_helper_context(ctx_);
    try {
      start_thread(device);
    } catch (...) {
      sw_ctx_destroy(ctx_);
      throw;
    }
  }
  void start_thread(int device) {
    th_ = std::thread([this, device] {
      (void)hipSetDevice(device);
      uint64_t seen = 0;
      for (;;) {
        if (!spin([&] { return posted_.load(std::memory_order_acquire) != seen || quit_.load(); })) {
          std::unique_lock<std::mutex> lk(mu_);
          cv_.wait(lk, [&] { return posted_.load(std::memory_order_acquire) != seen || quit_.load(); });
        }
        if (quit_.load()) return;
        std::function<void()> f;
        {
          std::lock_guard<std::mutex> lk(mu_);
          seen = posted_.load(std::memory_order_acquire);
          f = std::move(task_);
          task_ = nullptr;
        }
        std::exception_ptr e;
        try {
          f();
        } catch (...) {
          e = std::current_exception();
        }
        {
          std::lock_guard<std::mutex> lk(mu_);
          err_ = e;
          done_.store(seen, std::memory_order_release);
        }
        cv_.notify_all();
      }
    });
  }
  ~Helper() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      quit_.store(true);
    }
    cv_.notify_all();
    th_.join();
    sw_ctx_destroy(ctx_);
  }
  sw_ctx* ctx() const { return ctx_; }
  void start(std::function<void()> f) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      task_ = std::move(f);
      posted_.fetch_add(1, std::memory_order_release);
    }
    cv_.notify_all();
  }
  // Until the task has returned; rethrows what it threw.
  void wait() {
    const uint64_t want = posted_.load(std::memory_order_acquire);
    auto done = [&] { return done_.load(std::memory_order_acquire) == want; };
    if (!spin(done)) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, done);
    }
    std::lock_guard<std::mutex> lk(mu_);
    if (err_) {
      std::exception_ptr e = err_;
      err_ = nullptr;
      std::rethrow_exception(e);
    }
  }

 private:
  // Polls `ready` for up to two milliseconds.
  template <typename F>
  static bool spin(F ready) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0;; ++i) {
      if (ready()) return true;
      if ((i & 255) == 255 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) return false;
      __builtin_ia32_pause();
    }
  }
  sw_ctx* ctx_ = nullptr;
  std::thread th_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::function<void()> task_;
  std::atomic<uint64_t> posted_{0}, done_{0};
  std::atomic<bool> quit_{false};
  std::exception_ptr err_;
};

// The stop cell's new value, one 64-bit store the beam reads whole.
void set_stop(sw_ctx* ctx, uint64_t value) {
  __atomic_store_n(ctx->stop_cell, (unsigned long long)value, __ATOMIC_RELEASE);
}

// The context's helper, made on first use; NULL when one cannot be made
// (then the stages run in order, and no helper is tried again).
Helper* helper_of(sw_ctx* ctx) {
  if (!ctx->helper && !ctx->helper_failed) {
    try {
      ctx->helper = new Helper(ctx->device);
      ctx->helper_free = [](void* h) { delete static_cast<Helper*>(h); };
    } catch (const std::exception&) {
      ctx->helper_failed = true;
    }
  }
  return static_cast<Helper*>(ctx->helper);
}

// The state relabelling of the Beam's presort "indeg": states by ascending
// in-degree, ties by index (beam.hpp:351-369).  The order of equal-size sets
// depends on the labels.
DeviceBeam relabelled(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* d0, const BeamSettings& bs) {
  std::vector<uint32_t> order(n);
  if (bs.presort_indeg) {
    std::vector<std::pair<double, int>> indeg(n);
    for (uint32_t q = 0; q < n; ++q) indeg[q] = {0, int(q)};
    for (uint32_t q = 0; q < n; ++q)
      for (uint32_t a = 0; a < k; ++a) indeg[d0[size_t(q) * k + a]].first++;
    std::sort(indeg.begin(), indeg.end());
    for (uint32_t i = 0; i < n; ++i) order[indeg[i].second] = i;
  } else {
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
  }
  DeviceBeam br{ctx, n, k, std::vector<uint32_t>(size_t(n) * k, 0)};
  for (uint32_t q = 0; q < n; ++q)
    for (uint32_t a = 0; a < k; ++a) br.d[size_t(order[q]) * k + a] = order[d0[size_t(q) * k + a]];
  return br;
}

// The Beam's depth cap under the upper bound `upper` (max_iter, when set,
// caps the depths too).
uint64_t beam_maxlen(uint64_t upper, const BeamSettings& bs) {
  uint64_t maxlen = upper;
  if (bs.max_iter >= 0 && uint64_t(bs.max_iter) + 1 < maxlen) maxlen = uint64_t(bs.max_iter) + 1;
  return maxlen;
}

// Width of the Beam's first (or only) pass.
uint64_t beam_first_width(const BeamSettings& bs) {
  return bs.dynamic ? uint64_t(static_cast<unsigned>(bs.min_beam)) : uint64_t(bs.beam_size);
}

// `first`: the first pass already run (under this upper), else NULL.
BeamResult beam(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* d0, uint64_t upper, const BeamSettings& bs,
    const BeamResult* first) {
  const DeviceBeam br = relabelled(ctx, n, k, d0, bs);
  const uint64_t maxlen = beam_maxlen(upper, bs);
  // A pass under the bound `cap` that reached no depth below its depth cap
  // (max_iter may end it first) returns `cap` itself, with no word
  // (beam.hpp:125-131, 186).
  auto settle = [&](BeamResult o, uint64_t cap) {
    if (o.bound >= beam_maxlen(cap, bs)) o = BeamResult{cap, {}};
    return o;
  };
  BeamResult res;
  if (!bs.dynamic) {
    auto o = settle(first ? *first : br.run(uint64_t(bs.beam_size), maxlen), upper);
    res.bound = std::min(o.bound, upper);
    res.word = o.word;
    return res;
  }
  // Dynamic sizing (beam.hpp:66-89): a cheap run, then a wider one sized from
  // the estimated cost of the exact search.
  using cost_t = long double;
  const unsigned min_beam = static_cast<unsigned>(bs.min_beam);
  const unsigned max_beam = static_cast<unsigned>(bs.max_beam);
  auto o1 = settle(first ? *first : br.run(min_beam, maxlen), upper);
  unsigned bound = static_cast<unsigned>(o1.bound);
  res.bound = bound;
  res.word = o1.word;
  cost_t exact_cost = 2 * std::pow(static_cast<cost_t>(k), (bound + 1) / 2);
  cost_t proposed = std::round(exact_cost * bs.ratio / bound);
  unsigned size = static_cast<unsigned>(std::clamp<cost_t>(proposed, min_beam, max_beam));
  if (size > min_beam) {
    auto o2 = settle(br.run(size, beam_maxlen(bound, bs)), bound);
    if (o2.bound < bound) {
      res.bound = o2.bound;
      res.word = o2.word;
    }
  }
  return res;
}

const char* kDefaultPlan = R"PLAN({
  "upper_bound": "1ULL * (AUT_N - 1) * (AUT_N - 1) + 1",
  "threads": 1,
  "gpu": true,
  "algorithms": [
    {"name": "Brute", "config": {"max_n": "15"}},
    {"name": "Eppstein", "config": {"transition_tables": false, "find_word": false}},
    {"name": "Beam", "config": {"presort": "indeg", "dynamic": true, "min_beam_size": "AUT_N * std::log2(AUT_N)",
      "max_beam_size": "AUT_N * AUT_N * std::log2(AUT_N)", "beam_exact_ratio": 0.01}},
    {"name": "Exact", "config": {"dfs_min_list_size": "8192", "bfs_small_list_size": "AUT_N * AUT_K",
      "dfs": true, "dfs_shortcut": true, "max_memory_mb": "12 * 1024"}}
  ]
})PLAN";

}  // namespace

BeamSettings beam_settings(const Json* cfg, uint32_t n, uint32_t k) {
  BeamSettings bs;
  bs.beam_size = eval_num(cfg->get("beam_size"), "std::log2(AUT_N)", n, k);
  bs.dynamic = eval_bool(cfg->get("dynamic"), false, n, k);
  bs.min_beam = eval_num(cfg->get("min_beam_size"), "AUT_N * std::log2(AUT_N)", n, k);
  bs.max_beam = eval_num(cfg->get("max_beam_size"), "AUT_N * AUT_N * std::log2(AUT_N)", n, k);
  bs.ratio = eval_num(cfg->get("beam_exact_ratio"), "0.01", n, k);
  bs.max_iter = (long long)eval_num(cfg->get("max_iter"), "-1", n, k);
  const Json* ps = cfg->get("presort");
  std::string presort = ps && ps->kind == Json::STR ? ps->str : "none";
  SW_REQUIRE(presort == "none" || presort == "indeg", "presort value not in [none, indeg]");
  bs.presort_indeg = presort == "indeg";
  return bs;
}

// The reference's run() over a plan (jit/jitmain.cpp:11-49): bounds start at
// [0, upper_bound] and every algorithm may only tighten them; an algorithm
// whose bounds already meet returns at once (each plugin's early exit).
void pipeline(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* d, const char* plan_json, sw_result* out,
    std::vector<uint32_t>& word, std::vector<uint32_t>* result_word) {
  const std::string text = plan_json ? plan_json : kDefaultPlan;
  JsonParser jp{text.data(), text.data() + text.size()};
  Json plan = jp.parse();
  SW_REQUIRE(plan.kind == Json::OBJ, "plan must be a JSON object");
  const Json* algos = plan.get("algorithms");
  SW_REQUIRE(algos && algos->kind == Json::ARR, "plan needs an \"algorithms\" list");
  SW_REQUIRE(algos->arr.size() <= 8, "at most 8 algorithms in a plan");
  std::memset(out, 0, sizeof(*out));
  uint64_t lower = 0;
  long double ub = eval_num(plan.get("upper_bound"), "1ULL * AUT_N * AUT_N * AUT_N / 6", n, k);
  uint64_t upper = ub >= 1.8e19L ? UINT64_MAX : uint64_t(ub);
  bool non_synchro = false;
  word.clear();
  int idx = 0;
  // After a Reduce that shrank the automaton, only Exact acts, on the
  // reduced automaton from the Reduce's list (the reference's other
  // algorithms return at once then: brute.hpp:39, eppstein.hpp:36,
  // beam.hpp:48, reduce.hpp:35).
  ReduceState red;
  // The Beam's first pass, when it ran beside Eppstein (the stage before it;
  // both are latency-bound chains of small launches, or one workgroup each up
  // to 64 states).  It runs on the helper's context under the bound Eppstein
  // started from, and ends early at the bound Eppstein found (a stop cell the
  // beam reads); a depth at or past that bound counts as no improvement, so
  // the result is the pass run after Eppstein under its bound.
  struct {
    bool ran = false;
    BeamResult r;
  } first;
  std::chrono::steady_clock::duration carry{0};  // time of the next stage spent in this one
  for (size_t ai = 0; ai < algos->arr.size(); ++ai) {
    const Json& a = algos->arr[ai];
    const Json* nm = a.get("name");
    SW_REQUIRE(nm && nm->kind == Json::STR, "algorithm without a name");
    const Json empty;
    const Json* cfg = a.get("config") ? a.get("config") : &empty;
    const std::string name = nm->str;
    auto t0 = std::chrono::steady_clock::now() - carry;
    carry = std::chrono::steady_clock::duration{0};
    const bool open = !non_synchro && lower != upper;
    if (red.reduced && name != "Exact") {
      // Not run on the reduced automaton.
    } else if (name == "Brute") {
      const long long max_n = (long long)eval_num(cfg->get("max_n"), "20", n, k);
      SW_REQUIRE(max_n <= 32, "max_n can not be greater than 32");
      if (open && n <= uint64_t(max_n)) {
        int64_t r = 0;
        const uint64_t off = 0;
        brute_batch(ctx, 1, &n, &k, &off, d, &r);
        if (r < 0) non_synchro = true;
        else lower = upper = uint64_t(r);
      }
    } else if (name == "Eppstein") {
      const bool find_word = eval_bool(cfg->get("find_word"), false, n, k);
      const bool tables = eval_bool(cfg->get("transition_tables"), false, n, k);
      if (open) {
        // Eppstein's greedy bound on the device (k_eppstein.hip).
        const uint64_t off = 0;
        int32_t ns = 0, imp = 0;
        uint64_t bound = upper, wl = 0;
        const uint64_t cap = std::min<uint64_t>(upper, 1u << 22);
        std::vector<uint32_t> w(cap ? cap : 1);
        // The next stage a Beam that can run beside this one?
        const Json* next = ai + 1 < algos->arr.size() ? &algos->arr[ai + 1] : nullptr;
        const Json* next_name = next ? next->get("name") : nullptr;
        // Only for a context alone in its process: with several busy, the
        // device's few hardware queues are shared (two passes on one queue
        // run one after the other) and the plans already overlap each other.
        const bool beside = n >= 2 && next_name &&
            next_name->kind == Json::STR && next_name->str == "Beam" && user_contexts() == 1;
        BeamSettings bs;
        DeviceBeam br{nullptr, n, k, {}};
        uint64_t width = 0, pass_len = 0;
        BeamResult pass;
        Helper* helper = nullptr;
        uint32_t gen = 0;
        if (beside) {
          const Json empty_cfg;
          bs = beam_settings(next->get("config") ? next->get("config") : &empty_cfg, n, k);
          width = beam_first_width(bs);
          pass_len = beam_maxlen(upper, bs);
          if (width && pass_len > 1) helper = helper_of(ctx);
          if (helper) {
            br = relabelled(helper->ctx(), n, k, d, bs);
            if (!ctx->stop_cell) {
              // Pinned host memory: this thread writes it, the beam's kernel
              // (or its host loop) reads it.
              SW_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->stop_cell), 64, hipHostMallocDefault));
              set_stop(ctx, 0);
            }
            gen = ++ctx->stop_gen;
            if (gen == 0) gen = ++ctx->stop_gen;  // 0 is the cell's first value
            const BeamStop stop{ctx->stop_cell, gen};
            helper->start([&br, &pass, width, pass_len, stop] { pass = br.run(width, pass_len, &stop); });
          }
        }
        // Whatever ends this stage, the pass ends (at its next check) and is
        // waited for before the buffers it uses go.
        struct PassGuard {
          Helper* h;
          sw_ctx* ctx;
          uint32_t gen;
          ~PassGuard() {
            if (!h) return;
            set_stop(ctx, uint64_t(gen) << 32);
            try {
              h->wait();
            } catch (...) {
            }
          }
        } guard{helper, ctx, gen};
        eppstein_batch(ctx, 1, &n, &k, &off, d, &upper, cap, &ns, &imp, &bound, &wl, w.data());
        if (helper) {
          // The depth the pass would have stopped at, run after this stage.
          const auto t_split = std::chrono::steady_clock::now();
          const uint64_t after = ns ? 0 : beam_maxlen(imp ? bound : upper, bs);
          set_stop(ctx, (uint64_t(gen) << 32) | std::min<uint64_t>(after, 0xffffffffu));
          guard.h = nullptr;
          helper->wait();
          first.ran = !ns;
          first.r = pass.bound < after ? pass : BeamResult{after, {}};
          // The wait is the Beam's time.
          carry = std::chrono::steady_clock::now() - t_split;
          t0 += carry;
        }
        if (ns) {
          non_synchro = true;
        } else if (imp) {
          upper = bound;
          // The word is kept whatever find_word says: when Exact proves this
          // bound optimal, it is the shortest word.
          word.assign(w.begin(), w.begin() + std::min<uint64_t>(wl, cap));
          // The reference's result.word: Eppstein's, when the plan asks for it
          // (eppstein.hpp:70-72, FIND_WORD and not TRANSITION_TABLES); no
          // other algorithm sets it.
          if (result_word && find_word && !tables) *result_word = word;
        }
        out->eppstein_bound = upper;
      }
    } else if (name == "Beam") {
      if (open) {
        const BeamSettings bs = beam_settings(cfg, n, k);
        if (n == 1) {
          lower = upper = 0;
        } else {
          BeamResult b = beam(ctx, n, k, d, upper, bs, first.ran ? &first.r : nullptr);
          if (b.bound < upper || (b.bound == upper && word.size() != upper && b.word.size() == b.bound)) {
            upper = b.bound;
            word = b.word;
          }
        }
        out->beam_bound = upper;
      }
      first.ran = false;
    } else if (name == "Exact") {
      if (open) {
        sw_exact_config ec{};
        ec.dfs = eval_bool(cfg->get("dfs"), true, n, k) ? 1 : 0;
        // The device budget: gpu_max_memory_mb when the plan gives one, else
        // the reference's max_memory_mb (exact/config.hpp:24-46; 12 GiB in
        // configs/exact.json).
        const Json* mm = cfg->get("gpu_max_memory_mb");
        if (!mm) mm = cfg->get("max_memory_mb");
        if (mm) ec.max_memory_mb = uint64_t(eval_num(mm, "0", n, k));
        const Json* vb = cfg->get("verbose");
        ec.verbose = vb ? eval_bool(vb, false, n, k) : 0;
        // The backward deepening's smallest part and whether the budget is
        // strict (exact/config.hpp:24-46: dfs_min_list_size 8192,
        // strict_memory_limit false).
        ec.dfs_min_list_size = uint64_t(eval_num(cfg->get("dfs_min_list_size"), "8192", n, k));
        ec.strict_memory_limit = eval_bool(cfg->get("strict_memory_limit"), false, n, k) ? 1 : 0;
        const Json* gc = cfg->get("gpu_goal_check");
        if (gc) ec.goal_check = int32_t(eval_num(gc, "0", n, k));
        ExactOutcome r;
        if (red.reduced) {
          // The reduced automaton from the Reduce's list; lengths and words
          // get the Reduce's steps back (exact.cpp:83-87 of the reference).
          const uint64_t s = red.steps;
          ExactSeed seed;
          seed.rows = red.list.as<uint64_t>();
          seed.count = red.list_count;
          r = exact_search(ctx, red.n_red, k, red.delta_red.data(), lower > s ? lower - s : 0, upper - s, ec, true,
              &seed);
          r.r.length += s;
          if (!r.word.empty() && r.root >= 0) {
            std::vector<uint32_t> full = red.prefix(uint32_t(r.root));
            full.insert(full.end(), r.word.begin(), r.word.end());
            r.word = std::move(full);
            r.r.word_len = r.word.size();
          }
        } else {
          r = exact_search(ctx, n, k, d, lower, upper, ec);
        }
        out->exact = r.r;
        if (r.r.found) {
          lower = upper = r.r.length;
          if (!r.r.bound_word) word = r.word;
          else if (word.size() != upper) word.clear();
        } else {
          // The budget ran out ("mlsw >= x"): a lower bound, and any word the
          // search found on the way as a tighter upper bound.
          lower = std::max(lower, r.r.length);
          if (!r.word.empty() && r.word.size() < upper) {
            upper = r.word.size();
            word = r.word;
          }
        }
      }
    } else if (name == "Reduce") {
      // The reference's Reduce (algorithm/reduce/reduce.hpp:31-103): a short
      // forward search that may settle the length, or hand Exact a smaller
      // automaton and a list to start from (reduce.cpp).
      const long double min_n = eval_num(cfg->get("min_n"), "80", n, k);
      const long double thr = eval_num(cfg->get("list_size_threshold"), "AUT_N * AUT_K * 8", n, k);
      if (open && (long double)n >= min_n && upper != UINT64_MAX && !red.ran) {
        reduce_stage(ctx, n, k, d, upper, thr <= 0 ? 0 : uint64_t(thr), red);
        if (red.decided) {
          lower = upper = red.length;
          if (!red.word.empty()) word = red.word;
        } else if (red.reduced) {
          out->reduced_n = red.n_red;
          out->reduce_steps = uint32_t(red.steps);
        }
      }
    } else {
      throw Error{SW_ERR_UNSUPPORTED, "algorithm not supported by this library: " + name};
    }
    out->ms[idx] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::snprintf(out->names[idx], sizeof(out->names[idx]), "%s", name.c_str());
    ++idx;
  }
  out->n_algorithms = idx;
  out->non_synchro = non_synchro;
  out->lower = lower;
  out->upper = upper;
  if (non_synchro) word.clear();
  out->word_len = word.size();
}

}  // namespace sw

namespace {

int pipeline_api(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, const char* plan_json, sw_result* out,
    uint32_t* word, uint64_t word_cap, uint32_t* result_word, uint64_t result_word_cap, uint64_t* result_word_len) {
  try {
    SW_REQUIRE(ctx && delta && out, "NULL argument");
    SW_REQUIRE(n >= 1 && k >= 1 && k <= 255, "n must be >= 1 and k in [1, 255]");
    if (n > 512) throw sw::Error{SW_ERR_UNSUPPORTED, "n above 512 states is not compiled"};
    for (size_t i = 0; i < size_t(n) * k; ++i) SW_REQUIRE(delta[i] < n, "transition out of range");
    std::vector<uint32_t> w, rw;
    sw::pipeline(ctx, n, k, delta, plan_json, out, w, result_word_len ? &rw : nullptr);
    if (word && w.size() <= word_cap) std::copy(w.begin(), w.end(), word);
    if (result_word_len) {
      *result_word_len = out->non_synchro ? 0 : rw.size();
      if (result_word && rw.size() <= result_word_cap) std::copy(rw.begin(), rw.end(), result_word);
    }
    return SW_OK;
  } catch (const sw::Error& e) {
    sw::set_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    sw::set_error(e.what());
    return SW_ERR_INVALID;
  }
}

}  // namespace

extern "C" int sw_pipeline_solve(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, const char* plan_json,
    sw_result* out, uint32_t* word, uint64_t word_cap) {
  return pipeline_api(ctx, n, k, delta, plan_json, out, word, word_cap, nullptr, 0, nullptr);
}

extern "C" int sw_pipeline_solve_result_word(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta,
    const char* plan_json, sw_result* out, uint32_t* word, uint64_t word_cap, uint32_t* result_word,
    uint64_t result_word_cap, uint64_t* result_word_len) {
  if (!result_word_len) {
    sw::set_error("NULL argument");
    return SW_ERR_INVALID;
  }
  return pipeline_api(ctx, n, k, delta, plan_json, out, word, word_cap, result_word, result_word_cap, result_word_len);
}
