// The learned upper bound: a BF16 MLP that scores beam candidates.
//
// Role (and its limits): the reference's Beam keeps the candidates of largest
// cardinality (algorithm/beam/beam.hpp:98-188); any scoring keeps the bound
// valid, because only real preimages are followed and the bound is the depth
// at which a candidate equals Q.  Here the score is the network's estimate of
// the forward distance of a candidate S — the length of the shortest word u
// with delta(Q, u) inside S — learned from distances the exact search itself
// measures.  The exact search still decides every length.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <set>
#include <vector>

#include "sw_host.hpp"
#include "sw_kernels.hpp"
#include "sw_mlp.hpp"

struct sw_mlp {
  sw_ctx* ctx = nullptr;
  uint32_t n = 0, k = 0;
  sw::MlpDev p;
  sw::MlpAdam adam;
  sw::DevBuf params, bf, grads, moments;   // masters, bf16 copies, grads, Adam m|v
  sw::DevBuf h1f, h2f, y, dy, g1f, h1b, h2b, g1b, g2b, lossb;
  uint64_t work_rows = 0;
  uint64_t nparams = 0;
  uint32_t rank_group = 0;  // 0: squared error; G: ranking within groups of G
};

namespace sw {
namespace {

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double uniform(uint64_t& s) { return double(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

void bind(sw_mlp* m) {
  float* base = m->params.as<float>();
  const uint64_t D = m->p.D, H = m->p.H;
  m->p.w1 = base;
  m->p.b1 = m->p.w1 + H * D;
  m->p.w2 = m->p.b1 + H;
  m->p.b2 = m->p.w2 + H * H;
  m->p.w3 = m->p.b2 + H;
  m->p.b3 = m->p.w3 + H;
  m->p.w1b = m->bf.as<uint16_t>();
  m->p.w2b = m->p.w1b + H * D;
}

MlpGrads grads_of(sw_mlp* m) {
  float* base = m->grads.as<float>();
  const uint64_t D = m->p.D, H = m->p.H;
  MlpGrads g;
  g.w1 = base;
  g.b1 = g.w1 + H * D;
  g.w2 = g.b1 + H;
  g.b2 = g.w2 + H * H;
  g.w3 = g.b2 + H;
  g.b3 = g.w3 + H;
  return g;
}

MlpWork work_of(sw_mlp* m, uint64_t rows) {
  const uint64_t H = m->p.H;
  if (rows > m->work_rows) {
    m->h1f.ensure(rows * H * 4);
    m->h2f.ensure(rows * H * 4);
    m->g1f.ensure(rows * H * 4);
    m->h1b.ensure(rows * H * 2);
    m->h2b.ensure(rows * H * 2);
    m->g1b.ensure(rows * H * 2);
    m->g2b.ensure(rows * H * 2);
    m->y.ensure(rows * 4 + 4);
    m->dy.ensure(rows * 4 + 4);
    m->work_rows = rows;
  }
  MlpWork w;
  w.h1f = m->h1f.as<float>();
  w.h2f = m->h2f.as<float>();
  w.g1f = m->g1f.as<float>();
  w.h1b = m->h1b.as<uint16_t>();
  w.h2b = m->h2b.as<uint16_t>();
  w.g1b = m->g1b.as<uint16_t>();
  w.g2b = m->g2b.as<uint16_t>();
  w.y = m->y.as<float>();
  w.dy = m->dy.as<float>();
  return w;
}

}  // namespace

// The per-automaton side of the features: the images of Q by breadth-first
// search up to depth 2*kGuideDepths (distinct images, each at the first depth
// it appears), and the canonical state order — states by how often they
// occur in the deepest level's images, most frequent first, ties by state.
// Random automata share no state names; this order makes a feature column
// mean the same thing ("the r-th most often needed state") in every one.
struct Guide {
  DevBuf order, rows;
  FeatureGuide g;
};

void build_guide(uint32_t n, uint32_t k, const uint32_t* delta, Guide& out, hipStream_t st) {
  const uint32_t W = (n + 63) / 64;
  std::set<std::vector<uint64_t>> seen;
  std::vector<std::vector<uint64_t>> level(1, std::vector<uint64_t>(W, 0));
  for (uint32_t q = 0; q < n; ++q) level[0][q >> 6] |= 1ull << (q & 63);
  seen.insert(level[0]);
  std::vector<uint64_t> rows(level[0]);
  std::vector<std::vector<uint64_t>> last = level;
  uint32_t ends[kGuideDepths];
  const uint32_t depth = 2 * kGuideDepths;
  for (uint32_t d = 1; d <= depth; ++d) {
    std::vector<std::vector<uint64_t>> next;
    for (const auto& A : level)
      for (uint32_t a = 0; a < k; ++a) {
        std::vector<uint64_t> B(W, 0);
        for (uint32_t q = 0; q < n; ++q)
          if ((A[q >> 6] >> (q & 63)) & 1) {
            const uint32_t t = delta[size_t(q) * k + a];
            B[t >> 6] |= 1ull << (t & 63);
          }
        if (seen.insert(B).second) next.push_back(std::move(B));
      }
    for (const auto& B : next) rows.insert(rows.end(), B.begin(), B.end());
    if (!next.empty()) last = next;
    level.swap(next);
    if (d % 2 == 0) ends[d / 2 - 1] = uint32_t(rows.size() / W);
  }
  std::vector<uint32_t> freq(n, 0);
  for (const auto& A : last)
    for (uint32_t q = 0; q < n; ++q) freq[q] += (A[q >> 6] >> (q & 63)) & 1;
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return freq[a] > freq[b]; });
  std::vector<uint16_t> col_state(order.begin(), order.end());
  out.order.ensure(size_t(n) * 2 + 8);
  out.rows.ensure(rows.size() * 8 + 8);
  SW_CHECK_HIP(hipMemcpyAsync(out.order.p, col_state.data(), size_t(n) * 2, hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(hipMemcpyAsync(out.rows.p, rows.data(), rows.size() * 8, hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(hipStreamSynchronize(st));  // host vectors go out of scope
  out.g.order = out.order.as<uint16_t>();
  out.g.rows = out.rows.as<uint64_t>();
  for (int j = 0; j < kGuideDepths; ++j) out.g.ends[j] = ends[j];
}

// Device buffers of the beam, kept by the context between calls.
struct BeamWs {
  DevBuf dtf, dtb, list, cand, table, slot, keep, pos, uniq, uidx, x, scratch, keys, vals, vals2, small, scan_tmp,
      hist, mf, mb, sel;
  DevBuf sb_in, sb_meta, sb_word;  // one-workgroup beam: seed and transitions, depth table, word
  std::vector<uint64_t> host_mf, host_mb;  // per-state masks of the tables, until the next call
  size_t scan_bytes = 0;
  uint64_t* host = nullptr;  // pinned read-back
  BeamWs() { SW_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), 64, hipHostMallocDefault)); }
  ~BeamWs() {
    if (host) (void)hipHostFree(host);
  }
  BeamWs(const BeamWs&) = delete;
  BeamWs& operator=(const BeamWs&) = delete;
};

BeamWs& beam_workspace(sw_ctx* ctx) {
  if (!ctx->beam_ws) {
    ctx->beam_ws = new BeamWs();
    ctx->beam_ws_free = [](void* p) { delete static_cast<BeamWs*>(p); };
  }
  return *static_cast<BeamWs*>(ctx->beam_ws);
}

// Features of `count` device sets (images under every letter computed here).
void mlp_features(sw_mlp* m, const Guide& guide, const uint64_t* tables_fwd, const uint64_t* sets, uint64_t count,
    uint16_t* x, DevBuf& scratch, hipStream_t st) {
  const uint32_t n = m->n, k = m->k, W = (n + 63) / 64;
  scratch.ensure(count * k * W * 8 + 8);
  SW_CHECK_HIP(launch_expand(int(W), sets, count, tables_fwd, (n + 7) / 8, k, nullptr, n, 0, scratch.as<uint64_t>(),
      nullptr, st));
  SW_CHECK_HIP(launch_features(int(W), sets, scratch.as<uint64_t>(), count, n, k, m->p.D, guide.g, x, st));
}

}  // namespace sw

using sw::DevBuf;
using sw::Error;

namespace {

int fail(const Error& e) {
  sw::set_error(e.msg);
  return e.code;
}

}  // namespace

extern "C" {

int sw_mlp_create(sw_ctx* ctx, uint32_t n, uint32_t k, uint32_t hidden, uint64_t seed, sw_mlp** out) {
  try {
    SW_REQUIRE(ctx && out, "NULL argument");
    SW_REQUIRE(n >= 1 && n <= 512 && k >= 1 && k <= 255, "n in [1, 512], k in [1, 255]");
    SW_REQUIRE(hidden >= 16 && hidden <= 4096, "hidden in [16, 4096]");
    sw::use(ctx);
    auto m = std::make_unique<sw_mlp>();
    m->ctx = ctx;
    m->n = n;
    m->k = k;
    m->p.D = ((n + k + sw::kGuideFeatures) + 31) / 32 * 32;
    m->p.H = hidden;
    const uint64_t D = m->p.D, H = hidden;
    m->nparams = H * D + H + H * H + H + H + 1;
    m->params.ensure(m->nparams * 4);
    m->bf.ensure((H * D + H * H) * 2);
    m->grads.ensure(m->nparams * 4);
    m->moments.ensure(m->nparams * 8);
    m->lossb.ensure(16);
    sw::bind(m.get());
    // He-uniform initialisation from a splitmix64 stream; zero biases.
    std::vector<float> h(m->nparams, 0.f);
    uint64_t s = seed;
    auto fill = [&](uint64_t off, uint64_t cnt, double fan_in) {
      const double lim = std::sqrt(6.0 / fan_in);
      for (uint64_t i = 0; i < cnt; ++i) h[off + i] = float((2.0 * sw::uniform(s) - 1.0) * lim);
    };
    fill(0, H * D, double(n + k));
    fill(H * D + H, H * H, double(H));
    fill(H * D + H + H * H + H, H, double(H));
    SW_CHECK_HIP(hipMemcpyAsync(m->params.p, h.data(), h.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    SW_CHECK_HIP(hipMemsetAsync(m->moments.p, 0, m->nparams * 8, ctx->stream));
    m->adam.m = m->moments.as<float>();
    m->adam.v = m->adam.m + m->nparams;
    const uint64_t offs[6] = {0, H * D, H * D + H, H * D + H + H * H, H * D + H + H * H + H, H * D + 2 * H + H * H + H};
    for (int i = 0; i < 6; ++i) m->adam.off[i] = offs[i];
    SW_CHECK_HIP(sw::mlp_refresh_bf16(m->p, ctx->stream));
    SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    *out = m.release();
    return SW_OK;
  } catch (const Error& e) {
    return fail(e);
  }
}

void sw_mlp_destroy(sw_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  delete m;
}

int sw_mlp_shape(sw_mlp* m, uint32_t* in_dim, uint32_t* hidden, uint64_t* nparams) {
  if (!m) return SW_ERR_INVALID;
  if (in_dim) *in_dim = m->p.D;
  if (hidden) *hidden = m->p.H;
  if (nparams) *nparams = m->nparams;
  return SW_OK;
}

int sw_mlp_get_params(sw_mlp* m, float* out, uint64_t cap) {
  try {
    SW_REQUIRE(m && out && cap >= m->nparams, "NULL argument or buffer too small");
    sw::use(m->ctx);
    SW_CHECK_HIP(hipMemcpyAsync(out, m->params.p, m->nparams * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    SW_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
    return SW_OK;
  } catch (const Error& e) {
    return fail(e);
  }
}

int sw_mlp_set_params(sw_mlp* m, const float* in, uint64_t count) {
  try {
    SW_REQUIRE(m && in && count == m->nparams, "NULL argument or wrong parameter count");
    sw::use(m->ctx);
    SW_CHECK_HIP(hipMemcpyAsync(m->params.p, in, count * 4, hipMemcpyHostToDevice, m->ctx->stream));
    SW_CHECK_HIP(sw::mlp_refresh_bf16(m->p, m->ctx->stream));
    SW_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
    return SW_OK;
  } catch (const Error& e) {
    return fail(e);
  }
}

// Forward pass on raw BF16 features x[count][D] (host): y[count] (host).
int sw_mlp_forward(sw_mlp* m, const uint16_t* x, uint64_t count, float* y) {
  try {
    SW_REQUIRE(m && (x || !count) && (y || !count), "NULL argument");
    sw::use(m->ctx);
    if (!count) return SW_OK;
    hipStream_t st = m->ctx->stream;
    auto w = sw::work_of(m, count);
    m->ctx->a.ensure(count * m->p.D * 2);
    SW_CHECK_HIP(hipMemcpyAsync(m->ctx->a.p, x, count * m->p.D * 2, hipMemcpyHostToDevice, st));
    SW_CHECK_HIP(sw::mlp_forward(m->p, m->ctx->a.as<uint16_t>(), count, w, w.y, st));
    SW_CHECK_HIP(hipMemcpyAsync(y, w.y, count * 4, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    return SW_OK;
  } catch (const Error& e) {
    return fail(e);
  }
}

// Mean squared error and its gradient with respect to every parameter
// (flattened as sw_mlp_get_params) on one batch.
int sw_mlp_gradients(sw_mlp* m, const uint16_t* x, const float* target, uint64_t count, float* grads, float* loss) {
  try {
    SW_REQUIRE(m && x && target && grads && count, "NULL argument");
    sw::use(m->ctx);
    hipStream_t st = m->ctx->stream;
    auto w = sw::work_of(m, count);
    m->ctx->a.ensure(count * m->p.D * 2);
    m->ctx->b.ensure(count * 4);
    SW_CHECK_HIP(hipMemcpyAsync(m->ctx->a.p, x, count * m->p.D * 2, hipMemcpyHostToDevice, st));
    SW_CHECK_HIP(hipMemcpyAsync(m->ctx->b.p, target, count * 4, hipMemcpyHostToDevice, st));
    SW_CHECK_HIP(hipMemsetAsync(m->lossb.p, 0, 4, st));
    auto g = sw::grads_of(m);
    SW_CHECK_HIP(sw::mlp_forward(m->p, m->ctx->a.as<uint16_t>(), count, w, w.y, st));
    SW_REQUIRE(!m->rank_group || count % m->rank_group == 0, "count must be a multiple of the ranking group");
    SW_CHECK_HIP(sw::mlp_backward(m->p, m->ctx->a.as<uint16_t>(), count, m->ctx->b.as<float>(), w, g,
        m->lossb.as<float>(), m->rank_group, st));
    SW_CHECK_HIP(hipMemcpyAsync(grads, m->grads.p, m->nparams * 4, hipMemcpyDeviceToHost, st));
    if (loss) SW_CHECK_HIP(hipMemcpyAsync(loss, m->lossb.p, 4, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    return SW_OK;
  } catch (const Error& e) {
    return fail(e);
  }
}

int sw_mlp_set_loss(sw_mlp* m, uint32_t rank_group) {
  try {
    SW_REQUIRE(m, "NULL argument");
    SW_REQUIRE(rank_group == 0 || (rank_group >= 2 && rank_group <= sw::kRankMaxGroup), "rank_group 0 or 2..16");
    m->rank_group = rank_group;
    return SW_OK;
  } catch (const Error& e) {
    return fail(e);
  }
}

// Adam training on samples x[count][D], target[count] (host): `epochs` passes
// in shuffled minibatches; loss_curve[e] = mean loss of epoch e (may be NULL).
int sw_mlp_train(sw_mlp* m, const uint16_t* x, const float* target, uint64_t count, uint32_t epochs, uint32_t batch,
    float lr, uint64_t seed, float* loss_curve) {
  try {
    SW_REQUIRE(m && x && target && count && batch, "NULL argument");
    sw::use(m->ctx);
    hipStream_t st = m->ctx->stream;
    const uint64_t D = m->p.D;
    // Shuffle on the host once per epoch (whole groups under the ranking
    // loss); batches gathered on the host side.
    const uint32_t G = m->rank_group ? m->rank_group : 1;
    SW_REQUIRE(count % G == 0, "sample count must be a multiple of the ranking group");
    batch = batch / G * G;
    SW_REQUIRE(batch >= G, "batch smaller than one ranking group");
    DevBuf dx, dt;
    dx.ensure(uint64_t(batch) * D * 2);
    dt.ensure(uint64_t(batch) * 4);
    std::vector<uint64_t> idx(count / G);
    std::iota(idx.begin(), idx.end(), 0ull);
    std::vector<uint16_t> bx(uint64_t(batch) * D);
    std::vector<float> bt(batch);
    auto w = sw::work_of(m, batch);
    auto g = sw::grads_of(m);
    uint64_t s = seed;
    const uint64_t units = count / G, per = batch / G;
    for (uint32_t e = 0; e < epochs; ++e) {
      for (uint64_t i = units; i > 1; --i) std::swap(idx[i - 1], idx[sw::splitmix(s) % i]);
      SW_CHECK_HIP(hipMemsetAsync(m->lossb.p, 0, 4, st));
      uint64_t steps = 0;
      for (uint64_t b0 = 0; b0 + per <= units; b0 += per, ++steps) {
        for (uint32_t r = 0; r < batch; ++r) {
          const uint64_t src = idx[b0 + r / G] * G + r % G;
          std::memcpy(&bx[uint64_t(r) * D], x + src * D, D * 2);
          bt[r] = target[src];
        }
        SW_CHECK_HIP(hipMemcpyAsync(dx.p, bx.data(), bx.size() * 2, hipMemcpyHostToDevice, st));
        SW_CHECK_HIP(hipMemcpyAsync(dt.p, bt.data(), bt.size() * 4, hipMemcpyHostToDevice, st));
        SW_CHECK_HIP(sw::mlp_forward(m->p, dx.as<uint16_t>(), batch, w, w.y, st));
        SW_CHECK_HIP(sw::mlp_backward(m->p, dx.as<uint16_t>(), batch, dt.as<float>(), w, g, m->lossb.as<float>(),
            m->rank_group, st));
        SW_CHECK_HIP(sw::mlp_adam(m->p, g, m->adam, lr, st));
        SW_CHECK_HIP(hipStreamSynchronize(st));  // host batch buffers are reused
      }
      float l = 0.f;
      SW_CHECK_HIP(hipMemcpyAsync(&l, m->lossb.p, 4, hipMemcpyDeviceToHost, st));
      SW_CHECK_HIP(hipStreamSynchronize(st));
      if (loss_curve) loss_curve[e] = steps ? l / float(steps) : 0.f;
    }
    return SW_OK;
  } catch (const Error& e) {
    return fail(e);
  }
}

// BF16 features of host sets for automaton delta: x[count][D] (host).
int sw_mlp_features(sw_mlp* m, const uint32_t* delta, const uint64_t* sets, uint64_t count, uint16_t* x) {
  try {
    SW_REQUIRE(m && delta && (sets || !count) && (x || !count), "NULL argument");
    sw::use(m->ctx);
    if (!count) return SW_OK;
    hipStream_t st = m->ctx->stream;
    const uint32_t n = m->n, k = m->k, W = (n + 63) / 64;
    auto t = sw::build_tables(n, k, delta, 0);
    sw::Guide guide;
    sw::build_guide(n, k, delta, guide, st);
    DevBuf dtab, dsets, dx, scratch;
    dtab.ensure(t.size() * 8);
    dsets.ensure(count * W * 8);
    dx.ensure(count * m->p.D * 2);
    SW_CHECK_HIP(hipMemcpyAsync(dtab.p, t.data(), t.size() * 8, hipMemcpyHostToDevice, st));
    SW_CHECK_HIP(hipMemcpyAsync(dsets.p, sets, count * W * 8, hipMemcpyHostToDevice, st));
    sw::mlp_features(m, guide, dtab.as<uint64_t>(), dsets.as<uint64_t>(), count, dx.as<uint16_t>(), scratch, st);
    SW_CHECK_HIP(hipMemcpyAsync(x, dx.p, count * m->p.D * 2, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    return SW_OK;
  } catch (const Error& e) {
    return fail(e);
  }
}

}  // extern "C"

// ---------------------------------------------------------------- samples
namespace {

// Host beam (cardinality order, as the reference's) whose candidates are the
// training sets: per depth, `per_depth` children spread over the sorted list.
std::vector<uint64_t> beam_candidates(uint32_t n, uint32_t k, const uint32_t* d, uint32_t width, uint32_t max_depth,
    uint32_t per_depth) {
  const uint32_t W = (n + 63) / 64;
  auto tab = sw::build_tables(n, k, d, 1);
  const uint32_t S = (n + 7) / 8;
  auto apply = [&](const uint64_t* x, uint32_t a, uint64_t* r) {
    for (uint32_t w = 0; w < W; ++w) r[w] = 0;
    for (uint32_t s = 0; s < S; ++s) {
      uint32_t byte = uint32_t(x[s >> 3] >> ((s & 7) * 8)) & 255u;
      if (!byte) continue;
      const uint64_t* row = &tab[((size_t(a) * S + s) * 256 + byte) * W];
      for (uint32_t w = 0; w < W; ++w) r[w] |= row[w];
    }
  };
  auto pop = [&](const uint64_t* x) {
    int c = 0;
    for (uint32_t w = 0; w < W; ++w) c += __builtin_popcountll(x[w]);
    return c;
  };
  const std::vector<char> sink = sw::sink_states(n, k, d);
  std::vector<uint64_t> list, out;
  for (uint32_t q = 0; q < n; ++q)
    if (sink[q])
      for (uint32_t w = 0; w < W; ++w) list.push_back(w == q / 64 ? (1ull << (q % 64)) : 0ull);
  for (uint32_t depth = 1; depth <= max_depth && !list.empty(); ++depth) {
    const size_t m = list.size() / W;
    std::vector<uint64_t> next(m * k * W);
    for (uint32_t a = 0; a < k; ++a)
      for (size_t i = 0; i < m; ++i) apply(&list[i * W], a, &next[(a * m + i) * W]);
    std::vector<uint32_t> ord(m * k);
    std::iota(ord.begin(), ord.end(), 0u);
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
      int px = pop(&next[size_t(x) * W]), py = pop(&next[size_t(y) * W]);
      if (px != py) return px > py;
      return std::lexicographical_compare(&next[size_t(x) * W], &next[size_t(x) * W + W], &next[size_t(y) * W],
          &next[size_t(y) * W + W]);
    });
    std::vector<uint64_t> uniq;
    for (size_t j = 0; j < ord.size(); ++j) {
      const uint64_t* r = &next[size_t(ord[j]) * W];
      if (!uniq.empty() && std::equal(r, r + W, &uniq[uniq.size() - W])) continue;
      uniq.insert(uniq.end(), r, r + W);
    }
    const size_t u = uniq.size() / W;
    // Whole groups only: a depth with fewer candidates than per_depth adds none.
    const size_t take = u >= per_depth ? per_depth : 0;
    for (size_t t = 0; t < take; ++t) {
      const size_t j = take == 1 ? 0 : t * (u - 1) / (take - 1);
      out.insert(out.end(), &uniq[j * W], &uniq[j * W + W]);
    }
    if (pop(&uniq[0]) == int(n)) break;
    if (u > width) uniq.resize(size_t(width) * W);
    list.swap(uniq);
  }
  return out;
}

}  // namespace

extern "C" {

// Training samples from the exact search's own forward levels: candidate
// sets S of a cardinality beam, each labelled with its forward distance
// min{d : some image of Q at depth d lies inside S}, or fdepth + 1 when none
// up to fdepth.  Writes up to `cap` sets (W words each) and distances.
int sw_mlp_samples(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint32_t fdepth, uint32_t width,
    uint32_t bdepth, uint32_t per_depth, uint64_t* sets_out, float* dist_out, uint64_t cap, uint64_t* count) {
  try {
    SW_REQUIRE(ctx && delta && sets_out && dist_out && count, "NULL argument");
    SW_REQUIRE(n >= 2 && n <= 512 && k >= 1 && k <= 255, "n in [2, 512], k in [1, 255]");
    sw::use(ctx);
    hipStream_t st = ctx->stream;
    const uint32_t W = (n + 63) / 64;
    std::vector<uint64_t> cand = beam_candidates(n, k, delta, width, bdepth, per_depth);
    SW_REQUIRE(per_depth >= 1, "per_depth >= 1");
    uint64_t c = std::min<uint64_t>(cand.size() / W, cap / per_depth * per_depth);  // whole groups
    DevBuf levels, dcand, flags;
    std::vector<uint64_t> level_end;
    sw::forward_levels(ctx, n, k, delta, fdepth, 1ull << 24, levels, level_end);
    dcand.ensure(c * W * 8 + 8);
    SW_CHECK_HIP(hipMemcpyAsync(dcand.p, cand.data(), c * W * 8, hipMemcpyHostToDevice, st));
    std::vector<float> dist(c, float(level_end.size()));
    std::vector<char> set(c, 0);
    const size_t fb = sw::vsub_flag_bytes(int(W), c);
    flags.ensure(fb + 8);
    std::vector<uint8_t> bits(fb);
    uint64_t lo = 0;
    for (size_t d = 0; d < level_end.size(); ++d) {
      const uint64_t hi = level_end[d];
      SW_CHECK_HIP(hipMemsetAsync(flags.p, 0, fb, st));
      sw::VsubParams p;
      p.stream = levels.as<uint64_t>() + lo * W;
      p.s_count = hi - lo;
      p.tile = dcand.as<uint64_t>();
      p.t_count = c;
      p.n = n;
      p.mode = sw::VSUB_FLAGS;
      p.flags = flags.p;
      sw::apply_tuning(ctx, int(W), p);
      SW_CHECK_HIP(sw::launch_vsub(int(W), p, st));
      SW_CHECK_HIP(hipMemcpyAsync(bits.data(), flags.p, fb, hipMemcpyDeviceToHost, st));
      SW_CHECK_HIP(hipStreamSynchronize(st));
      for (uint64_t i = 0; i < c; ++i)
        if (!set[i] && ((bits[i >> 3] >> (i & 7)) & 1)) {
          set[i] = 1;
          dist[i] = float(d);
        }
      lo = hi;
    }
    std::copy(cand.begin(), cand.begin() + c * W, sets_out);
    std::copy(dist.begin(), dist.end(), dist_out);
    *count = c;
    return SW_OK;
  } catch (const Error& e) {
    return fail(e);
  }
}

}  // extern "C"

namespace sw {

// Beam search on the device: from the sink singletons with two preimages
// under a letter, expand by preimages, drop duplicates, keep the first
// `width` candidates — in the network's score order, or (m == NULL) in the
// reference's Beam order (larger sets first, then by their words,
// beam.hpp:98-188); the first depth producing Q bounds the length.  Returns
// *bound (upper when not improved) and the word of that length.  With a
// stop cell, the beam also ends at the cell's depth once the cell holds the
// stop's generation (read every 16 depths in the one-workgroup kernel, every
// depth on the device-wide path): the caller then treats any bound at or past
// that depth as no improvement.
int beam_device(sw_ctx* ctx, sw_mlp* m, uint32_t n, uint32_t k, const uint32_t* delta, uint32_t width,
    uint64_t upper, uint64_t* bound, uint32_t* word, uint64_t word_cap, uint64_t* word_len, const BeamStop* stop) {
  try {
    SW_REQUIRE(ctx && delta && bound && width, "NULL argument");
    SW_REQUIRE(!m || (m->n == n && m->k == k), "network built for another (n, k)");
    SW_REQUIRE(n >= 2 && n <= 512 && k >= 1 && k <= 255, "n in [2, 512], k in [1, 255]");
    sw::use(ctx);
    hipStream_t st = ctx->stream;
    const uint32_t W = (n + 63) / 64, S = (n + 7) / 8;
    *bound = upper;
    if (word_len) *word_len = 0;
    sw::Guide guide;
    if (m) sw::build_guide(n, k, delta, guide, st);
    sw::BeamWs& ws = sw::beam_workspace(ctx);
    DevBuf &dtf = ws.dtf, &dtb = ws.dtb, &list = ws.list, &cand = ws.cand, &table = ws.table, &slot = ws.slot,
           &keep = ws.keep, &pos = ws.pos, &uniq = ws.uniq, &uidx = ws.uidx, &x = ws.x, &scratch = ws.scratch,
           &keys = ws.keys, &vals = ws.vals, &vals2 = ws.vals2, &small = ws.small, &scan_tmp = ws.scan_tmp,
           &histd = ws.hist;
    size_t &scan_bytes = ws.scan_bytes;
    if (m) sw::build_tables_device(n, k, delta, 0, ws.host_mf, ws.mf, dtf, st);  // images: the features only
    bool tables_built = false;
    auto build_preimage_tables = [&]() {
      if (!tables_built) sw::build_tables_device(n, k, delta, 1, ws.host_mb, ws.mb, dtb, st);
      tables_built = true;
    };
    // The reference's start (beam.hpp:104-116): sink-component states with two
    // preimages under some letter, as singletons.
    const std::vector<char> sink = sw::sink_states(n, k, delta);
    std::vector<uint32_t> indeg(size_t(n) * k, 0);
    for (uint32_t q = 0; q < n; ++q)
      for (uint32_t a = 0; a < k; ++a) indeg[size_t(delta[size_t(q) * k + a]) * k + a]++;
    std::vector<uint64_t> seed;
    for (uint32_t q = 0; q < n; ++q) {
      if (!sink[q]) continue;
      bool two = false;
      for (uint32_t a = 0; a < k; ++a) two |= indeg[size_t(q) * k + a] > 1;
      if (two)
        for (uint32_t w = 0; w < W; ++w) seed.push_back(w == q / 64 ? (1ull << (q % 64)) : 0ull);
    }
    if (seed.empty()) return SW_OK;  // no state with two preimages: nothing to follow
    uint64_t m0 = seed.size() / W;
    list.ensure(seed.size() * 8 + 8);
    SW_CHECK_HIP(hipMemcpyAsync(list.p, seed.data(), seed.size() * 8, hipMemcpyHostToDevice, st));
    small.ensure(64);
    // Word history on the device: per depth, the candidate index of every
    // kept set (its parent is index % parents, its letter index / parents).
    std::vector<uint64_t> hist_off, hist_m;
    uint64_t hist_len = 0;
    uint64_t len0 = 1;
    // Automata of at most 64 states without a network: depths in one
    // workgroup, one launch (k_beam_small.hip), while they fit its LDS; a
    // depth that does not is continued below from the list and history the
    // kernel hands back.
    constexpr uint64_t kSmallDepths = 1u << 20;
    constexpr uint32_t kSmallHist = 1u << 22;
    if (!m && n <= 64 && k <= sw::kBeamSmallMaxK && upper <= kSmallDepths && m0 <= sw::kBeamSmallList) {
      DevBuf &sin = ws.sb_in, &smeta = ws.sb_meta, &sword = ws.sb_word;
      const size_t din_bytes = size_t(n) * k * 4 + 64 * 8;
      sin.ensure(din_bytes);
      smeta.ensure(size_t(upper) * 8 + 8);
      sword.ensure(size_t(upper) * 4 + 16 + 4);
      histd.ensure(size_t(kSmallHist) * 4, st, hist_len * 4);
      list.ensure(size_t(sw::kBeamSmallCand) * 8 + 8);
      std::vector<uint8_t> hin(din_bytes, 0);
      std::memcpy(hin.data(), seed.data(), seed.size() * 8);
      std::memcpy(hin.data() + 64 * 8, delta, size_t(n) * k * 4);
      SW_CHECK_HIP(hipMemcpyAsync(sin.p, hin.data(), din_bytes, hipMemcpyHostToDevice, st));
      sw::BeamSmallArgs a{};
      a.n = n;
      a.k = k;
      a.width = width;
      a.upper = uint32_t(upper);
      a.m0 = uint32_t(m0);
      a.seed = sin.as<uint64_t>();
      a.delta = reinterpret_cast<const uint32_t*>(sin.as<uint8_t>() + 64 * 8);
      a.hist = histd.as<uint32_t>();
      a.hist_cap = kSmallHist;
      a.meta = smeta.as<uint32_t>();
      a.meta_cap = uint32_t(upper);
      // The four result words, then the word: one read-back for both.
      a.out = sword.as<uint32_t>();
      a.word = sword.as<uint32_t>() + 4;
      a.list_out = list.as<uint64_t>();
      if (stop) {
        a.stop = stop->cell;
        a.stop_gen = stop->gen;
      }
      SW_CHECK_HIP(sw::launch_beam_small(a, st));
      std::vector<uint32_t> back(4 + size_t(upper));
      SW_CHECK_HIP(hipMemcpyAsync(back.data(), sword.p, back.size() * 4, hipMemcpyDeviceToHost, st));
      SW_CHECK_HIP(hipStreamSynchronize(st));
      const uint32_t* o = back.data();
      if (o[0] == sw::kBeamSmallFound) {
        std::vector<uint32_t> w(back.begin() + 4, back.begin() + 4 + o[1]);
        *bound = o[1];
        if (word_len) *word_len = w.size();
        if (word && w.size() <= word_cap) std::copy(w.begin(), w.end(), word);
        return SW_OK;
      }
      if (o[0] != sw::kBeamSmallHandover) return SW_OK;  // no depth below upper produced Q
      // Continue at depth o[1] with the kernel's list and history.
      std::vector<uint32_t> meta(size_t(o[1] - 1) * 2);
      if (!meta.empty())
        SW_CHECK_HIP(hipMemcpy(meta.data(), smeta.p, meta.size() * 4, hipMemcpyDeviceToHost));
      for (size_t d = 0; d + 1 < size_t(o[1]); ++d) {
        hist_off.push_back(meta[2 * d]);
        hist_m.push_back(meta[2 * d + 1]);
      }
      hist_len = o[3];
      m0 = o[2];
      len0 = o[1];
    }
    build_preimage_tables();
    // The hash table is emptied whole once (and when it grows); after that
    // each depth's compaction empties the slots it filled.  The full-set
    // index is reset once: a depth that sets it ends the beam.  The index and
    // the depth's number of distinct sets are written by the kernels straight
    // into pinned host memory: a depth waits once and copies nothing.
    unsigned long long* first = reinterpret_cast<unsigned long long*>(ws.host + 6);
    uint32_t* total = reinterpret_cast<uint32_t*>(ws.host + 5);
    bool table_clean = false;
    for (uint64_t len = len0; len < upper; ++len) {
      const uint64_t cnt = m0 * k;
      cand.ensure(cnt * W * 8 + 8);
      uint64_t tcap = 1024;
      while (tcap < 2 * cnt) tcap <<= 1;
      const void* table_before = table.p;
      table.ensure(tcap * 4);
      slot.ensure(cnt * 4 + 4);
      keep.ensure(cnt * 4 + 4);
      pos.ensure(cnt * 4 + 4);
      if (!table_clean || table.p != table_before) {
        SW_CHECK_HIP(sw::launch_beam_prep(table.as<uint32_t>(), table.cap / 4, first, st));
        table_clean = true;
      }
      SW_CHECK_HIP(sw::launch_expand(int(W), list.as<uint64_t>(), m0, dtb.as<uint64_t>(), S, k, nullptr, n, 0,
          cand.as<uint64_t>(), nullptr, st));
      // Dedup within the step (the smallest candidate index of equal sets
      // survives), launched before the one read-back of the step.
      SW_CHECK_HIP(sw::launch_hash_insert(int(W), cand.as<uint64_t>(), nullptr, cnt, table.as<uint32_t>(), tcap - 1,
          cand.as<uint64_t>(), slot.as<uint32_t>(), nullptr, st));
      size_t need = 0;
      SW_CHECK_HIP(sw::keep_scan_temp_bytes(cnt, &need));
      if (need > scan_bytes) {
        scan_tmp.release();
        scan_tmp.ensure(need);
        scan_bytes = need;
      }
      SW_CHECK_HIP(sw::launch_keep_scan(1, slot.as<uint32_t>(), table.as<uint32_t>(), nullptr, cnt,
          keep.as<uint32_t>(), pos.as<uint32_t>(), total, scan_tmp.p, &scan_bytes, st));
      uniq.ensure(cnt * W * 8 + 8);
      uidx.ensure(cnt * 4 + 4);
      // The compaction also finds Q (the first full candidate) and writes the
      // history indices of this depth in place, which stand when every
      // distinct set is kept.
      histd.ensure((hist_len + cnt) * 4 + 4, st, hist_len * 4);
      SW_CHECK_HIP(sw::launch_beam_select(int(W), cand.as<uint64_t>(), keep.as<uint32_t>(), pos.as<uint32_t>(), cnt,
          n, uniq.as<uint64_t>(), uidx.as<uint32_t>(), histd.as<uint32_t>() + hist_len, first, table.as<uint32_t>(),
          slot.as<uint32_t>(), st));
      SW_CHECK_HIP(hipStreamSynchronize(st));
      const unsigned long long f = *reinterpret_cast<volatile unsigned long long*>(first);
      const uint32_t nu = *reinterpret_cast<volatile uint32_t*>(total);
      if (stop) {
        const unsigned long long v = __atomic_load_n(stop->cell, __ATOMIC_ACQUIRE);
        if (uint32_t(v >> 32) == stop->gen && len >= uint32_t(v)) break;  // capped there
      }
      if (f != ~0ull) {
        // Word: letter of this candidate, then back through the kept parents.
        std::vector<uint32_t> h(hist_len);
        if (hist_len) {
          SW_CHECK_HIP(hipMemcpyAsync(h.data(), histd.p, hist_len * 4, hipMemcpyDeviceToHost, st));
          SW_CHECK_HIP(hipStreamSynchronize(st));
        }
        std::vector<uint32_t> w;
        uint64_t c = f;
        w.push_back(uint32_t(c / m0));
        uint64_t parent = c % m0;
        for (size_t dd = hist_off.size(); dd-- > 0;) {
          const uint64_t ci = h[hist_off[dd] + parent];
          w.push_back(uint32_t(ci / hist_m[dd]));
          parent = ci % hist_m[dd];
        }
        *bound = len;
        if (word_len) *word_len = w.size();
        if (word && w.size() <= word_cap) std::copy(w.begin(), w.end(), word);
        break;
      }
      // Order the distinct candidates and keep the first `width`: by the
      // network's score, or (m == NULL) in the reference's Beam order.  Up to
      // the counted maximum the kept sets are written straight from their
      // places; beyond it a sort gives the order and two gathers follow.
      const uint64_t keepn = std::min<uint64_t>(width, nu);
      histd.ensure((hist_len + keepn) * 4 + 4, st, hist_len * 4);
      if (nu <= width) {
        // Every distinct set is kept.  Which sets a depth keeps never depends
        // on their stored order (the order is total on distinct sets, and a
        // network's only says which `width` come first), and neither does
        // the bound, so they stay in candidate order: no ranking, and the
        // list is the compacted buffer itself.
        std::swap(list.p, uniq.p);
        std::swap(list.cap, uniq.cap);
        hist_off.push_back(hist_len);  // this depth's history: written by the compaction
        hist_m.push_back(m0);
        hist_len += keepn;
        m0 = keepn;
        if (keepn == 0) break;  // every continuation was empty
        continue;
      }
      list.ensure(keepn * W * 8 + 8);
      const bool counted = sw::beam_rank_fits(nu);
      const float* scores = nullptr;
      if (m) {
        x.ensure(uint64_t(nu) * m->p.D * 2 + 8);
        sw::mlp_features(m, guide, dtf.as<uint64_t>(), uniq.as<uint64_t>(), nu, x.as<uint16_t>(), scratch, st);
        auto wk = sw::work_of(m, nu);
        SW_CHECK_HIP(sw::mlp_forward(m->p, x.as<uint16_t>(), nu, wk, wk.y, st));
        scores = wk.y;
      }
      if (counted) {
        vals.ensure(uint64_t(nu) * 4 + 4);
        if (m)
          SW_CHECK_HIP(sw::beam_select_scores(scores, uniq.as<uint64_t>(), uidx.as<uint32_t>(), nu, W, keepn,
              vals.as<uint32_t>(), list.as<uint64_t>(), histd.as<uint32_t>() + hist_len, st));
        else
          SW_CHECK_HIP(sw::beam_select_reference(uniq.as<uint64_t>(), uidx.as<uint32_t>(), nu, W, n, keepn,
              vals.as<uint32_t>(), list.as<uint64_t>(), histd.as<uint32_t>() + hist_len, st));
      } else {
        // Past the counted maximum: the kept sets found by selection (no
        // sort; their stored order does not matter), then compacted in
        // candidate order with their history indices.
        vals.ensure(uint64_t(nu) * 4 + 4);
        vals2.ensure(uint64_t(nu) * 4 + 4);
        keys.ensure(uint64_t(nu) + 8);
        pos.ensure(uint64_t(nu) * 4 + 4);
        ws.sel.ensure(sw::beam_select_ctl_bytes());
        SW_CHECK_HIP(sw::beam_select_first(uniq.as<uint64_t>(), scores, nu, W, n, keepn, keys.as<uint8_t>(),
            vals.as<uint32_t>(), ws.sel.p, reinterpret_cast<uint32_t*>(ws.host + 7), st));
        size_t need = 0;
        SW_CHECK_HIP(sw::keep_scan_temp_bytes(nu, &need));
        if (need > scan_bytes) {
          scan_tmp.release();
          scan_tmp.ensure(need);
          scan_bytes = need;
        }
        uint32_t* kept_total = small.as<uint32_t>() + 4;  // (keepn, known already)
        SW_CHECK_HIP(sw::launch_keep_scan(0, nullptr, nullptr, nullptr, nu, vals.as<uint32_t>(), pos.as<uint32_t>(),
            kept_total, scan_tmp.p, &scan_bytes, st));
        SW_CHECK_HIP(sw::launch_select(int(W), uniq.as<uint64_t>(), 0, vals.as<uint32_t>(), pos.as<uint32_t>(), nu,
            list.as<uint64_t>(), vals2.as<uint32_t>(), st));
        SW_CHECK_HIP(sw::launch_gather_u32(uidx.as<uint32_t>(), vals2.as<uint32_t>(), keepn,
            histd.as<uint32_t>() + hist_len, st));
      }
      hist_off.push_back(hist_len);
      hist_m.push_back(m0);
      hist_len += keepn;
      m0 = keepn;
      if (keepn == 0) break;  // every continuation was empty
    }
    return SW_OK;
  } catch (const Error& e) {
    return fail(e);
  }
}

}  // namespace sw

extern "C" {

int sw_beam_device(sw_ctx* ctx, sw_mlp* m, uint32_t n, uint32_t k, const uint32_t* delta, uint32_t width,
    uint64_t upper, uint64_t* bound, uint32_t* word, uint64_t word_cap, uint64_t* word_len) {
  return sw::beam_device(ctx, m, n, k, delta, width, upper, bound, word, word_cap, word_len, nullptr);
}

}  // extern "C"
