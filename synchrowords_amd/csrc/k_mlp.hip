// The learned scorer: a BF16 MLP on MFMA with FP32 accumulation.
//
// Inputs of one sample: the state bitset of a set S (n values in {0, 1}) and,
// per letter a, |delta_a(S)| / n, zero-padded to D (a multiple of 32).
// Network: D -> H -> H -> 1 with ReLU, weights in BF16 (FP32 masters kept for
// training), FP32 accumulation everywhere.  Target: the forward distance of S
// (the length of the shortest word u with delta(Q, u) inside S), as measured
// by the exact search's own forward levels.
//
// GEMM tiles use gfx950's double-rate v_mfma_f32_16x16x32_bf16: a workgroup
// of 4 waves computes a 64x64 tile of C, each wave a 32x32 quadrant of 2x2
// MFMA blocks; K advances 32 at a time through LDS, one MFMA per block and
// step, its operands one 16-byte LDS read per lane.  The network is small (about 0.4 M parameters), so
// the layers are latency- rather than MFMA-bound at beam batch sizes.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "sw_kernels.hpp"
#include "sw_mlp.hpp"


#include <algorithm>
#include <utility>

namespace sw {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {
  // Round to nearest even.
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16) | ((u & 0xffffu) ? 0x40 : 0);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

constexpr int TM = 64, TN = 64, TK = 32;

// C[M x N] (+)= op(A)[M x K] * op(B)[K x N].
//   A: row-major [M][K] (lda) or, with TA, [K][M].
//   B: row-major [K][N] (ldb) or, with TB, [N][K].
// Epilogue: + bias[n] (if bias), ReLU (if relu); write f32 (Cf) and/or bf16 (Cb).
// VEC (no TA, TB, K and both leading dimensions multiples of 8, 16-byte
// aligned rows): both tiles are contiguous along k and stage as one 16-byte
// load per thread each.
template <bool TA, bool TB, bool VEC>
__global__ __launch_bounds__(256) void k_gemm(const uint16_t* __restrict__ A, int lda, const uint16_t* __restrict__ B,
    int ldb, float* __restrict__ Cf, uint16_t* __restrict__ Cb, int ldc, int M, int N, int K,
    const float* __restrict__ bias, int relu, int accumulate) {
  __shared__ __align__(16) uint16_t As[TM][TK + 8];  // [m][k], rows of 80 bytes
  __shared__ __align__(16) uint16_t Bs[TN][TK + 8];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < K; k0 += TK) {
    // Stage the tiles (zero outside the matrices).
    if (VEC) {
      static_assert(TM * TK / 8 == 256 && TN * TK / 8 == 256, "one 16-byte chunk per thread and tile");
      const int i = tid >> 2, c = (tid & 3) * 8;
      uint4 va = make_uint4(0u, 0u, 0u, 0u), vb = make_uint4(0u, 0u, 0u, 0u);
      if (m0 + i < M && k0 + c < K) va = *reinterpret_cast<const uint4*>(A + size_t(m0 + i) * lda + k0 + c);
      if (n0 + i < N && k0 + c < K) vb = *reinterpret_cast<const uint4*>(B + size_t(n0 + i) * ldb + k0 + c);
      *reinterpret_cast<uint4*>(&As[i][c]) = va;
      *reinterpret_cast<uint4*>(&Bs[i][c]) = vb;
    } else {
    for (int e = tid; e < TM * TK; e += 256) {
      int i, k;
      if (TA) {
        k = e / TM;
        i = e % TM;
      } else {
        i = e / TK;
        k = e % TK;
      }
      const int gm = m0 + i, gk = k0 + k;
      uint16_t v = 0;
      if (gm < M && gk < K) v = TA ? A[size_t(gk) * lda + gm] : A[size_t(gm) * lda + gk];
      As[i][k] = v;
    }
    for (int e = tid; e < TN * TK; e += 256) {
      int j, k;
      if (TB) {
        j = e / TK;
        k = e % TK;
      } else {
        k = e / TN;
        j = e % TN;
      }
      const int gn = n0 + j, gk = k0 + k;
      uint16_t v = 0;
      if (gn < N && gk < K) v = TB ? B[size_t(gn) * ldb + gk] : B[size_t(gk) * ldb + gn];
      Bs[j][k] = v;
    }
    }
    __syncthreads();
    static_assert(TK % 32 == 0 && (TK + 8) % 8 == 0, "k blocks of 32; rows of whole 16-byte lines");
#pragma unroll
    for (int kk = 0; kk < TK; kk += 32) {
      // gfx950's double-rate 16x16x32 block: lane l holds A[l%16][8*(l/16) .. +7]
      // and B[8*(l/16) .. +7][l%16], one 16-byte LDS read each.
      const int r = lane & 15, kq = kk + 8 * (lane >> 4);
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + i * 16 + r][kq]);
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + j * 16 + r][kq]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // D block layout: lane l holds D[4*(l/16) + v][l%16], v = 0..3.
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = m0 + wm + i * 16 + 4 * (lane >> 4) + v;
        const int gn = n0 + wn + j * 16 + (lane & 15);
        if (gm >= M || gn >= N) continue;
        float x = acc[i][j][v];
        if (accumulate && Cf) x += Cf[size_t(gm) * ldc + gn];
        if (bias) x += bias[gn];
        if (relu) x = x > 0.f ? x : 0.f;
        if (Cf) Cf[size_t(gm) * ldc + gn] = x;
        if (Cb) Cb[size_t(gm) * ldc + gn] = f2bf(x);
      }
}

// Features of set rows S_i (D columns, BF16):
//   x[perm[q]]      = 1 if state q is in S_i (states in the automaton's canonical order),
//   x[n + a]        = s |delta_a(S_i)| / n for each letter a,
//   x[n + k + j]    = s min over images A of Q of depth <= 2(j+1) of |A \ S_i| / n (j < 4),
//   x[n + k + 4]    = s |S_i| / n,
// zero elsewhere; s = kFeatureScale puts these few count features on a scale
// that competes with the n bits.  One wave per row: lanes take the columns
// (column r is state order[r], the inverse of perm) and split the guide
// images, whose minima meet in a wave reduction.
template <int W>
__global__ __launch_bounds__(256) void k_features(const uint64_t* __restrict__ sets,
    const uint64_t* __restrict__ images, uint64_t count, uint32_t n, uint32_t k, uint32_t D, FeatureGuide g,
    uint16_t* __restrict__ x) {
  const uint64_t i = uint64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  if (i >= count) return;  // whole waves
  uint16_t* row = x + i * D;
  uint64_t s[W];
  int card = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    s[w] = sets[i * W + w];
    card += __popcll(s[w]);
  }
  for (uint32_t r = lane; r < n; r += 64) {
    const uint32_t q = g.order[r];
    uint64_t word = 0;
#pragma unroll
    for (int w = 0; w < W; ++w)  // constant indices: s stays in registers
      if (uint32_t(w) == (q >> 6)) word = s[w];
    row[r] = ((word >> (q & 63)) & 1ull) ? uint16_t(0x3f80) : uint16_t(0);  // 1.0 or 0
  }
  for (uint32_t a = lane; a < k; a += 64) {
    const uint64_t* im = images + (uint64_t(a) * count + i) * W;
    int c = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) c += __popcll(im[w]);
    row[n + a] = f2bf(kFeatureScale * float(c) / float(n));
  }
  uint32_t best[kGuideDepths];
#pragma unroll
  for (int j = 0; j < kGuideDepths; ++j) best[j] = n;
  for (uint32_t r = lane; r < g.ends[kGuideDepths - 1]; r += 64) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) c += __popcll(g.rows[uint64_t(r) * W + w] & ~s[w]);
#pragma unroll
    for (int j = 0; j < kGuideDepths; ++j)
      if (r < g.ends[j]) best[j] = c < best[j] ? c : best[j];
  }
#pragma unroll
  for (int j = 0; j < kGuideDepths; ++j)
    for (int off = 32; off > 0; off >>= 1) {
      const uint32_t o = __shfl_xor(best[j], off, 64);
      best[j] = o < best[j] ? o : best[j];
    }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < kGuideDepths; ++j) row[n + k + j] = f2bf(kFeatureScale * float(best[j]) / float(n));
    row[n + k + kGuideDepths] = f2bf(kFeatureScale * float(card) / float(n));
  }
  for (uint32_t r = n + k + kGuideDepths + 1 + lane; r < D; r += 64) row[r] = 0;
}

// y[i] = h[i] . w3 + b3 (FP32), one wave per row.
__global__ void k_head(const float* __restrict__ h, const float* __restrict__ w3, const float* __restrict__ b3,
    uint64_t count, uint32_t H, float* __restrict__ y) {
  const uint64_t row = uint64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= count) return;
  float s = 0.f;
  for (uint32_t j = lane; j < H; j += 64) s += h[row * H + j] * w3[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) y[row] = s + b3[0];
}

// Loss gradient of the head and first hidden gradient:
//   dy = 2 (y - t) / B;  dH2 = dy * w3 (masked by ReLU(H2) > 0), bf16 + f32.
// With dy_given, dy already holds the loss gradient (the ranking loss).
__global__ void k_head_back(const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ h2,
    const float* __restrict__ w3, uint64_t count, uint32_t H, float inv_b, int dy_given, float* __restrict__ dy,
    uint16_t* __restrict__ dh2, float* __restrict__ loss) {
  const uint64_t row = blockIdx.x;
  if (row >= count) return;
  float d;
  if (dy_given) {
    d = dy[row];
  } else {
    d = 2.f * (y[row] - t[row]) * inv_b;
    if (threadIdx.x == 0) {
      dy[row] = d;
      atomicAdd(loss, (y[row] - t[row]) * (y[row] - t[row]) * inv_b);
    }
  }
  for (uint32_t j = threadIdx.x; j < H; j += blockDim.x) {
    const float g = h2[row * H + j] > 0.f ? d * w3[j] : 0.f;
    dh2[row * H + j] = f2bf(g);
  }
}

// Pairwise ranking loss over groups of G consecutive rows (one thread per
// group): for every pair (a, b) of a group with t_a < t_b (a is closer to Q),
// loss += softplus(y_a - y_b), so a learns to score below b.  Normalised by
// the number of ordered pairs the batch could hold, groups * G(G-1)/2.
__global__ void k_rank_dy(const float* __restrict__ y, const float* __restrict__ t, uint64_t groups, uint32_t G,
    float inv_pairs, float* __restrict__ dy, float* __restrict__ loss) {
  const uint64_t gi = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gi >= groups) return;
  const uint64_t base = gi * G;
  float d[kRankMaxGroup];
  for (uint32_t a = 0; a < G; ++a) d[a] = 0.f;
  float l = 0.f;
  for (uint32_t a = 0; a < G; ++a)
    for (uint32_t b = 0; b < G; ++b) {
      if (!(t[base + a] < t[base + b])) continue;
      const float z = y[base + a] - y[base + b];
      const float sg = 1.f / (1.f + expf(-z));
      l += (z > 0.f ? z : 0.f) + log1pf(expf(-fabsf(z)));
      d[a] += sg;
      d[b] -= sg;
    }
  for (uint32_t a = 0; a < G; ++a) dy[base + a] = d[a] * inv_pairs;
  atomicAdd(loss, l * inv_pairs);
}

// dW3[j] = sum_i dy[i] * H2[i][j]; db3 = sum dy.  One block per 256 columns.
__global__ void k_head_wgrad(const float* __restrict__ dy, const float* __restrict__ h2, uint64_t count, uint32_t H,
    float* __restrict__ gw3, float* __restrict__ gb3) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < H) {
    float s = 0.f;
    for (uint64_t i = 0; i < count; ++i) s += dy[i] * h2[i * H + j];
    gw3[j] = s;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float s = 0.f;
    for (uint64_t i = 0; i < count; ++i) s += dy[i];
    gb3[0] = s;
  }
}

// g (bf16 in, f32 master) masked by ReLU of the layer output; column sums -> bias grads.
__global__ void k_relu_mask(const uint16_t* __restrict__ g, const float* __restrict__ h, uint64_t count, uint32_t H,
    uint16_t* __restrict__ out) {
  const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= count * H) return;
  out[e] = h[e] > 0.f ? g[e] : uint16_t(0);
}

__global__ void k_colsum_bf16(const uint16_t* __restrict__ g, uint64_t count, uint32_t H, float* __restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= H) return;
  float s = 0.f;
  for (uint64_t i = 0; i < count; ++i) s += bf2f(g[i * H + j]);
  out[j] = s;
}

// Adam on FP32 masters, then refresh the BF16 copy.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
    uint64_t count, float lr, float b1, float b2, float eps, float c1, float c2, uint16_t* __restrict__ wb) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float* __restrict__ x, uint64_t count, uint16_t* __restrict__ out) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) out[i] = f2bf(x[i]);
}

inline unsigned blocks(uint64_t n, unsigned bs = 256) { return unsigned((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_gemm(bool ta, bool tb, const uint16_t* A, int lda, const uint16_t* B, int ldb, float* Cf,
    uint16_t* Cb, int ldc, int M, int N, int K, const float* bias, int relu, int accumulate, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
#define L(a, b, v) hipLaunchKernelGGL((k_gemm<a, b, v>), grid, dim3(256), 0, st, A, lda, B, ldb, Cf, Cb, ldc, M, N, K, \
    bias, relu, accumulate)
  const bool vec = !ta && tb && K % 8 == 0 && lda % 8 == 0 && ldb % 8 == 0 &&
      (reinterpret_cast<uintptr_t>(A) & 15u) == 0 && (reinterpret_cast<uintptr_t>(B) & 15u) == 0;
  if (vec) L(false, true, true);
  else if (!ta && !tb) L(false, false, false);
  else if (!ta && tb) L(false, true, false);
  else if (ta && !tb) L(true, false, false);
  else L(true, true, false);
#undef L
  return hipGetLastError();
}

hipError_t launch_features(int W, const uint64_t* sets, const uint64_t* images, uint64_t count, uint32_t n, uint32_t k,
    uint32_t D, const FeatureGuide& g, uint16_t* x, hipStream_t st) {
  if (!count) return hipSuccess;
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_features<w>, dim3(unsigned((count + 3) / 4)), dim3(256), 0, st, sets, images, count, n, k, D, g, x); break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

namespace {

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t count,
    uint32_t* __restrict__ out) {
  const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < count) out[j] = src[idx[j]];
}

__global__ void k_neg_popcount(const uint64_t* __restrict__ sets, uint64_t count, uint32_t W, float* __restrict__ y) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  int c = 0;
  for (uint32_t w = 0; w < W; ++w) c += __popcll(sets[i * W + w]);
  y[i] = -float(c);
}

}  // namespace

hipError_t launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint64_t count, uint32_t* out,
    hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_gather_u32, dim3(blocks(count)), dim3(256), 0, st, src, idx, count, out);
  return hipGetLastError();
}

hipError_t launch_neg_popcount(const uint64_t* sets, uint64_t count, uint32_t W, float* y, hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_neg_popcount, dim3(blocks(count)), dim3(256), 0, st, sets, count, W, y);
  return hipGetLastError();
}

// A beam depth of up to kBeamRankMax candidates is ordered by counting
// (beam_select_reference, beam_select_scores) rather than sorted.
constexpr uint32_t kRankChunk = 256;
constexpr uint64_t kBeamRankMax = 16384;

// Stable ascending order of float scores, by counting (as k_beam_rank):
// rank[i] = #{j : key j below key i, or equal with j < i}.  Keys compare as
// the radix sort orders them: the float bits mapped to an unsigned order.
__device__ __forceinline__ uint32_t float_order(float f) {
  const uint32_t u = __float_as_uint(f);
  return u ^ ((u >> 31) ? 0xffffffffu : 0x80000000u);
}

// Grid of (64 items i) x (kRankChunk items j): the j chunk is staged in LDS,
// each of the four waves counts a quarter of it for the same 64 items, and
// one atomic add per item and workgroup adds the four partial counts.
__global__ __launch_bounds__(256) void k_score_rank(const float* __restrict__ keys, uint32_t count,
    uint32_t* __restrict__ rank) {
  __shared__ uint32_t jk[kRankChunk];
  __shared__ uint32_t part[4][64];
  const uint32_t j0 = blockIdx.y * kRankChunk;
  const uint32_t jn = count - j0 < kRankChunk ? count - j0 : kRankChunk;
  if (threadIdx.x < jn) jk[threadIdx.x] = float_order(keys[j0 + threadIdx.x]);
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * 64 + lane;
  uint32_t before = 0;
  if (i < count) {
    const uint32_t ki = float_order(keys[i]);
    const uint32_t tb = wv * (kRankChunk / 4), te = min(jn, tb + kRankChunk / 4);
    for (uint32_t t = tb; t < te; ++t) {
      const uint32_t kj = jk[t];
      before += (kj < ki || (kj == ki && j0 + t < i)) ? 1u : 0u;
    }
  }
  part[wv][lane] = before;
  __syncthreads();
  if (wv == 0 && i < count) {
    const uint32_t tot = part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
    if (tot) atomicAdd(&rank[i], tot);
  }
}


// Up to kBeamRankMax sets, the place of each set in that order is counted
// directly: rank[i] = #{j : set j comes before set i}, over a grid of (64
// sets i) x (kRankChunk sets j, staged in LDS with their keys), as
// k_score_rank.  A clear, the count and k_rank_select, where the merge sort
// and two gathers take seven launches or more; past kBeamRankMax the
// quadratic count is left to the sort (beam_reference_order).

// Sort key of a set: (n - |S|) in the top 10 bits (n <= 512, so the empty
// set's 512 still fits), then the top 54 bits of word 0; sets with equal keys
// are ordered by all their words.
template <int W>
__device__ __forceinline__ uint64_t beam_key(const uint64_t (&x)[W], uint32_t n) {
  uint32_t c = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) c += uint32_t(__popcll(x[w]));
  return (uint64_t(n - c) << 54) | (x[0] >> 10);
}

#ifndef SW_BEAM_RANK_J
#define SW_BEAM_RANK_J 256
#endif
constexpr uint32_t kBeamRankJ = SW_BEAM_RANK_J;  // sets j staged per workgroup

template <int W>
__global__ __launch_bounds__(256) void k_beam_rank(const uint64_t* __restrict__ sets, uint32_t count, uint32_t n,
    uint32_t* __restrict__ rank) {
  __shared__ uint64_t js[kBeamRankJ * W];
  __shared__ uint64_t jkey[kBeamRankJ];
  __shared__ uint32_t part[4][64];
  const uint32_t j0 = blockIdx.y * kBeamRankJ;
  const uint32_t jn = count - j0 < kBeamRankJ ? count - j0 : kBeamRankJ;
  for (uint32_t t = threadIdx.x; t < jn; t += blockDim.x) {
    uint64_t x[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      x[w] = sets[uint64_t(j0 + t) * W + w];
      js[t * W + w] = x[w];
    }
    jkey[t] = beam_key<W>(x, n);
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * 64 + lane;
  uint32_t before = 0;
  if (i < count) {
    uint64_t xi[W];
#pragma unroll
    for (int w = 0; w < W; ++w) xi[w] = sets[uint64_t(i) * W + w];
    const uint64_t ki = beam_key<W>(xi, n);
    const uint32_t tb = wv * (kBeamRankJ / 4), te = min(jn, tb + kBeamRankJ / 4);
    for (uint32_t t = tb; t < te; ++t) {
      const uint64_t kj = jkey[t];
      if (kj != ki) {
        before += kj < ki ? 1u : 0u;  // larger first, then the high bits of word 0
        continue;
      }
#pragma unroll
      for (int w = 0; w < W; ++w) {  // then word 0, word 1, ... ascending
        const uint64_t y = js[t * W + w];
        if (y != xi[w]) {
          before += y < xi[w] ? 1u : 0u;
          break;
        }
      }
    }
  }
  part[wv][lane] = before;
  __syncthreads();
  if (wv == 0 && i < count) {
    const uint32_t tot = part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
    if (tot) atomicAdd(&rank[i], tot);
  }
}

// The kept sets of a depth straight from their places: out_rows[r] =
// sets[i] and out_idx[r] = idx[i] for the set i of place r < keep (the
// scatter and both gathers of the sorted path in one launch).
template <int W>
__global__ __launch_bounds__(256) void k_rank_select(const uint32_t* __restrict__ rank,
    const uint64_t* __restrict__ sets, const uint32_t* __restrict__ idx, uint32_t count, uint32_t keep,
    uint64_t* __restrict__ out_rows, uint32_t* __restrict__ out_idx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t r = rank[i];
  if (r >= keep) return;
#pragma unroll
  for (int w = 0; w < W; ++w) out_rows[uint64_t(r) * W + w] = sets[uint64_t(i) * W + w];
  out_idx[r] = idx[i];
}

static hipError_t rank_select(const uint32_t* rank, const uint64_t* sets, const uint32_t* idx, uint64_t count,
    uint32_t W, uint64_t keep, uint64_t* out_rows, uint32_t* out_idx, hipStream_t st) {
  switch (W) {
#define CASE(w) \
  case w: hipLaunchKernelGGL(k_rank_select<w>, dim3(blocks(count)), dim3(256), 0, st, rank, sets, idx, \
      uint32_t(count), uint32_t(keep), out_rows, out_idx); return hipGetLastError();
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
  }
  return hipErrorInvalidValue;
}

bool beam_rank_fits(uint64_t count) { return count <= kBeamRankMax; }

hipError_t beam_select_reference(const uint64_t* sets, const uint32_t* idx, uint64_t count, uint32_t W, uint32_t n,
    uint64_t keep, uint32_t* rank, uint64_t* out_rows, uint32_t* out_idx, hipStream_t st) {
  if (!count) return hipSuccess;
  if (count > kBeamRankMax || W < 1 || W > 8) return hipErrorInvalidValue;
  hipError_t e = launch_fill_u32(rank, 0u, count, st);
  if (e != hipSuccess) return e;
  const dim3 grid(unsigned((count + 63) / 64), unsigned((count + kBeamRankJ - 1) / kBeamRankJ));
  switch (W) {
#define CASE(w) \
  case w: hipLaunchKernelGGL(k_beam_rank<w>, grid, dim3(256), 0, st, sets, uint32_t(count), n, rank); break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
  }
  return rank_select(rank, sets, idx, count, W, keep, out_rows, out_idx, st);
}

hipError_t beam_select_scores(const float* scores, const uint64_t* sets, const uint32_t* idx, uint64_t count,
    uint32_t W, uint64_t keep, uint32_t* rank, uint64_t* out_rows, uint32_t* out_idx, hipStream_t st) {
  if (!count) return hipSuccess;
  if (count > kBeamRankMax) return hipErrorInvalidValue;
  hipError_t e = launch_fill_u32(rank, 0u, count, st);
  if (e != hipSuccess) return e;
  const dim3 grid(unsigned((count + 63) / 64), unsigned((count + kRankChunk - 1) / kRankChunk));
  hipLaunchKernelGGL(k_score_rank, grid, dim3(256), 0, st, scores, uint32_t(count), rank);
  return rank_select(rank, sets, idx, count, W, keep, out_rows, out_idx, st);
}

// A depth's hash table cleared and its full-set index reset, in one launch.
__global__ void k_beam_prep(uint32_t* __restrict__ table, uint64_t slots, unsigned long long* __restrict__ first) {
  const uint64_t i0 = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i0 == 0) *first = ~0ull;
  for (uint64_t i = i0; i < slots; i += uint64_t(gridDim.x) * blockDim.x) table[i] = kEmpty;
}

hipError_t launch_beam_prep(uint32_t* table, uint64_t slots, unsigned long long* first, hipStream_t st) {
  const uint64_t b = std::min<uint64_t>(blocks(slots), 1024);
  hipLaunchKernelGGL(k_beam_prep, dim3(unsigned(b ? b : 1)), dim3(256), 0, st, table, slots, first);
  return hipGetLastError();
}

// Past the counted maximum a depth keeps its first `keep` sets by selection
// instead of a sort: which sets a depth keeps does not depend on their
// stored order, so one workgroup finds them by a most-significant-digit
// radix selection over the order's key, 8 bits a pass, and the kept sets
// are then compacted in candidate order.  The key of a set in the
// reference's order is the bit string (n - |S|) in 10 bits, word 0, word 1,
// ... (64 bits each), read as 64-bit chunks: chunk 0 = (n - |S|) << 54 |
// word 0 >> 10, chunk c = word c-1 << 54 | word c >> 10, chunk W = word W-1
// << 54.  Distinct sets have distinct keys.  Scores (the learned order) have
// the one chunk float_order(score) << 32 | index (ascending, stable).
// Each pass histograms the digit of the items still undecided, finds the
// digit d where the `need` first end, keeps those below d, drops those above
// and continues with d's bucket, or stops when all of d's bucket is needed.
// state[i]: 0 undecided, 1 kept, 2 dropped; keep[i] = state != 2 at the end.
template <int W, bool SCORES>
__device__ __forceinline__ uint64_t select_chunk(const uint64_t* __restrict__ sets, const float* __restrict__ scores,
    uint32_t i, uint32_t n, uint32_t c) {
  if (SCORES) return (uint64_t(float_order(scores[i])) << 32) | i;
  const uint64_t* x = sets + uint64_t(i) * W;
  if (c == 0) {
    uint32_t pc = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) pc += uint32_t(__popcll(x[w]));
    return (uint64_t(n - pc) << 54) | (x[0] >> 10);
  }
  return (x[c - 1] << 54) | (c < uint32_t(W) ? (x[c] >> 10) : 0ull);
}

constexpr uint32_t kSelectThreads = 1024;
constexpr uint64_t kSelectOneWorkgroup = 1u << 17;  // past this many items, the device-wide passes

template <int W, bool SCORES>
__global__ __launch_bounds__(kSelectThreads) void k_select_first(const uint64_t* __restrict__ sets,
    const float* __restrict__ scores, uint32_t count, uint32_t n, uint32_t keep, uint8_t* __restrict__ state,
    uint32_t* __restrict__ keep_flag) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t s_need, s_digit, s_done;
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < count; i += kSelectThreads) state[i] = 0;
  if (tid == 0) {
    s_need = keep;
    s_done = keep >= count ? 1u : 0u;
  }
  __syncthreads();
  const uint32_t chunks = SCORES ? 1u : uint32_t(W) + 1u;
  for (uint32_t c = 0; c < chunks && !s_done; ++c) {
    for (int dp = 7; dp >= 0 && !s_done; --dp) {
      const uint32_t shift = uint32_t(dp) * 8u;
      for (uint32_t b = tid; b < 256; b += kSelectThreads) hist[b] = 0;
      __syncthreads();
      for (uint32_t i = tid; i < count; i += kSelectThreads)
        if (state[i] == 0) atomicAdd(&hist[(select_chunk<W, SCORES>(sets, scores, i, n, c) >> shift) & 255u], 1u);
      __syncthreads();
      if (tid == 0) {
        // The undecided items number at least s_need, so some digit ends the need.
        uint32_t below = 0, d = 0;
        for (; d < 255; ++d) {
          if (below + hist[d] >= s_need) break;
          below += hist[d];
        }
        s_digit = d;
        s_need -= below;
        s_done = hist[d] == s_need ? 1u : 0u;
      }
      __syncthreads();
      const uint32_t ds = s_digit;
      for (uint32_t i = tid; i < count; i += kSelectThreads) {
        if (state[i] != 0) continue;
        const uint32_t dg = uint32_t(select_chunk<W, SCORES>(sets, scores, i, n, c) >> shift) & 255u;
        if (dg != ds) state[i] = dg < ds ? 1 : 2;
      }
      __syncthreads();
    }
  }
  for (uint32_t i = tid; i < count; i += kSelectThreads) keep_flag[i] = state[i] != 2 ? 1u : 0u;
}

// The same selection over the whole device, for lists too long for one
// workgroup (a wide Beam at large n: the plan's second run reaches widths of
// 10^5 and more at n = 512, k = 2).  Each pass is three launches: a
// histogram of the undecided items' digit (LDS histograms added into
// ctl.hist), the decision (one workgroup: the boundary digit, the need left,
// done when the whole bucket is needed), and the marking of the items
// decided by it.  A launch finds the pass done or not yet decided in ctl and
// does nothing then, so the host enqueues a chunk's eight passes at once and
// reads `done` back once per chunk.
struct SelCtl {
  uint32_t need, digit, done, pass;
  uint32_t hist[256];
};

template <int W, bool SCORES>
__global__ __launch_bounds__(256) void k_sel_hist(const uint64_t* __restrict__ sets,
    const float* __restrict__ scores, uint32_t count, uint32_t n, uint32_t c, uint32_t shift,
    const uint8_t* __restrict__ state, SelCtl* ctl) {
  __shared__ uint32_t h[256];
  if (ctl->done) return;  // uniform: written by an earlier launch
  h[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256)
    if (state[i] == 0) atomicAdd(&h[(select_chunk<W, SCORES>(sets, scores, i, n, c) >> shift) & 255u], 1u);
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&ctl->hist[threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_sel_decide(SelCtl* ctl, uint32_t pass) {
  __shared__ uint32_t s_done;
  if (threadIdx.x == 0) {
    s_done = ctl->done;
    if (!s_done) {
      uint32_t below = 0, d = 0;
      for (; d < 255; ++d) {
        if (below + ctl->hist[d] >= ctl->need) break;
        below += ctl->hist[d];
      }
      ctl->digit = d;
      ctl->need -= below;
      ctl->done = ctl->hist[d] == ctl->need ? 1u : 0u;
      ctl->pass = pass;
    }
  }
  __syncthreads();
  if (!s_done) ctl->hist[threadIdx.x] = 0;
}

template <int W, bool SCORES>
__global__ __launch_bounds__(256) void k_sel_mark(const uint64_t* __restrict__ sets,
    const float* __restrict__ scores, uint32_t count, uint32_t n, uint32_t c, uint32_t shift,
    uint8_t* __restrict__ state, const SelCtl* ctl, uint32_t pass) {
  if (ctl->pass != pass) return;  // this pass made no decision
  const uint32_t ds = ctl->digit;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) {
    if (state[i] != 0) continue;
    const uint32_t dg = uint32_t(select_chunk<W, SCORES>(sets, scores, i, n, c) >> shift) & 255u;
    if (dg != ds) state[i] = dg < ds ? 1 : 2;
  }
}

__global__ __launch_bounds__(256) void k_sel_init(uint8_t* __restrict__ state, uint32_t count, SelCtl* ctl,
    uint32_t keep) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) state[i] = 0;
  if (blockIdx.x == 0) {
    ctl->hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
      ctl->need = keep;
      ctl->done = keep >= count ? 1u : 0u;
      ctl->digit = 0;
      ctl->pass = ~0u;
    }
  }
}

__global__ __launch_bounds__(256) void k_sel_flags(const uint8_t* __restrict__ state, uint32_t count,
    uint32_t* __restrict__ keep_flag) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256)
    keep_flag[i] = state[i] != 2 ? 1u : 0u;
}

template <int W, bool SCORES>
static hipError_t select_wide(const uint64_t* sets, const float* scores, uint32_t count, uint32_t n, uint32_t keep,
    uint8_t* state, uint32_t* keep_flag, void* ctl_mem, uint32_t* host_done, hipStream_t st) {
  SelCtl* ctl = static_cast<SelCtl*>(ctl_mem);
  const unsigned g = unsigned(std::min<uint64_t>(blocks(count), 2048));
  hipLaunchKernelGGL(k_sel_init, dim3(g), dim3(256), 0, st, state, count, ctl, keep);
  const uint32_t chunks = SCORES ? 1u : uint32_t(W) + 1u;
  uint32_t pass = 0;
  for (uint32_t c = 0; c < chunks; ++c) {
    for (int dp = 7; dp >= 0; --dp, ++pass) {
      const uint32_t shift = uint32_t(dp) * 8u;
      hipLaunchKernelGGL((k_sel_hist<W, SCORES>), dim3(g), dim3(256), 0, st, sets, scores, count, n, c, shift, state,
          ctl);
      hipLaunchKernelGGL(k_sel_decide, dim3(1), dim3(256), 0, st, ctl, pass);
      hipLaunchKernelGGL((k_sel_mark<W, SCORES>), dim3(g), dim3(256), 0, st, sets, scores, count, n, c, shift, state,
          ctl, pass);
    }
    hipError_t e = hipMemcpyAsync(host_done, &ctl->done, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (*host_done) break;
  }
  hipLaunchKernelGGL(k_sel_flags, dim3(g), dim3(256), 0, st, state, count, keep_flag);
  return hipGetLastError();
}

size_t beam_select_ctl_bytes() { return sizeof(SelCtl); }

hipError_t beam_select_first(const uint64_t* sets, const float* scores, uint64_t count, uint32_t W, uint32_t n,
    uint64_t keep, uint8_t* state, uint32_t* keep_flag, void* ctl, uint32_t* host_done, hipStream_t st) {
  if (!count) return hipSuccess;
  if (count > 0xffffffffull || W < 1 || W > 8) return hipErrorInvalidValue;
  const dim3 g(1), b(kSelectThreads);
  const uint32_t c32 = uint32_t(count), k32 = uint32_t(keep < count ? keep : count);
  if (count > kSelectOneWorkgroup) {
    if (!ctl || !host_done) return hipErrorInvalidValue;
    if (scores) return select_wide<1, true>(sets, scores, c32, n, k32, state, keep_flag, ctl, host_done, st);
    switch (W) {
#define CASE(w) case w: return select_wide<w, false>(sets, scores, c32, n, k32, state, keep_flag, ctl, host_done, st);
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    }
    return hipErrorInvalidValue;
  }
  if (scores) {
    hipLaunchKernelGGL((k_select_first<1, true>), g, b, 0, st, sets, scores, c32, n, k32, state, keep_flag);
    return hipGetLastError();
  }
  switch (W) {
#define CASE(w) \
  case w: hipLaunchKernelGGL((k_select_first<w, false>), g, b, 0, st, sets, scores, c32, n, k32, state, keep_flag); break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
  }
  return hipGetLastError();
}

hipError_t mlp_forward(const MlpDev& p, const uint16_t* x, uint64_t B, MlpWork& w, float* y, hipStream_t st) {
  hipError_t e;
  if ((e = launch_gemm(false, true, x, p.D, p.w1b, p.D, w.h1f, w.h1b, p.H, int(B), p.H, p.D, p.b1, 1, 0, st))) return e;
  if ((e = launch_gemm(false, true, w.h1b, p.H, p.w2b, p.H, w.h2f, w.h2b, p.H, int(B), p.H, p.H, p.b2, 1, 0, st)))
    return e;
  hipLaunchKernelGGL(k_head, dim3(unsigned((B + 3) / 4)), dim3(256), 0, st, w.h2f, p.w3, p.b3, B, p.H, y);
  return hipGetLastError();
}

hipError_t mlp_backward(const MlpDev& p, const uint16_t* x, uint64_t B, const float* target, MlpWork& w, MlpGrads& g,
    float* loss, uint32_t rank_group, hipStream_t st) {
  hipError_t e;
  const float inv_b = 1.f / float(B);
  if (rank_group) {
    if (rank_group < 2 || rank_group > kRankMaxGroup || B % rank_group) return hipErrorInvalidValue;
    const uint64_t groups = B / rank_group;
    const float inv_pairs = 1.f / (float(groups) * float(rank_group * (rank_group - 1) / 2));
    hipLaunchKernelGGL(k_rank_dy, dim3(blocks(groups)), dim3(256), 0, st, w.y, target, groups, rank_group, inv_pairs,
        w.dy, loss);
  }
  hipLaunchKernelGGL(k_head_back, dim3(unsigned(B)), dim3(256), 0, st, w.y, target, w.h2f, p.w3, B, p.H, inv_b,
      rank_group ? 1 : 0, w.dy, w.g2b, loss);
  hipLaunchKernelGGL(k_head_wgrad, dim3(blocks(p.H)), dim3(256), 0, st, w.dy, w.h2f, B, p.H, g.w3, g.b3);
  // dW2 = dH2^T H1, db2 = colsum dH2
  if ((e = launch_gemm(true, false, w.g2b, p.H, w.h1b, p.H, g.w2, nullptr, p.H, p.H, p.H, int(B), nullptr, 0, 0, st)))
    return e;
  hipLaunchKernelGGL(k_colsum_bf16, dim3(blocks(p.H)), dim3(256), 0, st, w.g2b, B, p.H, g.b2);
  // dH1 = (dH2 W2) masked by ReLU(H1)
  if ((e = launch_gemm(false, false, w.g2b, p.H, p.w2b, p.H, w.g1f, w.g1b, p.H, int(B), p.H, p.H, nullptr, 0, 0, st)))
    return e;
  hipLaunchKernelGGL(k_relu_mask, dim3(blocks(B * p.H)), dim3(256), 0, st, w.g1b, w.h1f, B, p.H, w.g1b);
  hipLaunchKernelGGL(k_colsum_bf16, dim3(blocks(p.H)), dim3(256), 0, st, w.g1b, B, p.H, g.b1);
  // dW1 = dH1^T X
  if ((e = launch_gemm(true, false, w.g1b, p.H, x, p.D, g.w1, nullptr, p.D, p.H, p.D, int(B), nullptr, 0, 0, st)))
    return e;
  return hipGetLastError();
}

hipError_t mlp_adam(MlpDev& p, MlpGrads& g, MlpAdam& a, float lr, hipStream_t st) {
  a.t += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, float(a.t)), c2 = 1.f - powf(b2, float(a.t));
  struct T {
    float* w;
    float* g;
    float* m;
    float* v;
    uint64_t n;
    uint16_t* wb;
  } ts[] = {
      {p.w1, g.w1, a.m + a.off[0], a.v + a.off[0], uint64_t(p.H) * p.D, p.w1b},
      {p.b1, g.b1, a.m + a.off[1], a.v + a.off[1], uint64_t(p.H), nullptr},
      {p.w2, g.w2, a.m + a.off[2], a.v + a.off[2], uint64_t(p.H) * p.H, p.w2b},
      {p.b2, g.b2, a.m + a.off[3], a.v + a.off[3], uint64_t(p.H), nullptr},
      {p.w3, g.w3, a.m + a.off[4], a.v + a.off[4], uint64_t(p.H), nullptr},
      {p.b3, g.b3, a.m + a.off[5], a.v + a.off[5], 1, nullptr},
  };
  for (auto& t : ts)
    hipLaunchKernelGGL(k_adam, dim3(blocks(t.n)), dim3(256), 0, st, t.w, t.g, t.m, t.v, t.n, lr, b1, b2, eps, c1, c2,
        t.wb);
  return hipGetLastError();
}

hipError_t mlp_refresh_bf16(MlpDev& p, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(uint64_t(p.H) * p.D)), dim3(256), 0, st, p.w1, uint64_t(p.H) * p.D, p.w1b);
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(uint64_t(p.H) * p.H)), dim3(256), 0, st, p.w2, uint64_t(p.H) * p.H, p.w2b);
  return hipGetLastError();
}

}  // namespace sw
