// The learned scorer (k_mlp.hip, mlp.cpp): device parameter and work buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sw {

// Parameters on the device: FP32 masters and BF16 copies of the two hidden
// weight matrices.  W1 is [H][D], W2 is [H][H] (row = output unit).
struct MlpDev {
  uint32_t D = 0, H = 0;
  float *w1 = nullptr, *b1 = nullptr, *w2 = nullptr, *b2 = nullptr, *w3 = nullptr, *b3 = nullptr;
  uint16_t *w1b = nullptr, *w2b = nullptr;
};

// Activations of one batch.
struct MlpWork {
  float *h1f = nullptr, *h2f = nullptr, *y = nullptr, *dy = nullptr, *g1f = nullptr;
  uint16_t *h1b = nullptr, *h2b = nullptr, *g1b = nullptr, *g2b = nullptr;
};

struct MlpGrads {
  float *w1 = nullptr, *b1 = nullptr, *w2 = nullptr, *b2 = nullptr, *w3 = nullptr, *b3 = nullptr;
};

struct MlpAdam {
  float *m = nullptr, *v = nullptr;
  uint64_t off[6] = {0, 0, 0, 0, 0, 0};
  int t = 0;
};

// Per-automaton inputs of the feature kernel: the canonical order of the
// states and the images of Q up to depth 2*kGuideDepths.
constexpr int kGuideDepths = 4;
constexpr uint32_t kGuideFeatures = kGuideDepths + 1;  // min-missing counts + cardinality
struct FeatureGuide {
  const uint16_t* order = nullptr;  // order[r]: the state of feature column r (canonical order)
  const uint64_t* rows = nullptr;  // images of Q, by depth (W words each)
  uint32_t ends[kGuideDepths] = {0, 0, 0, 0};  // rows of depth <= 2, 4, 6, 8
};
constexpr uint32_t kRankMaxGroup = 16;
constexpr float kFeatureScale = 8.f;  // count features are kFeatureScale * count / n

hipError_t launch_gemm(bool ta, bool tb, const uint16_t* A, int lda, const uint16_t* B, int ldb, float* Cf,
    uint16_t* Cb, int ldc, int M, int N, int K, const float* bias, int relu, int accumulate, hipStream_t st);
hipError_t launch_features(int W, const uint64_t* sets, const uint64_t* images, uint64_t count, uint32_t n, uint32_t k,
    uint32_t D, const FeatureGuide& g, uint16_t* x, hipStream_t st);
hipError_t mlp_forward(const MlpDev& p, const uint16_t* x, uint64_t B, MlpWork& w, float* y, hipStream_t st);
// rank_group 0: mean squared error against target; G >= 2: pairwise ranking
// loss within consecutive groups of G rows (B a multiple of G).
hipError_t mlp_backward(const MlpDev& p, const uint16_t* x, uint64_t B, const float* target, MlpWork& w, MlpGrads& g,
    float* loss, uint32_t rank_group, hipStream_t st);
hipError_t mlp_adam(MlpDev& p, MlpGrads& g, MlpAdam& a, float lr, hipStream_t st);
hipError_t mlp_refresh_bf16(MlpDev& p, hipStream_t st);
// out[j] = src[idx[j]].
hipError_t launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint64_t count, uint32_t* out,
    hipStream_t st);
// y[i] = -|sets[i]| (the cardinality order as an ascending key).
hipError_t launch_neg_popcount(const uint64_t* sets, uint64_t count, uint32_t W, float* y, hipStream_t st);
// Which `keep` of `count` distinct sets come first, by selection (one
// workgroup, no sort): in the reference's Beam order (scores NULL: larger
// first, then word 0, word 1, ... ascending, beam.hpp's comparator) or in
// ascending stable order of float scores.  keep_flag[i] = 1 for the kept;
// state: count bytes of scratch.  Past 2^17 items the passes run over the
// whole device: ctl (beam_select_ctl_bytes() bytes of device memory) and
// host_done (one pinned word) are then required, and the call synchronizes
// the stream once per key chunk.
size_t beam_select_ctl_bytes();
hipError_t beam_select_first(const uint64_t* sets, const float* scores, uint64_t count, uint32_t W, uint32_t n,
    uint64_t keep, uint8_t* state, uint32_t* keep_flag, void* ctl, uint32_t* host_done, hipStream_t st);
// The first `keep` sets of a depth, by place, written to out_rows/out_idx
// (sets[i], idx[i] of the set at place r < keep), for count <= the counted
// maximum (beam_rank_fits): in the reference's Beam order, or in ascending
// stable order of float scores.  rank: count words of scratch.
bool beam_rank_fits(uint64_t count);
hipError_t beam_select_reference(const uint64_t* sets, const uint32_t* idx, uint64_t count, uint32_t W, uint32_t n,
    uint64_t keep, uint32_t* rank, uint64_t* out_rows, uint32_t* out_idx, hipStream_t st);
hipError_t beam_select_scores(const float* scores, const uint64_t* sets, const uint32_t* idx, uint64_t count,
    uint32_t W, uint64_t keep, uint32_t* rank, uint64_t* out_rows, uint32_t* out_idx, hipStream_t st);
// table[0 .. slots) = empty and *first = ~0 (the start of a beam depth).
hipError_t launch_beam_prep(uint32_t* table, uint64_t slots, unsigned long long* first, hipStream_t st);

}  // namespace sw
