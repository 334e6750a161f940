// The Beam bound of an automaton of at most 64 states in one workgroup, every
// depth inside one launch (the reference's Beam, algorithm/beam/beam.hpp:98-188:
// preimages of the list under every letter, the first depth producing Q
// bounds the length, duplicates dropped, the `width` largest sets kept, ties
// by their word ascending).  Sets are one 64-bit word; the preimage masks, the
// list, the candidates and the dedup table live in LDS (about 156 KB).
// Depths of more than kBeamSmallCand candidates, or that keep more than
// kBeamSmallList sets, are handed back to the device-wide beam (mlp.cpp)
// with the list and the word history so far, which it continues.
//
// Same bounds as the device-wide beam in the reference's order: candidates
// are numbered letter-major (c = a m + i), among equal sets the smallest
// number survives, each depth keeps the same `width` distinct sets (chosen
// by selection, kept in candidate order), and the history keeps each kept
// set's candidate number.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sw_device.hpp"
#include "sw_kernels.hpp"

namespace sw {

namespace {

constexpr uint32_t kThreadsBS = 1024;
constexpr uint32_t kTab = 2 * kBeamSmallCand;  // dedup slots, at most half full
constexpr uint32_t kNone = 0xffffffffu;

__device__ __forceinline__ uint32_t mix32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  return uint32_t(x);
}

// Exclusive scan of 0/1 flags over the 1024 threads (ballots, two barriers).
__device__ __forceinline__ uint32_t flag_scan(bool v, uint32_t* wsum, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t b = __ballot(v);
  if (lane == 0) wsum[wv] = uint32_t(__popcll(b));
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (uint32_t w = 0; w < kThreadsBS / 64; ++w) {
    before += w < wv ? wsum[w] : 0u;
    tot += wsum[w];
  }
  *total = tot;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t r = before + uint32_t(__popcll(b & below));
  __syncthreads();
  return r;
}

}  // namespace

__global__ __launch_bounds__(kThreadsBS) void k_beam_small(BeamSmallArgs a) {
  __shared__ uint64_t pre[kBeamSmallMaxK][64];
  __shared__ uint64_t lists[2][kBeamSmallList];
  __shared__ uint64_t cand[kBeamSmallCand];
  __shared__ uint64_t uval[kBeamSmallCand];
  __shared__ uint32_t uidx[kBeamSmallCand];
  __shared__ uint32_t tab[kTab];
  __shared__ uint32_t wsum[kThreadsBS / 64];
  __shared__ uint32_t pch[65];  // sets per size, for the selection
  __shared__ uint32_t s_first, s_cls, s_cstar, s_need, s_halt;
  const uint32_t tid = threadIdx.x;
  const uint32_t n = a.n, k = a.k;
  const uint64_t full = n == 64 ? ~0ull : ((1ull << n) - 1);
  for (uint32_t i = tid; i < kBeamSmallMaxK * 64; i += kThreadsBS) (&pre[0][0])[i] = 0;
  for (uint32_t i = tid; i < a.m0; i += kThreadsBS) lists[0][i] = a.seed[i];
  __syncthreads();
  for (uint32_t i = tid; i < n * k; i += kThreadsBS) {
    const uint32_t q = i / k, l = i - q * k;
    atomicOr(reinterpret_cast<unsigned long long*>(&pre[l][a.delta[i]]), 1ull << q);
  }
  __syncthreads();
  uint32_t m = a.m0, cur = 0, hist_len = 0, len = 1;
  uint32_t status = kBeamSmallExhausted;
  for (; len < a.upper; ++len) {
    if (a.stop && (len & 15u) == 0) {
      // Every 16 depths: has the caller's other stream set a smaller cap?
      if (tid == 0) {
        const unsigned long long v = __hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        s_halt = (uint32_t(v >> 32) == a.stop_gen && len >= uint32_t(v)) ? 1u : 0u;
      }
      __syncthreads();
      if (s_halt) break;  // as if upper had been that depth: exhausted
    }
    const uint32_t cnt = m * k;
    if (cnt > kBeamSmallCand || hist_len + cnt > a.hist_cap || len > a.meta_cap) {
      status = kBeamSmallHandover;
      break;
    }
    const uint64_t* list = lists[cur];
    if (tid == 0) s_first = kNone;
    for (uint32_t i = tid; i < kTab; i += kThreadsBS) tab[i] = kNone;
    __syncthreads();
    // Preimages, letter-major; the first full set bounds the length.
    for (uint32_t c = tid; c < cnt; c += kThreadsBS) {
      const uint32_t l = c / m, i = c - l * m;
      uint64_t x = list[i], r = 0;
      while (x) {
        r |= pre[l][__builtin_ctzll(x)];
        x &= x - 1;
      }
      cand[c] = r;
      if (r == full) atomicMin(&s_first, c);
    }
    __syncthreads();
    if (s_first != kNone) {
      status = kBeamSmallFound;
      if (tid == 0) {
        // The word: this candidate's letter, then back through the kept
        // parents (their candidate numbers at each earlier depth).
        uint32_t c = s_first, mm = m, wl = 0;
        a.word[wl++] = c / mm;
        uint32_t parent = c % mm;
        for (uint32_t d = len - 1; d >= 1; --d) {
          const uint32_t ci = a.hist[a.meta[2 * (d - 1)] + parent];
          const uint32_t md = a.meta[2 * (d - 1) + 1];
          a.word[wl++] = ci / md;
          parent = ci % md;
        }
      }
      break;
    }
    // Dedup: the smallest candidate number of each distinct set holds its slot.
    for (uint32_t c = tid; c < cnt; c += kThreadsBS) {
      const uint64_t r = cand[c];
      uint32_t pos = mix32(r) & (kTab - 1);
      while (true) {
        uint32_t v = atomicCAS(&tab[pos], kNone, c);
        if (v == kNone) break;
        if (cand[v] == r) {
          atomicMin(&tab[pos], c);
          break;
        }
        pos = (pos + 1) & (kTab - 1);
      }
    }
    __syncthreads();
    // Survivors, in candidate order.
    uint32_t nu = 0;
    for (uint32_t c0 = 0; c0 < cnt; c0 += kThreadsBS) {
      const uint32_t c = c0 + tid;
      bool keep = false;
      if (c < cnt) {
        const uint64_t r = cand[c];
        uint32_t pos = mix32(r) & (kTab - 1);
        while (cand[tab[pos]] != r) pos = (pos + 1) & (kTab - 1);
        keep = tab[pos] == c;
      }
      uint32_t tot;
      const uint32_t at = flag_scan(keep, wsum, &tot);
      if (keep) {
        uval[nu + at] = cand[c];
        uidx[nu + at] = c;
      }
      nu += tot;
    }
    __syncthreads();
    // The `width` first distinct sets in the reference's order (larger
    // first, then word ascending), kept in candidate order: that order is
    // total on distinct sets, so which sets a depth keeps does not depend on
    // the order the list is stored in, and the bound does not either (the
    // word is any word of the bound's length).  All are kept when they fit;
    // otherwise the sets larger than the boundary size are kept, and of the
    // boundary size those with fewer than `need` smaller words in it.
    const uint32_t keepn = nu < a.width ? nu : a.width;
    if (keepn > kBeamSmallList) {  // this depth continues on the device-wide beam
      status = kBeamSmallHandover;
      break;
    }
    uint64_t* next = lists[cur ^ 1];
    if (nu <= a.width) {
      for (uint32_t u = tid; u < nu; u += kThreadsBS) {
        next[u] = uval[u];
        a.hist[hist_len + u] = uidx[u];
      }
    } else {
      for (uint32_t i = tid; i < 65; i += kThreadsBS) pch[i] = 0;
      if (tid == 0) s_cls = 0;
      __syncthreads();
      for (uint32_t u = tid; u < nu; u += kThreadsBS) atomicAdd(&pch[__popcll(uval[u])], 1u);
      __syncthreads();
      if (tid == 0) {
        uint32_t above = 0, c = 64;
        while (above + pch[c] < a.width) above += pch[c--];
        s_cstar = c;
        s_need = a.width - above;
      }
      __syncthreads();
      const uint32_t cs = s_cstar;
      // The boundary class, gathered (its order only feeds counts).
      for (uint32_t u = tid; u < nu; u += kThreadsBS)
        if (uint32_t(__popcll(uval[u])) == cs) cand[atomicAdd(&s_cls, 1u)] = uval[u];
      __syncthreads();
      const uint32_t ncls = s_cls, need = s_need;
      uint32_t kept = 0;
      for (uint32_t u0 = 0; u0 < nu; u0 += kThreadsBS) {
        const uint32_t u = u0 + tid;
        bool keep = false;
        if (u < nu) {
          const uint64_t x = uval[u];
          const uint32_t px = uint32_t(__popcll(x));
          if (px > cs) {
            keep = true;
          } else if (px == cs) {
            uint32_t below = 0;
            for (uint32_t j = 0; j < ncls; ++j) below += cand[j] < x ? 1u : 0u;
            keep = below < need;
          }
        }
        uint32_t tot;
        const uint32_t at = flag_scan(keep, wsum, &tot);
        if (keep) {
          next[kept + at] = uval[u];
          a.hist[hist_len + kept + at] = uidx[u];
        }
        kept += tot;
      }
    }
    if (tid == 0) {
      a.meta[2 * (len - 1)] = hist_len;
      a.meta[2 * (len - 1) + 1] = m;
    }
    hist_len += keepn;
    m = keepn;
    cur ^= 1;
    __syncthreads();
    if (m == 0) {
      status = kBeamSmallEmpty;
      break;
    }
  }
  // State for the caller: where the beam stopped and the list it holds.
  for (uint32_t i = tid; i < m; i += kThreadsBS) a.list_out[i] = lists[cur][i];
  if (tid == 0) {
    a.out[0] = status;
    a.out[1] = len;
    a.out[2] = m;
    a.out[3] = hist_len;
  }
}

hipError_t launch_beam_small(const BeamSmallArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_beam_small, dim3(1), dim3(kThreadsBS), 0, st, a);
  return hipGetLastError();
}

}  // namespace sw
