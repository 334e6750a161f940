// Device subset trie: host driver of the k_trie.hip kernels (trie.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "sw_host.hpp"
#include "sw_kernels.hpp"

namespace sw {

// Growable pinned host buffer for the build's small uploads and read-backs.
struct PinnedBuf {
  void* p = nullptr;
  size_t cap = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  template <typename T>
  T* get(size_t count) {
    const size_t bytes = count * sizeof(T) + 64;
    if (bytes > cap) {
      if (p) SW_CHECK_HIP(hipHostFree(p));
      p = nullptr;
      size_t nc = cap ? cap : 4096;
      while (nc < bytes) nc *= 2;
      SW_CHECK_HIP(hipHostMalloc(&p, nc, hipHostMallocDefault));
      cap = nc;
    }
    return static_cast<T*>(p);
  }
};

// A subset trie over a list of sets, resident in HBM (the reference's
// SubsetsTrie<N,Threads>, synchrolib/data_structures/subsets_trie.hpp).
class DeviceTrie {
 public:
  uint32_t n = 0, W = 0;
  uint64_t m = 0;              // sets in the trie
  uint32_t node_count = 0;     // node ids used (some ids of a local range may stay unused)
  uint32_t top_levels = 0;     // level-synchronous passes of the last build
  DevBuf sets, tags, nodes;    // sets in trie order, their tags, the nodes

  // Build over `count` sets at src (device; complemented on load when
  // `complement`), tagged by tags_in[i] (NULL: the row index i).
  void build(hipStream_t st, uint32_t n, const uint64_t* src, uint64_t count, int complement,
      const uint32_t* tags_in);

  // Query `qcount` sets at q (device; complemented on load when
  // q_complement) against the trie.  superset: does some trie set contain
  // the query (else: lie inside it).  With `found` (ANY): stop at the first
  // hit, *found = 1 and *witness = min over hits of (query << 32 | tag) when
  // query_high, else (tag << 32 | query); found and witness must hold 0 and
  // ~0.  Without: flags[query] = 1 for every query with a hit (flags must
  // hold zeros).  sort: walk the queries in cardinality order.  Enqueued on
  // st; nothing is read back.
  void query(hipStream_t st, uint32_t grid, int superset, const uint64_t* q, uint64_t qcount, int q_complement,
      int sort, int query_high, int* found, unsigned long long* witness, uint8_t* flags);

 private:
  DevBuf buf1_, buf2_, tags1_, tags2_;
  DevBuf segs_, chunks_, hist_, segmin_, segmax_, bits_, ones_, czero_, coff_, jobs_, bigjobs_;
  DevBuf qhist_, qcursor_, qperm_, qnext_;
  PinnedBuf h_segs_, h_chunks_, h_back_, h_coff_, h_nodes_, h_jobs_;
};

}  // namespace sw
