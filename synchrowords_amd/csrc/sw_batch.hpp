// Batched exact search (k_batch.hip): shared host/device declarations.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace sw {

constexpr uint32_t kBatchMaxK = 4;  // letters handled inside the persistent kernel
constexpr uint32_t kBatchMaxW = 3;  // words per set (n <= 192)
constexpr uint32_t kBatchThreads = 256;  // threads per workgroup of the persistent launch

enum BatchStatus : uint32_t {
  kBatchPending = 0,
  kBatchFound = 1,        // shortest word found by the search
  kBatchBound = 2,        // no word below `upper`: upper is the length
  kBatchOverflow = 3,     // outgrew the workgroup's scratch
  kBatchUnsupported = 4,  // n or k outside the kernel's range
};

struct BatchInstance {
  uint32_t n, k;
  uint32_t delta_off;  // into the byte array of transitions
  uint32_t upper;      // length of a known reset word (search below it)
  uint64_t b0[kBatchMaxW];  // sink states with two preimages under a letter
};

// Why an instance left the workgroup (BatchResult::detail for kBatchOverflow).
enum BatchOverflow : uint32_t {
  kOverflowNone = 0,
  kOverflowCand = 1,   // one step's candidates above cap_cand
  kOverflowRows = 2,   // a side above cap_rows
  kOverflowWord = 3,   // length reached the word buffer
  kOverflowHash = 4,   // hash table above hash_slots
  kOverflowPairs = 5,  // goal check above max_pairs
};

struct BatchResult {
  uint32_t status, length, word_len, steps, detail;
};

// Capacities of one workgroup's scratch.
struct BatchLimits {
  uint32_t cap_rows;    // sets per side
  uint32_t cap_cand;    // candidates of one step
  uint32_t hash_slots;  // largest hash table (power of two)
  uint64_t max_pairs;   // largest goal check inside a workgroup
};

size_t batch_scratch_bytes(int W, const BatchLimits& L);
hipError_t launch_batch(int W, const BatchInstance* inst, const uint8_t* deltas, const uint32_t* order,
    uint32_t count, uint32_t* head, BatchResult* out, uint8_t* words, uint32_t word_cap, uint8_t* scratch,
    size_t scratch_stride, const BatchLimits& L, uint32_t grid, hipStream_t st, uint32_t threads = kBatchThreads);

}  // namespace sw
