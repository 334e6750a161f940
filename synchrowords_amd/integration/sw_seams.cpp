// The reference's four CUDA entry points, defined on this library's C ABI.
//
// A SynchroLib GPU build (plan "gpu": true, app/jit.hpp:120-121) compiles
// the four files synchrolib/data_structures/cuda/*.cu with nvcc and links
// cudart (jit/makefile_jit_gpu).  This one file replaces them: it defines
// exactly the functions and class members their headers declare, compiled
// by g++ like the rest of the JIT build and linked against
// libsynchrowords_amd.so.  Every caller in SynchroLib stays as it is:
//   preprocessed_transition_kernel  <- PreprocessedTransition::apply
//                                      (preprocessed_transitions.cpp:58-70)
//   SubsetsImplicitTrieKernel       <- SubsetsImplicitTrie's leaf test
//                                      (subsets_implicit_trie.hpp:74-84, 170-185)
//   check_subset_intersection_kernel_impl, reset_cuda_device_impl
//                                   <- SubsetsChecker (subsets_checker.hpp:25-57)
//   count_bits_kernel               (declared, count_bits_kernel.hpp:9-10)
// The templates are instantiated for the JIT's AUT_N / AUT_K, as the .cu
// files instantiate them.
#include <synchrolib/data_structures/cuda/count_bits_kernel.hpp>
#include <synchrolib/data_structures/cuda/preprocessed_transition_kernel.hpp>
#include <synchrolib/data_structures/cuda/subsets_checker_kernel.hpp>
#include <synchrolib/data_structures/cuda/subsets_implicit_trie_kernel.hpp>

#include <cstring>
#include <vector>

#include "sw_context.hpp"

namespace synchrolib {

namespace {

template <uint N>
const uint64_t* words(const Subset<N>* s) {
  static_assert(sizeof(Subset<N>) == sizeof(uint64_t) * Subset<N>::buckets(), "Subset<N> is its words");
  return reinterpret_cast<const uint64_t*>(s);
}

}  // namespace

// count_bits_kernel.cu:46-66: acc[q] = the number of sets holding state q.
template <uint N>
void count_bits_kernel(const Subset<N>* subsets, int* acc, size_t cnt) {
  std::vector<uint64_t> a(N, 0);
  if (cnt) synchrowords_amd::check(sw_count_bits(synchrowords_amd::context(), N, words(subsets), cnt, a.data(), SW_HOST));
  for (uint q = 0; q < N; ++q) acc[q] = static_cast<int>(a[q]);
}

// preprocessed_transition_kernel.cu: to[i] = OR over 8-state slices of
// trans[slice][byte of from[i]].
template <uint N, uint K>
void preprocessed_transition_kernel(const PreprocessedTransition<N, K>& trans, const Subset<N>* from, Subset<N>* to,
    size_t cnt) {
  if (!cnt) return;
  synchrowords_amd::check(sw_transition_apply(synchrowords_amd::context(), N, words(&trans.trans[0][0]),
      words(from), reinterpret_cast<uint64_t*>(to), cnt, SW_HOST));
}

// subsets_checker_kernel.cu:40-62: results[i] = (#{j : small_i subset of
// large_j} >= thresh); `buckets` 64-bit words per set.
std::unique_ptr<bool[]> check_subset_intersection_kernel_impl(const uint64* small, const uint64* large, int small_cnt,
    int large_cnt, int buckets, int thresh) {
  static_assert(sizeof(bool) == 1, "bool results are bytes");
  auto results = std::make_unique<bool[]>(small_cnt > 0 ? small_cnt : 0);
  if (small_cnt <= 0) return results;
  std::vector<uint8_t> r(small_cnt, 0);
  // The checker's lists carry `buckets` words; any n that needs exactly as
  // many words tests the same bits.
  const uint32_t n = uint32_t(buckets) * 64;
  synchrowords_amd::check(sw_subset_check(synchrowords_amd::context(), n, small, uint64_t(small_cnt), large,
      uint64_t(large_cnt > 0 ? large_cnt : 0), uint32_t(thresh > 0 ? thresh : 0), SW_HOST, r.data(), nullptr,
      nullptr));
  for (int i = 0; i < small_cnt; ++i) results[i] = r[i] != 0;
  return results;
}

// subsets_checker_kernel.cu:64-66 resets the device; the library frees only
// its own buffers (sw_ctx_reset).
void reset_cuda_device_impl() { synchrowords_amd::check(sw_ctx_reset(synchrowords_amd::context())); }

// subsets_implicit_trie_kernel.cu:160-238.  The class's members are the
// reference's; its first device-pointer member carries the library's leaf
// handle (an opaque pointer, never dereferenced as a set).
template <uint N, bool Proper>
SubsetsImplicitTrieKernel<N, Proper>::SubsetsImplicitTrieKernel()
    : allocated(false), small_ptr(nullptr), large_ptr(nullptr), results(nullptr) {}

template <uint N, bool Proper>
SubsetsImplicitTrieKernel<N, Proper>::~SubsetsImplicitTrieKernel() {
  deallocate();
}

template <uint N, bool Proper>
void SubsetsImplicitTrieKernel<N, Proper>::allocate(const Subset<N>* small, uint small_cnt, uint large_cnt) {
  deallocate();
  sw_leaf* leaf = nullptr;
  synchrowords_amd::check(sw_leaf_create(synchrowords_amd::context(), N, small_cnt ? words(small) : nullptr,
      small_cnt, large_cnt, &leaf));
  small_ptr = reinterpret_cast<Subset<N>*>(leaf);
  allocated = true;
}

template <uint N, bool Proper>
void SubsetsImplicitTrieKernel<N, Proper>::run(uint small_pos, uint small_cnt, const Subset<N>* large,
    uint large_cnt, bool* ret) {
  static_assert(sizeof(bool) == 1, "bool results are bytes");
  if (!large_cnt) return;
  sw_leaf* leaf = reinterpret_cast<sw_leaf*>(small_ptr);
  synchrowords_amd::check(sw_leaf_run(leaf, Proper ? 1 : 0, small_pos, small_cnt, words(large), large_cnt,
      reinterpret_cast<uint8_t*>(ret)));
}

template <uint N, bool Proper>
void SubsetsImplicitTrieKernel<N, Proper>::deallocate() {
  if (!allocated) return;
  sw_leaf_destroy(reinterpret_cast<sw_leaf*>(small_ptr));
  small_ptr = nullptr;
  allocated = false;
}

template void count_bits_kernel<AUT_N>(const Subset<AUT_N>* subsets, int* acc, size_t cnt);
template void preprocessed_transition_kernel<AUT_N, AUT_K>(const PreprocessedTransition<AUT_N, AUT_K>& trans,
    const Subset<AUT_N>* from, Subset<AUT_N>* to, size_t cnt);
template class SubsetsImplicitTrieKernel<AUT_N, false>;
template class SubsetsImplicitTrieKernel<AUT_N, true>;

}  // namespace synchrolib
