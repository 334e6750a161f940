// One library context per host thread, and the status check, for the
// SynchroLib-side adapters in this directory (sw_seams.cpp, exact_gpu.hpp).
// Errors surface as exceptions, which the reference's JIT driver already
// catches (app/jit.hpp:61-77), in place of the exit(EXIT_FAILURE) of its CUDA
// helper (external/cuda/helper_cuda.h).
#pragma once
#include <synchrowords_amd.h>

#include <memory>
#include <stdexcept>
#include <string>

namespace synchrowords_amd {

inline sw_ctx* context() {
  thread_local std::unique_ptr<sw_ctx, void (*)(sw_ctx*)> ctx{nullptr, sw_ctx_destroy};
  if (!ctx) {
    sw_ctx* c = nullptr;
    if (sw_ctx_create(0, &c) != SW_OK) throw std::runtime_error(std::string("sw_ctx_create: ") + sw_last_error());
    ctx.reset(c);
  }
  return ctx.get();
}

inline void check(int rc) {
  if (rc != SW_OK) throw std::runtime_error(sw_last_error());
}

}  // namespace synchrowords_amd
