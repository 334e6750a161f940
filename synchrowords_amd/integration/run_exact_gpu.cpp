// The JIT library entry `run` (the reference's jit/jitmain.cpp:11-49) with
// the Exact stage served by ExactGpu: the registration a SynchroLib build
// makes in make_algorithm (synchrolib/synchrolib.hpp:53-92), written here as
// the one branch of the plan loop that differs, so that the reference's own
// files stay as they are.  Same signature, same bounds, same per-algorithm
// timing and Reduce hand-back as the reference's loop.
#include <jitdefines.hpp>

#include <synchrolib/synchrolib.hpp>

#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "exact_gpu.hpp"

extern "C" {

using namespace synchrolib;

void run(const std::string& aut_encoded, const std::vector<std::string>& algorithms, AlgoResult& result,
    Logger::LogLevel log_level) {
  Timer timer("algorithms");
  Logger::set_log_level(log_level);
  AlgoData<AUT_N, AUT_K> data = AlgoData<AUT_N, AUT_K>(Automaton<AUT_N, AUT_K>::decode(aut_encoded));
  if (result.reduce && result.reduce->done) {
    data.result = result;
  } else {
    data.result.non_synchro = false;
    data.result.mlsw_lower_bound = 0;
    using BoundType = decltype(data.result.mlsw_upper_bound);
    data.result.mlsw_upper_bound = std::numeric_limits<BoundType>::max();
    if (UPPER_BOUND < data.result.mlsw_upper_bound) data.result.mlsw_upper_bound = static_cast<BoundType>(UPPER_BOUND);
  }
  while (data.result.algorithms_run.size() < algorithms.size()) {
    const std::string& name = algorithms[data.result.algorithms_run.size()];
    Timer algo_timer("algo" + name);
    if (name == "Exact") {
      synchrowords_amd::ExactGpu<AUT_N, AUT_K>().run(data);
    } else {
      make_algorithm<AUT_N, AUT_K>(name)->run(data);
    }
    data.result.algorithms_run.emplace_back(name, algo_timer.stop());
    if (data.result.reduce && !data.result.reduce->done) {
      result = data.result;
      return;
    }
  }
  result = data.result;
}

}  // extern "C"
