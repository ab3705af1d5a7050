// Learned warm start (see mlp.h).
#include "mlp.h"

namespace evoc {

void warm_start_init(const float* dX, int64_t n, int d, int64_t ld, const Graph& g,
                     const evoc_params_t& p, const evoc_hip_nn_params_t& nn, int dim,
                     uint64_t seed, MlpWarmStart& ws, float* d_init, hipStream_t s,
                     DevPool& pool, int verbose) {
  throw std::invalid_argument("warm_start: learned initialisation not available in this build");
}

}  // namespace evoc
