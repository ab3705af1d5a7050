// Learned warm start: a small MLP (input -> 256 -> 256 -> embedding dim, GELU) that
// predicts the initial layout from the input vectors.
#pragma once
#include <functional>

#include "../../include/evoc_hip.h"
#include "kernels.h"

namespace evoc {

struct MlpWarmStart {
  int64_t subsample = 50000;
  int train_epochs = 0;
  double train_ms = 0.0;
  double final_loss = 0.0;
  // the acceptance check on the training subsample
  bool accepted = false;
  double judge_ari = 0.0;        // ARI of the warm-started labels against the plain labels
  double judge_mst_ratio = 0.0;  // MST total weight, warm-started / plain
};

// Decides on the training subsample whether the learned layout may replace the plain one.
// Arguments: the subsample's fuzzy graph, its plain layout (label propagation + the full
// epoch count) and the MLP's predicted layout, all in device memory (S x dim).
using WarmJudge = std::function<bool(const Graph& g, const float* d_plain, const float* d_pred,
                                     int64_t S, MlpWarmStart& ws)>;

// Trains the MLP on a subsample's plain layout, asks `judge`, and only if it accepts writes
// the predicted layout of all n points to d_init. Returns whether it did.
bool warm_start_init(const float* dX, int64_t n, int d, int64_t ld, const evoc_params_t& p,
                     const evoc_hip_nn_params_t& nn, int dim, uint64_t seed, MlpWarmStart& ws,
                     const WarmJudge& judge, float* d_init, hipStream_t s, DevPool& pool,
                     int verbose);

// MSE loss, prediction and gradients of the MLP on one batch (weights given as host
// fp32 arrays, layer l: W[l] (out x in), b[l] (out)); for gradient checks.
void mlp_loss_and_grads(const float* x, const float* t, int rows, int D, int O,
                        const float* const* w, const float* const* b, float* const* gw,
                        float* const* gb, float* out_y, double* loss, hipStream_t s,
                        DevPool& pool);

}  // namespace evoc
