// C ABI of the evoc_hip library: argument checking, device context, and the stage
// entry points. The pipeline orchestration (evoc_hip_run*) lives in pipeline.cpp.
#include "api_internal.h"

#include <cstring>
#include <map>
#include <memory>

namespace evoc {

namespace {
thread_local std::string g_err;
std::mutex g_call_mu;
std::mutex g_ctx_mu;
std::map<int, std::unique_ptr<Ctx>> g_ctx;
}  // namespace

Ctx& context(int device) {
  std::lock_guard<std::mutex> g(g_ctx_mu);
  auto it = g_ctx.find(device);
  if (it != g_ctx.end()) {
    EVOC_HIP(hipSetDevice(device));
    return *it->second;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    throw NoDevice("no HIP device visible");
  }
  EVOC_CHECK(device >= 0 && device < count, "device_id out of range");
  EVOC_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  EVOC_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    throw NoDevice(std::string("evoc_hip is built for gfx950, device is ") + prop.gcnArchName);
  auto c = std::make_unique<Ctx>();
  c->device = device;
  EVOC_HIP(hipStreamCreate(&c->stream));
  EVOC_HIP(hipStreamCreate(&c->aux));
  Ctx& ref = *c;
  g_ctx[device] = std::move(c);
  return ref;
}

void release_all() {
  std::lock_guard<std::mutex> g(g_ctx_mu);
  for (auto& kv : g_ctx) {
    (void)hipSetDevice(kv.first);
    (void)hipStreamSynchronize(kv.second->stream);
    (void)hipStreamSynchronize(kv.second->aux);
    kv.second->pool.release();
  }
}

int guarded(const std::function<void()>& f) {
  std::lock_guard<std::mutex> lk(g_call_mu);
  try {
    f();
    g_err.clear();
    return EVOC_OK;
  } catch (const NoDevice& e) {
    g_err = e.what();
    return EVOC_ERR_NO_DEVICE;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return EVOC_ERR_INVALID_ARGUMENT;
  } catch (const HipError& e) {
    g_err = e.what();
    return EVOC_ERR_HIP;
  } catch (const std::exception& e) {
    g_err = e.what();
    return EVOC_ERR_INTERNAL;
  } catch (...) {
    g_err = "unknown error";
    return EVOC_ERR_INTERNAL;
  }
}

const char* last_error() { return g_err.c_str(); }

}  // namespace evoc

using namespace evoc;

extern "C" {

const char* evoc_hip_last_error(void) { return last_error(); }

void evoc_hip_default_params(evoc_params_t* p) {
  if (!p) return;
  p->n_neighbours = 15;
  p->noise_level = 0.5f;
  p->n_epochs = 50;
  p->embedding_dim = 0;
  p->neighbour_scale = 1.0f;
  p->symmetrise = 1;
  p->min_samples = 5;
  p->base_min_cluster_size = 5;
  p->approx_n_clusters = 0;
  p->min_similarity_threshold = 0.2;
  p->max_layers = 10;
}

void evoc_hip_default_nn_params(evoc_hip_nn_params_t* p) {
  if (!p) return;
  p->metric = EVOC_METRIC_EUCLIDEAN;
  p->knn_precision = EVOC_KNN_AUTO;
  p->warm_start = 0;
  p->warm_epochs = 0;
  p->mlp_subsample = 0;
  p->device_id = 0;
}

int evoc_hip_device_count(int32_t* count) {
  return guarded([&] {
    EVOC_CHECK(count, "count is NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
      (void)hipGetLastError();
      c = 0;
    }
    *count = c;
  });
}

int evoc_hip_release(void) {
  return guarded([&] { release_all(); });
}

int evoc_hip_knn(const float* x, int64_t n, int64_t d, int32_t metric, int32_t k,
                 int32_t precision, int32_t* idx_out, float* dist_out, int64_t* fallback_rows,
                 double* kernel_ms) {
  return guarded([&] {
    EVOC_CHECK(x && idx_out && dist_out, "knn: NULL buffer");
    EVOC_CHECK(n >= 1 && d >= 1, "knn: n and d must be >= 1");
    EVOC_CHECK(metric == 0 || metric == 1, "knn: unknown metric");
    EVOC_CHECK(k >= 1 && k <= 64, "knn: k must be in [1, 64]");
    check_finite_host(x, n * d, "input");
    Ctx& c = context(0);
    Buf<float> dx(c.pool, n * d);
    Buf<int32_t> di(c.pool, n * k);
    Buf<float> dd(c.pool, n * k);
    EVOC_HIP(hipMemcpyAsync(dx.p, x, n * d * sizeof(float), hipMemcpyHostToDevice, c.stream));
    KnnStats st;
    knn_device(dx.p, n, (int)d, d, metric, k, precision, di.p, dd.p, &st, c.stream, c.pool);
    EVOC_HIP(hipMemcpyAsync(idx_out, di.p, n * k * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    EVOC_HIP(hipMemcpyAsync(dist_out, dd.p, n * k * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    EVOC_HIP(hipStreamSynchronize(c.stream));
    if (fallback_rows) *fallback_rows = st.fallback_rows;
    if (kernel_ms) *kernel_ms = st.mfma_ms;
  });
}

int evoc_hip_knn_tile_pass(const float* x, int64_t n, int64_t d, int32_t metric, int32_t k,
                           int32_t precision, int32_t* cand_idx, float* cand_key,
                           float* key_scale, int32_t* ncand, float* gamma) {
  return guarded([&] {
    EVOC_CHECK(x && cand_idx && cand_key && key_scale, "knn_tile_pass: NULL buffer");
    EVOC_CHECK(n >= 1 && d >= 1, "knn_tile_pass: n and d must be >= 1");
    EVOC_CHECK(metric == 0 || metric == 1, "knn_tile_pass: unknown metric");
    EVOC_CHECK(k >= 1 && k <= 16, "knn_tile_pass: k must be in [1, 16]");
    EVOC_CHECK(precision == EVOC_KNN_SPLIT || precision == EVOC_KNN_FP8,
               "knn_tile_pass: precision must be EVOC_KNN_SPLIT or EVOC_KNN_FP8");
    check_finite_host(x, n * d, "input");
    Ctx& c = context(0);
    Buf<float> dx(c.pool, n * d);
    Buf<int32_t> di(c.pool, n * k);
    Buf<float> dd(c.pool, n * k);
    EVOC_HIP(hipMemcpyAsync(dx.p, x, n * d * sizeof(float), hipMemcpyHostToDevice, c.stream));
    KnnStats st;
    KnnCapture cap;
    knn_device(dx.p, n, (int)d, d, metric, k, precision, di.p, dd.p, &st, c.stream, c.pool, 0,
               &cap);
    EVOC_CHECK(cap.C == 32 && (int64_t)cap.idx.size() == n * 32, "knn_tile_pass: no capture");
    std::memcpy(cand_idx, cap.idx.data(), cap.idx.size() * sizeof(int32_t));
    std::memcpy(cand_key, cap.key.data(), cap.key.size() * sizeof(float));
    std::memcpy(key_scale, cap.scale.data(), n * sizeof(float));
    if (ncand) *ncand = cap.C;
    if (gamma) *gamma = cap.gamma;
  });
}

int evoc_hip_knn_last_stats(double* kernel_flop, double* kernel_ms, int32_t* symmetric,
                            int64_t* escalated_rows) {
  return guarded([&] {
    const KnnStats& st = last_knn_stats();
    if (kernel_flop) *kernel_flop = st.flops;
    if (kernel_ms) *kernel_ms = st.mfma_ms;
    if (symmetric) *symmetric = st.symmetric;
    if (escalated_rows) *escalated_rows = st.escalated_rows;
  });
}

}  // extern "C"
