// Learned warm start for the node embedding: an MLP  D -> 256 -> 256 -> dim  (GELU),
// bf16 weights on v_mfma_f32_32x32x16_bf16 with fp32 accumulation, fp32 master
// weights and Adam moments in HBM. Trained on (input row, layout row) pairs that the
// plain pipeline produces on a subsample; its prediction for every point replaces the
// label-propagation initial layout, so the full embedding needs far fewer epochs.
//
// All GEMMs are C[M x N] = A[M x K] * B[N x K]^T with row-major bf16 A and B (the
// "NT" form); transposed operands are materialised by a transpose kernel.
#include <cstring>
#include <numeric>

#include "host_stages.h"
#include "mlp.h"

namespace evoc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int HID = 256;

__device__ __forceinline__ uint16_t f2b(float f) {
  uint32_t b = __float_as_uint(f);
  if ((b & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float b2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// C = A B^T (+ bias). One wavefront per 32x32 tile of C; M, N multiples of 32, K of 16.
// mode 0: C32 = acc + bias; mode 1: also Z32 = acc + bias and Cb = bf16(gelu(Z)).
__global__ __launch_bounds__(256) void gemm_nt_kernel(const uint16_t* __restrict__ A,
                                                      const uint16_t* __restrict__ B, int M, int N,
                                                      int K, const float* __restrict__ bias,
                                                      float* __restrict__ C32,
                                                      uint16_t* __restrict__ Cb, int mode) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tiles_n = N / 32;
  const int tile = blockIdx.x * 4 + wave;
  if (tile >= (M / 32) * tiles_n) return;
  const int m0 = (tile / tiles_n) * 32, n0 = (tile % tiles_n) * 32;
  const uint16_t* ar = A + (size_t)(m0 + (lane & 31)) * K + (lane >> 5) * 8;
  const uint16_t* br = B + (size_t)(n0 + (lane & 31)) * K + (lane >> 5) * 8;
  f32x16 acc = {};
  for (int k = 0; k < K; k += 16) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(ar + k);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(br + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  const int col = n0 + (lane & 31);
  const float bv = bias ? bias[col] : 0.0f;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = m0 + j * 8 + (lane >> 5) * 4 + i;
      const float z = acc[j * 4 + i] + bv;
      C32[(size_t)row * N + col] = z;
      if (mode == 1) Cb[(size_t)row * N + col] = f2b(gelu(z));
    }
}

// bf16 transpose: out[c][r] = in[r][c]
__global__ void transpose_kernel(const uint16_t* __restrict__ in, int R, int Cc,
                                 uint16_t* __restrict__ out) {
  __shared__ uint16_t t[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int r = by + j, c = bx + tx;
    t[j][tx] = (r < R && c < Cc) ? in[(size_t)r * Cc + c] : 0;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = bx + j, r = by + tx;
    if (c < Cc && r < R) out[(size_t)c * R + r] = t[tx][j];
  }
}

// rows of x (f32, leading dimension ld) gathered by idx (or contiguous from row0) -> bf16,
// zero-padded to rows_pad x K
__global__ void gather_bf16_kernel(const float* __restrict__ x, int64_t ld, int d,
                                   const int64_t* __restrict__ idx, int64_t row0, int rows,
                                   int rows_pad, int K, uint16_t* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)rows_pad * K) return;
  const int r = (int)(q / K), c = (int)(q - (int64_t)r * K);
  float v = 0.0f;
  if (r < rows && c < d) {
    const int64_t src = idx ? idx[row0 + r] : row0 + r;
    v = x[src * ld + c];
  }
  out[q] = f2b(v);
}

// dY = 2/(B*O) (Y - T) on the first O columns of a rows_pad x NP output, stored bf16
// (padded columns/rows zero) and the per-batch loss sum
__global__ void mse_grad_kernel(const float* __restrict__ Y, int NP, const float* __restrict__ T,
                                int O, int rows, int rows_pad, float scale,
                                uint16_t* __restrict__ dYb, float* __restrict__ dY32,
                                float* __restrict__ loss) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rows_pad * NP) return;
  const int r = q / NP, c = q - r * NP;
  float g = 0.0f;
  if (r < rows && c < O) {
    const float e = Y[q] - T[(size_t)r * O + c];
    g = scale * e;
    atomicAdd(loss, e * e);
  }
  dY32[q] = g;
  dYb[q] = f2b(g);
}

// dZ = dA * gelu'(Z) (f32 and bf16 copies); rows >= valid rows are zeroed
__global__ void gelu_back_kernel(const float* __restrict__ dA, const float* __restrict__ Z,
                                 int rows, int rows_pad, int N, float* __restrict__ dZ32,
                                 uint16_t* __restrict__ dZb) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rows_pad * N) return;
  const int r = q / N;
  const float g = r < rows ? dA[q] * gelu_grad(Z[q]) : 0.0f;
  dZ32[q] = g;
  dZb[q] = f2b(g);
}

// column sums of a rows x N matrix (bias gradients); one thread per column
__global__ void colsum_kernel(const float* __restrict__ G, int rows, int N, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.0f;
  for (int r = 0; r < rows; ++r) s += G[(size_t)r * N + c];
  out[c] = s;
}

// Adam on fp32 master weights, refreshing the bf16 copy
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                            float b1, float b2, float c1, float c2, float eps,
                            uint16_t* __restrict__ wb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2b(wi);
}

__global__ void to_bf16_kernel(const float* __restrict__ x, int64_t n, uint16_t* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2b(x[i]);
}

// out[r][c] = Y[r][c] * sd[c] + mu[c] for c < O (strip column padding)
__global__ void unstandardise_kernel(const float* __restrict__ Y, int NP, int O, int rows,
                                     const float* __restrict__ mu, const float* __restrict__ sd,
                                     float* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rows * O) return;
  const int r = q / O, c = q - r * O;
  out[q] = Y[(size_t)r * NP + c] * sd[c] + mu[c];
}

__global__ void gather_rows_kernel(const float* __restrict__ x, int64_t ld, int d,
                                   const int64_t* __restrict__ idx, int64_t rows,
                                   float* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rows * d) return;
  const int64_t r = q / d;
  out[q] = x[idx[r] * ld + (q - r * d)];
}

inline int64_t pad32(int64_t v) { return (v + 31) / 32 * 32; }

}  // namespace

// ------------------------------------------------------------------ host-side model
struct MlpModel {
  int D = 0, Dp = 0, O = 0, Op = 0;  // input and output widths padded to 32
  // layer l: W (out x in) fp32 master + bf16 copy, bias fp32
  Buf<float> W[3], Bv[3], mW[3], vW[3], mB[3], vB[3], gW[3], gB[3];
  Buf<uint16_t> Wb[3];
  int rows_of(int l) const { return l == 0 ? HID : (l == 1 ? HID : Op); }
  int cols_of(int l) const { return l == 0 ? Dp : HID; }
  int64_t step = 0;
};

void mlp_init(MlpModel& M, int D, int O, uint64_t seed, hipStream_t s, DevPool& pool) {
  M.D = D;
  M.Dp = (int)pad32(D);  // K of layer 1 and N of its weight gradient
  M.O = O;
  M.Op = (int)pad32(O);
  for (int l = 0; l < 3; ++l) {
    const int ro = M.rows_of(l), ci = M.cols_of(l);
    const int fan_in = l == 0 ? D : HID;
    const int real_out = l == 2 ? O : HID;
    std::vector<float> w((size_t)ro * ci, 0.0f), b(ro, 0.0f);
    const double bound = std::sqrt(6.0 / (double)(fan_in + real_out));  // Glorot uniform
    for (int r = 0; r < real_out; ++r)
      for (int c = 0; c < fan_in; ++c)
        w[(size_t)r * ci + c] =
            (float)((2.0 * Philox::unit(seed, (uint32_t)l, (uint32_t)r, (uint32_t)c, TAG_MLPW) - 1.0) * bound);
    const size_t nw = w.size();
    M.W[l].alloc(pool, nw);
    M.Wb[l].alloc(pool, nw);
    M.Bv[l].alloc(pool, ro);
    M.mW[l].alloc(pool, nw);
    M.vW[l].alloc(pool, nw);
    M.gW[l].alloc(pool, nw);
    M.mB[l].alloc(pool, ro);
    M.vB[l].alloc(pool, ro);
    M.gB[l].alloc(pool, ro);
    EVOC_HIP(hipMemcpyAsync(M.W[l].p, w.data(), nw * sizeof(float), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemcpyAsync(M.Bv[l].p, b.data(), ro * sizeof(float), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemsetAsync(M.mW[l].p, 0, nw * sizeof(float), s));
    EVOC_HIP(hipMemsetAsync(M.vW[l].p, 0, nw * sizeof(float), s));
    EVOC_HIP(hipMemsetAsync(M.mB[l].p, 0, ro * sizeof(float), s));
    EVOC_HIP(hipMemsetAsync(M.vB[l].p, 0, ro * sizeof(float), s));
    to_bf16_kernel<<<(unsigned)ceil_div((int64_t)nw, 256), 256, 0, s>>>(M.W[l].p, (int64_t)nw, M.Wb[l].p);
  }
  EVOC_HIP(hipGetLastError());
}

void mlp_set_weights(MlpModel& M, const float* const* w_host, const float* const* b_host,
                     hipStream_t s) {
  for (int l = 0; l < 3; ++l) {
    const int ro = M.rows_of(l), ci = M.cols_of(l);
    const int real_out = l == 2 ? M.O : HID, real_in = l == 0 ? M.D : HID;
    std::vector<float> w((size_t)ro * ci, 0.0f), b(ro, 0.0f);
    for (int r = 0; r < real_out; ++r) {
      for (int c = 0; c < real_in; ++c) w[(size_t)r * ci + c] = w_host[l][(size_t)r * real_in + c];
      b[r] = b_host[l][r];
    }
    EVOC_HIP(hipMemcpyAsync(M.W[l].p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemcpyAsync(M.Bv[l].p, b.data(), ro * sizeof(float), hipMemcpyHostToDevice, s));
    to_bf16_kernel<<<(unsigned)ceil_div((int64_t)w.size(), 256), 256, 0, s>>>(M.W[l].p, (int64_t)w.size(), M.Wb[l].p);
  }
  EVOC_HIP(hipStreamSynchronize(s));
}

// Activations of one batch (rows_pad rows).
struct MlpBatch {
  int rows = 0, rows_pad = 0;
  Buf<uint16_t> X, A1, A2, T0, T1, dYb, dZb, Wt;
  Buf<float> Z1, Z2, Y, dY, dA, dZ;
  void ensure(const MlpModel& M, int rp, DevPool& pool) {
    if (rp <= rows_pad) return;
    rows_pad = rp;
    X.alloc(pool, (size_t)rp * M.Dp);
    A1.alloc(pool, (size_t)rp * HID);
    A2.alloc(pool, (size_t)rp * HID);
    Z1.alloc(pool, (size_t)rp * HID);
    Z2.alloc(pool, (size_t)rp * HID);
    Y.alloc(pool, (size_t)rp * M.Op);
    dY.alloc(pool, (size_t)rp * M.Op);
    dYb.alloc(pool, (size_t)rp * M.Op);
    dA.alloc(pool, (size_t)rp * HID);
    dZ.alloc(pool, (size_t)rp * HID);
    dZb.alloc(pool, (size_t)rp * HID);
    T0.alloc(pool, (size_t)rp * std::max(M.Dp, HID));
    T1.alloc(pool, (size_t)rp * std::max(M.Dp, HID));
    Wt.alloc(pool, (size_t)HID * std::max(M.Dp, HID));
  }
};

static void gemm(const uint16_t* A, const uint16_t* B, int M, int N, int K, const float* bias,
                 float* C32, uint16_t* Cb, int mode, hipStream_t s) {
  const int tiles = (M / 32) * (N / 32);
  gemm_nt_kernel<<<(unsigned)ceil_div(tiles, 4), 256, 0, s>>>(A, B, M, N, K, bias, C32, Cb, mode);
}
static void transpose(const uint16_t* in, int R, int Cc, uint16_t* out, hipStream_t s) {
  dim3 g((unsigned)ceil_div(Cc, 32), (unsigned)ceil_div(R, 32));
  transpose_kernel<<<g, 256, 0, s>>>(in, R, Cc, out);
}

// forward on Bt.X (already filled); leaves Z1, A1, Z2, A2, Y
static void mlp_forward(const MlpModel& M, MlpBatch& Bt, hipStream_t s) {
  const int R = Bt.rows_pad;
  gemm(Bt.X.p, M.Wb[0].p, R, HID, M.Dp, M.Bv[0].p, Bt.Z1.p, Bt.A1.p, 1, s);
  gemm(Bt.A1.p, M.Wb[1].p, R, HID, HID, M.Bv[1].p, Bt.Z2.p, Bt.A2.p, 1, s);
  gemm(Bt.A2.p, M.Wb[2].p, R, M.Op, HID, M.Bv[2].p, Bt.Y.p, nullptr, 0, s);
}

// backward of the MSE loss against T (rows x O, f32 device); gradients into M.gW / M.gB
static void mlp_backward(MlpModel& M, MlpBatch& Bt, const float* T, float* loss, hipStream_t s) {
  const int R = Bt.rows_pad, rows = Bt.rows;
  const float scale = 2.0f / (float)((int64_t)rows * M.O);
  mse_grad_kernel<<<(unsigned)ceil_div((int64_t)R * M.Op, 256), 256, 0, s>>>(
      Bt.Y.p, M.Op, T, M.O, rows, R, scale, Bt.dYb.p, Bt.dY.p, loss);
  // layer 3: dW3 = dY^T A2, db3 = colsum dY, dA2 = dY W3
  transpose(Bt.dYb.p, R, M.Op, Bt.T0.p, s);   // Op x R
  transpose(Bt.A2.p, R, HID, Bt.T1.p, s);     // HID x R
  gemm(Bt.T0.p, Bt.T1.p, M.Op, HID, R, nullptr, M.gW[2].p, nullptr, 0, s);
  colsum_kernel<<<(unsigned)ceil_div(M.Op, 256), 256, 0, s>>>(Bt.dY.p, R, M.Op, M.gB[2].p);
  transpose(M.Wb[2].p, M.Op, HID, Bt.Wt.p, s);  // HID x Op
  gemm(Bt.dYb.p, Bt.Wt.p, R, HID, M.Op, nullptr, Bt.dA.p, nullptr, 0, s);
  gelu_back_kernel<<<(unsigned)ceil_div((int64_t)R * HID, 256), 256, 0, s>>>(
      Bt.dA.p, Bt.Z2.p, rows, R, HID, Bt.dZ.p, Bt.dZb.p);
  // layer 2: dW2 = dZ2^T A1, db2, dA1 = dZ2 W2
  transpose(Bt.dZb.p, R, HID, Bt.T0.p, s);
  transpose(Bt.A1.p, R, HID, Bt.T1.p, s);
  gemm(Bt.T0.p, Bt.T1.p, HID, HID, R, nullptr, M.gW[1].p, nullptr, 0, s);
  colsum_kernel<<<(unsigned)ceil_div(HID, 256), 256, 0, s>>>(Bt.dZ.p, R, HID, M.gB[1].p);
  transpose(M.Wb[1].p, HID, HID, Bt.Wt.p, s);
  gemm(Bt.dZb.p, Bt.Wt.p, R, HID, HID, nullptr, Bt.dA.p, nullptr, 0, s);
  gelu_back_kernel<<<(unsigned)ceil_div((int64_t)R * HID, 256), 256, 0, s>>>(
      Bt.dA.p, Bt.Z1.p, rows, R, HID, Bt.dZ.p, Bt.dZb.p);
  // layer 1: dW1 = dZ1^T X, db1
  transpose(Bt.dZb.p, R, HID, Bt.T0.p, s);
  transpose(Bt.X.p, R, M.Dp, Bt.T1.p, s);     // Dp x R
  gemm(Bt.T0.p, Bt.T1.p, HID, M.Dp, R, nullptr, M.gW[0].p, nullptr, 0, s);
  colsum_kernel<<<(unsigned)ceil_div(HID, 256), 256, 0, s>>>(Bt.dZ.p, R, HID, M.gB[0].p);
  EVOC_HIP(hipGetLastError());
}

static void mlp_adam(MlpModel& M, float lr, hipStream_t s) {
  M.step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)M.step), c2 = 1.0f - std::pow(b2, (float)M.step);
  for (int l = 0; l < 3; ++l) {
    const int64_t nw = (int64_t)M.rows_of(l) * M.cols_of(l);
    adam_kernel<<<(unsigned)ceil_div(nw, 256), 256, 0, s>>>(M.W[l].p, M.gW[l].p, M.mW[l].p,
                                                           M.vW[l].p, nw, lr, b1, b2, c1, c2,
                                                           1e-8f, M.Wb[l].p);
    adam_kernel<<<(unsigned)ceil_div(M.rows_of(l), 256), 256, 0, s>>>(
        M.Bv[l].p, M.gB[l].p, M.mB[l].p, M.vB[l].p, M.rows_of(l), lr, b1, b2, c1, c2, 1e-8f,
        nullptr);
  }
}

// ------------------------------------------------------------------ warm start
bool warm_start_init(const float* dX, int64_t n, int d, int64_t ld, const evoc_params_t& p,
                     const evoc_hip_nn_params_t& nn, int dim, uint64_t seed, MlpWarmStart& ws,
                     const WarmJudge& judge, float* d_init, hipStream_t s, DevPool& pool,
                     int verbose) {
  const int64_t S = std::min<int64_t>(ws.subsample, n);
  // 1. subsample (sorted, Philox-keyed) and its plain layout
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  for (int64_t i = 0; i < S; ++i) {  // partial Fisher-Yates
    const uint64_t r = ((uint64_t)Philox::first(seed, (uint32_t)i, (uint32_t)(i >> 32), 1, TAG_MLPS) << 32) |
                       Philox::first(seed, (uint32_t)i, (uint32_t)(i >> 32), 2, TAG_MLPS);
    const int64_t j = i + (int64_t)(r % (uint64_t)(n - i));
    std::swap(order[i], order[j]);
  }
  std::vector<int64_t> sub(order.begin(), order.begin() + S);
  std::sort(sub.begin(), sub.end());
  Buf<int64_t> dsub(pool, S);
  EVOC_HIP(hipMemcpyAsync(dsub.p, sub.data(), S * sizeof(int64_t), hipMemcpyHostToDevice, s));
  Buf<float> xs(pool, (size_t)S * d);
  Graph sg_keep;
  {
    gather_rows_kernel<<<(unsigned)ceil_div(S * d, 256), 256, 0, s>>>(dX, ld, d, dsub.p, S, xs.p);
    EVOC_HIP(hipGetLastError());
    const int k = p.n_neighbours;
    Buf<int32_t> si(pool, S * k);
    Buf<float> sd(pool, S * k);
    knn_device(xs.p, S, d, d, nn.metric, k, nn.knn_precision, si.p, sd.p, nullptr, s, pool);
    Graph& sg = sg_keep;
    fuzzy_device(si.p, sd.p, S, k, p.neighbour_scale * (float)k, true, nullptr, nullptr, sg, s, pool);
    std::vector<int64_t> ptr(S + 1);
    std::vector<int32_t> col(sg.nnz);
    std::vector<float> w(sg.nnz);
    EVOC_HIP(hipMemcpyAsync(ptr.data(), sg.indptr.p, (S + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(col.data(), sg.indices.p, sg.nnz * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(w.data(), sg.weights.p, sg.nnz * sizeof(float), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
    std::vector<float> init((size_t)S * dim);
    label_prop_init_host(ptr.data(), col.data(), w.data(), S, dim, xs.p, d, d, seed, init.data(),
                         s, pool, 0);
    Buf<float> lay(pool, (size_t)S * dim);
    EVOC_HIP(hipMemcpyAsync(lay.p, init.data(), init.size() * sizeof(float), hipMemcpyHostToDevice, s));
    EmbedParams ep;
    ep.n_epochs = p.n_epochs;
    ep.noise_level = p.noise_level;
    embed_device(sg, dim, ep, lay.p, false, seed, s, pool);
    // standardise the targets per coordinate
    std::vector<float> hl((size_t)S * dim);
    EVOC_HIP(hipMemcpyAsync(hl.data(), lay.p, hl.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
    std::vector<float> mu(dim, 0.0f), sdv(dim, 0.0f);
    for (int c = 0; c < dim; ++c) {
      double m = 0, v = 0;
      for (int64_t i = 0; i < S; ++i) m += hl[i * dim + c];
      m /= (double)S;
      for (int64_t i = 0; i < S; ++i) v += (hl[i * dim + c] - m) * (hl[i * dim + c] - m);
      mu[c] = (float)m;
      sdv[c] = (float)std::sqrt(std::max(v / (double)S, 1e-12));
    }
    for (int64_t i = 0; i < S; ++i)
      for (int c = 0; c < dim; ++c) hl[i * dim + c] = (hl[i * dim + c] - mu[c]) / sdv[c];
    Buf<float> T(pool, (size_t)S * dim), dmu(pool, dim), dsd(pool, dim);
    EVOC_HIP(hipMemcpyAsync(T.p, hl.data(), hl.size() * sizeof(float), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemcpyAsync(dmu.p, mu.data(), dim * sizeof(float), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemcpyAsync(dsd.p, sdv.data(), dim * sizeof(float), hipMemcpyHostToDevice, s));
    // 2. train
    MlpModel M;
    mlp_init(M, d, dim, seed, s, pool);
    MlpBatch Bt;
    const int batch = 512;
    Bt.ensure(M, batch, pool);
    Buf<float> loss(pool, 1);
    const int epochs = ws.train_epochs > 0 ? ws.train_epochs : 20;
    const int64_t nb = ceil_div(S, batch);
    std::vector<int64_t> perm(S);
    std::iota(perm.begin(), perm.end(), 0);
    Buf<int64_t> dperm(pool, S);
    Buf<float> Tb(pool, (size_t)batch * dim);
    // targets permuted each epoch on the host, gathered per batch
    double last_loss = 0.0;
    std::vector<float> tb((size_t)S * dim);
    Buf<float> Tperm(pool, (size_t)S * dim);
    hipEvent_t e0, e1;
    EVOC_HIP(hipEventCreate(&e0));
    EVOC_HIP(hipEventCreate(&e1));
    EVOC_HIP(hipEventRecord(e0, s));
    for (int ep_i = 0; ep_i < epochs; ++ep_i) {
      for (int64_t i = S - 1; i > 0; --i) {
        const uint32_t r = Philox::first(seed, (uint32_t)i, (uint32_t)ep_i, 3, TAG_MLPS);
        std::swap(perm[i], perm[Philox::range(r, (uint64_t)i + 1)]);
      }
      for (int64_t i = 0; i < S; ++i)
        std::memcpy(&tb[i * dim], &hl[perm[i] * dim], dim * sizeof(float));
      EVOC_HIP(hipMemcpyAsync(dperm.p, perm.data(), S * sizeof(int64_t), hipMemcpyHostToDevice, s));
      EVOC_HIP(hipMemcpyAsync(Tperm.p, tb.data(), tb.size() * sizeof(float), hipMemcpyHostToDevice, s));
      EVOC_HIP(hipMemsetAsync(loss.p, 0, sizeof(float), s));
      const float lr = 2e-3f * (1.0f - 0.9f * (float)ep_i / (float)epochs);
      for (int64_t b = 0; b < nb; ++b) {
        const int rows = (int)std::min<int64_t>(batch, S - b * batch);
        const int rp = (int)pad32(rows);
        Bt.rows = rows;
        Bt.rows_pad = std::max(Bt.rows_pad, rp);
        const int R = (int)pad32(batch);
        gather_bf16_kernel<<<(unsigned)ceil_div((int64_t)R * M.Dp, 256), 256, 0, s>>>(
            xs.p, d, d, dperm.p, b * batch, rows, R, M.Dp, Bt.X.p);
        const int save = Bt.rows_pad;
        Bt.rows_pad = R;
        mlp_forward(M, Bt, s);
        mlp_backward(M, Bt, Tperm.p + (size_t)b * batch * dim, loss.p, s);
        mlp_adam(M, lr, s);
        Bt.rows_pad = save;
      }
      float hloss = 0.0f;
      EVOC_HIP(hipMemcpyAsync(&hloss, loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
      EVOC_HIP(hipStreamSynchronize(s));
      last_loss = hloss / (double)(S * dim);
      if (verbose >= 2) std::printf("   MLP epoch %d/%d  mse %.5f\n", ep_i + 1, epochs, last_loss);
    }
    EVOC_HIP(hipEventRecord(e1, s));
    EVOC_HIP(hipStreamSynchronize(s));
    float tms = 0.0f;
    EVOC_HIP(hipEventElapsedTime(&tms, e0, e1));
    EVOC_HIP(hipEventDestroy(e0));
    EVOC_HIP(hipEventDestroy(e1));
    ws.train_ms = tms;
    ws.final_loss = last_loss;
    ws.train_epochs = epochs;
    if (verbose) std::printf("  MLP trained on %lld points: %d epochs, mse %.4f, %.1f ms\n",
                             (long long)S, epochs, last_loss, tms);
    // 3. predict: the subsample first (for the judge), then, if accepted, every point
    const int chunk = 16384;
    MlpBatch P;
    P.ensure(M, chunk, pool);
    auto predict = [&](const float* src, int64_t lds, int64_t rows_total, float* dst) {
      for (int64_t r0 = 0; r0 < rows_total; r0 += chunk) {
        const int rows = (int)std::min<int64_t>(chunk, rows_total - r0);
        const int R = (int)pad32(rows);
        gather_bf16_kernel<<<(unsigned)ceil_div((int64_t)R * M.Dp, 256), 256, 0, s>>>(
            src, lds, d, nullptr, r0, rows, R, M.Dp, P.X.p);
        const int save = P.rows_pad;
        P.rows_pad = R;
        mlp_forward(M, P, s);
        P.rows_pad = save;
        unstandardise_kernel<<<(unsigned)ceil_div((int64_t)rows * dim, 256), 256, 0, s>>>(
            P.Y.p, M.Op, dim, rows, dmu.p, dsd.p, dst + r0 * dim);
      }
      EVOC_HIP(hipGetLastError());
    };
    Buf<float> pred(pool, (size_t)S * dim);
    predict(xs.p, d, S, pred.p);
    EVOC_HIP(hipStreamSynchronize(s));
    ws.accepted = judge(sg_keep, lay.p, pred.p, S, ws);
    if (verbose)
      std::printf("  MLP layout on the subsample: ARI %.4f, MST ratio %.4f -> %s\n", ws.judge_ari,
                  ws.judge_mst_ratio, ws.accepted ? "accepted" : "rejected (plain init)");
    if (!ws.accepted) return false;
    predict(dX, ld, n, d_init);
    EVOC_HIP(hipStreamSynchronize(s));
    return true;
  }
}

// ------------------------------------------------------------------ gradient check entry
// Loss and gradients of the MSE of the MLP on one batch, for checking the hand-written
// backward pass against an fp64 reference.
void mlp_loss_and_grads(const float* x, const float* t, int rows, int D, int O,
                        const float* const* w, const float* const* b, float* const* gw,
                        float* const* gb, float* out_y, double* loss, hipStream_t s,
                        DevPool& pool) {
  MlpModel M;
  mlp_init(M, D, O, 1, s, pool);
  mlp_set_weights(M, w, b, s);
  MlpBatch Bt;
  const int R = (int)pad32(rows);
  Bt.ensure(M, R, pool);
  Bt.rows = rows;
  Bt.rows_pad = R;
  Buf<float> dx(pool, (size_t)rows * D), dt(pool, (size_t)rows * O), dl(pool, 1);
  EVOC_HIP(hipMemcpyAsync(dx.p, x, (size_t)rows * D * sizeof(float), hipMemcpyHostToDevice, s));
  EVOC_HIP(hipMemcpyAsync(dt.p, t, (size_t)rows * O * sizeof(float), hipMemcpyHostToDevice, s));
  EVOC_HIP(hipMemsetAsync(dl.p, 0, sizeof(float), s));
  gather_bf16_kernel<<<(unsigned)ceil_div((int64_t)R * M.Dp, 256), 256, 0, s>>>(
      dx.p, D, D, nullptr, 0, rows, R, M.Dp, Bt.X.p);
  mlp_forward(M, Bt, s);
  mlp_backward(M, Bt, dt.p, dl.p, s);
  EVOC_HIP(hipStreamSynchronize(s));
  for (int l = 0; l < 3; ++l) {
    const int ro = M.rows_of(l), ci = M.cols_of(l);
    const int real_out = l == 2 ? O : HID, real_in = l == 0 ? D : HID;
    std::vector<float> hw((size_t)ro * ci), hb(ro);
    EVOC_HIP(hipMemcpy(hw.data(), M.gW[l].p, hw.size() * sizeof(float), hipMemcpyDeviceToHost));
    EVOC_HIP(hipMemcpy(hb.data(), M.gB[l].p, hb.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int r = 0; r < real_out; ++r) {
      for (int c = 0; c < real_in; ++c) gw[l][(size_t)r * real_in + c] = hw[(size_t)r * ci + c];
      gb[l][r] = hb[r];
    }
  }
  std::vector<float> hy((size_t)R * M.Op);
  EVOC_HIP(hipMemcpy(hy.data(), Bt.Y.p, hy.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < O; ++c) out_y[(size_t)r * O + c] = hy[(size_t)r * M.Op + c];
  float hl = 0.0f;
  EVOC_HIP(hipMemcpy(&hl, dl.p, sizeof(float), hipMemcpyDeviceToHost));
  *loss = (double)hl / (double)((int64_t)rows * O);
}

}  // namespace evoc
