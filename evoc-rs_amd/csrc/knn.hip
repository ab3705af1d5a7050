// Exhaustive kNN self-join for gfx950 (the exhaustive arm of run_ann_search,
// src/nearest_neighbours/nearest_neighbour_cpu.rs:313-317, with the j != i self
// filter of src/nearest_neighbours/nearest_neighbour_gpu.rs:291-303).
//
//  1. knn_prep:   canonical f32 norms + bf16 copies (hi, and lo = bf16(x - hi) when
//                 the input is not bf16-representable).
//  2. knn_topc:   S = Q X^T on v_mfma_f32_32x32x16_bf16, 128x128 tiles, per-row
//                 top-C candidate lists kept in registers (never materialises S).
//  3. knn_rerank: exact f32 distances of the C candidates in the canonical order
//                 (64 strided lane sums + halving tree), sort by (dist, idx), and a
//                 per-row certificate that no non-candidate can enter the top k.
//  4. knn_exact:  brute-force exact rows for the (rare) rows whose certificate fails.
#include <cstdio>
#include <cstdlib>

#include "kernels.h"

namespace evoc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128;   // query rows per workgroup
constexpr int BN = 128;   // database columns per tile
constexpr int KC = 32;    // k-chunk staged through LDS
constexpr int SCS = 132;  // padded LDS row stride (floats) of the key tile
constexpr float FINF = __builtin_huge_valf();

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t b = __float_as_uint(f);
  if ((b & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  uint32_t r = b + 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// canonical halving-tree reduction across the 64 lanes: result valid in lane 0
__device__ __forceinline__ float tree64(float acc) {
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) acc = acc + __shfl_down(acc, s, 64);
  return acc;
}

__device__ __forceinline__ float cos_dist(float dot, float na, float nb) {
  float den = na * nb;
  if (!(den > 0.0f)) return 1.0f;
  float c = dot / den;
  return 1.0f - c;
}

// ------------------------------------------------------------------ 1. prep
__global__ __launch_bounds__(256) void knn_prep_kernel(
    const float* __restrict__ X, int64_t n, int d, int64_t ld, int64_t npad, int dpad,
    uint16_t* __restrict__ Xh, uint16_t* __restrict__ Xl, float* __restrict__ nrm,
    float* __restrict__ sqn, float* __restrict__ colscale, int metric,
    unsigned* __restrict__ flags /* [0]=inexact, [1]=max norm bits */) {
  // one wavefront per row, rows strided over the grid; the flags are folded per
  // wavefront so that each wavefront issues one atomic of each kind
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  bool any_lo = false;
  unsigned maxbits = 0;
  for (int64_t row = (int64_t)blockIdx.x * 4 + wave; row < npad; row += nwaves) {
    float acc = 0.0f;
    for (int t = lane; t < dpad; t += 64) {
      float v = (row < n && t < d) ? X[row * ld + t] : 0.0f;
      float p = v * v;
      acc = acc + p;
      uint16_t h = f2bf(v);
      float r = v - bf2f(h);
      uint16_t lo = f2bf(r);
      Xh[row * dpad + t] = h;
      if (Xl) Xl[row * dpad + t] = lo;
      any_lo |= (lo & 0x7fffu) != 0;
    }
    acc = tree64(acc);
    if (lane == 0) {
      float nv = sqrtf(acc);
      if (row < n) {
        nrm[row] = nv;
        sqn[row] = acc;
        maxbits = max(maxbits, __float_as_uint(nv));
      }
      float cs;
      if (metric == 1) cs = (row < n && nv > 0.0f) ? 1.0f / nv : 0.0f;
      else cs = row < n ? acc : 0.0f;
      colscale[row] = cs;
    }
  }
  if (__any(any_lo) && lane == 0) atomicOr(&flags[0], 1u);
  if (lane == 0 && maxbits) atomicMax(&flags[1], maxbits);
}

// FP8-e4m3 copy for the FP8 first pass (cosine): per row the scale s = max|x| / 448 maps
// the largest magnitude onto e4m3's largest finite value; q = e4m3(x / s) rounded to
// nearest even (hardware conversion, inputs clamped to +-448, subnormal results stored
// as zero); the pass's key scale is
// s / ||x||. The certificate needs the quantisation error: rho = ||x - s q|| / ||x|| per
// row, formed in f64 from the decoded values (s q is exact in f64), rounded up and folded
// into the maximum over rows (positive floats order as their bit patterns).
__global__ __launch_bounds__(256) void knn_prep_fp8_kernel(
    const float* __restrict__ X, int64_t n, int d, int64_t ld, int64_t npad, int dpad8,
    const float* __restrict__ nrm, uint8_t* __restrict__ Xq, float* __restrict__ scale8,
    unsigned* __restrict__ rho_bits) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  unsigned rmax = 0;
  for (int64_t row = (int64_t)blockIdx.x * 4 + wave; row < npad; row += nwaves) {
    float amax = 0.0f;
    if (row < n)
      for (int t = lane; t < d; t += 64) amax = fmaxf(amax, fabsf(X[row * ld + t]));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
    const float sc = amax > 0.0f ? amax / 448.0f : 0.0f;
    const float inv = amax > 0.0f ? 448.0f / amax : 0.0f;
    double e2 = 0.0, x2 = 0.0;
    for (int t = lane; t < dpad8; t += 64) {
      const float v = (row < n && t < d) ? X[row * ld + t] : 0.0f;
      const float y = fminf(fmaxf(v * inv, -448.0f), 448.0f);
      int pk = __builtin_amdgcn_cvt_pk_fp8_f32(y, 0.0f, 0, false);
      float back = __builtin_amdgcn_cvt_f32_fp8(pk, 0);
      if (fabsf(back) < 0.015625f) {  // subnormal e4m3 (below 2^-6) stored as zero, so the
        pk = 0;                       // MFMA's reading of the bytes cannot differ from the
        back = 0.0f;                  // values the error bound is formed from
      }
      const double r = (double)v - (double)sc * (double)back;
      e2 += r * r;
      x2 += (double)v * (double)v;
      Xq[row * dpad8 + t] = (uint8_t)(pk & 0xff);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      e2 += __shfl_xor(e2, o, 64);
      x2 += __shfl_xor(x2, o, 64);
    }
    if (lane == 0) {
      float rho = 0.0f;
      if (row < n && x2 > 0.0) rho = (float)(sqrt(e2 / x2) * (1.0 + 1e-6));
      rmax = max(rmax, __float_as_uint(rho));
      scale8[row] = (row < n && nrm[row] > 0.0f) ? sc / nrm[row] : 0.0f;
    }
  }
  if (lane == 0 && rmax) atomicMax(rho_bits, rmax);
}

// ------------------------------------------------------------------ 2. MFMA top-C
template <int C>
__device__ __forceinline__ void list_insert(float (&lk)[C], int (&li)[C], float key, int idx) {
  float nk = key;
  int ni = idx;
#pragma unroll
  for (int q = 0; q < C; ++q) {
    const bool sw = (nk < lk[q]) | ((nk == lk[q]) & (ni < li[q]));  // no short-circuit branches
    float tk = sw ? lk[q] : nk;
    int ti = sw ? li[q] : ni;
    lk[q] = sw ? nk : lk[q];
    li[q] = sw ? ni : li[q];
    nk = tk;
    ni = ti;
  }
}

// vmcnt wait with a compile-time count (gfx9 encoding: vmcnt in bits 3:0 and 15:14,
// expcnt and lgkmcnt left at their maxima so only vector-memory counts are awaited)
template <int N>
__device__ __forceinline__ void wait_vm() {
  __builtin_amdgcn_s_waitcnt((N & 15) | (((N >> 4) & 3) << 14) | (7 << 4) | (15 << 8));
}

// 16-byte LDS slot of piece q (8 bf16 along k) of staged row r: XOR swizzle so that the
// 16 rows a half-wavefront reads with one ds_read_b128 hit 16 distinct bank groups
__device__ __forceinline__ int slot_of(int r, int q) { return r * 4 + (q ^ ((r >> 2) & 3)); }

typedef __attribute__((address_space(3))) void lds_void;

// Workgroup barrier for the direct-to-LDS ring: each wavefront has already waited for
// its own loads of the stage with wait_vm<N>(), so no memory fence is wanted here
// (__syncthreads() fences LDS, and with loads in flight that means vmcnt(0), draining
// the whole prefetch). The empty asm statements keep the compiler from moving memory
// operations across the barrier.
__device__ __forceinline__ void ring_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

typedef int v8i __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

template <bool SPLIT, int C, bool F8 = false>  // C = list length per column half
// Query rows are the rows of Xh / Xl themselves (self-join), or, with qmap, the rows of
// Qh / Ql: nq gathered rows whose global indices qmap holds in ascending order (the
// re-search of rows whose certificate failed, with longer lists). F8: the rows are
// FP8-e4m3 bytes (per-row scales folded into colscale), 64 per 64-byte chunk, multiplied
// on v_mfma_scale_f32_32x32x64_f8f6f4 with unit scales and f32 accumulation.
__global__ __launch_bounds__(256, 1) void knn_topc_kernel(
    const uint16_t* __restrict__ Xh, const uint16_t* __restrict__ Xl, int64_t n, int64_t npad,
    int dpad, const float* __restrict__ colscale, int metric, int32_t* __restrict__ cand_idx,
    float* __restrict__ cand_key, const uint16_t* __restrict__ Qh,
    const uint16_t* __restrict__ Ql, const int32_t* __restrict__ qmap, int64_t nq, int64_t ct0,
    int64_t ct1) {
  static_assert(!(F8 && SPLIT), "the FP8 pass has no split form");
  // column tiles [ct0, ct1) of BN columns are swept (a shard of the columns, or all of them)
  constexpr int NOPS = SPLIT ? 4 : 2;           // A hi, B hi, (A lo, B lo)
  constexpr int NST = SPLIT ? 2 : 4;            // stages in the LDS ring
  constexpr int OPB = BM * 4 * 16;              // bytes per staged operand chunk (8 KB)
  constexpr int STB = NOPS * OPB;               // bytes per stage
  constexpr int REGION = BM * SCS * 4;          // key tile / candidate queue
  constexpr int QCAP = REGION / 8;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  char* ring = smem;
  float* sc = reinterpret_cast<float*>(smem + NST * STB);
  float2* queue = reinterpret_cast<float2*>(sc);
  float* thr = reinterpret_cast<float*>(smem + NST * STB + REGION);  // [2][BM]
  int* qcount = reinterpret_cast<int*>(thr + 2 * BM);                 // [2]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t rb = blockIdx.x;
  const int64_t row0 = rb * BM;
  // global index of each query row of the block (-1 past the last one), for the self mask
  __shared__ int32_t qg[BM];
  const uint16_t* const QH = qmap ? Qh : Xh;
  const uint16_t* const QL = qmap ? Ql : Xl;
  const int64_t nrows = qmap ? nq : n;
  if (tid < BM) qg[tid] = row0 + tid < nrows ? (qmap ? qmap[row0 + tid] : (int32_t)(row0 + tid)) : -1;
  const int64_t qlo = qmap ? qmap[row0] : row0;
  const int64_t qhi = qmap ? qmap[std::min<int64_t>(row0 + BM, nq) - 1] : row0 + BM - 1;
  // 64-byte chunks of a row: 32 bf16 or 64 fp8 values
  const int64_t rowb = F8 ? (int64_t)dpad : 2 * (int64_t)dpad;  // bytes per row
  const int nck = (int)(rowb / 64);
  const int64_t g0 = ct0 * nck;
  const int64_t total = ct1 * nck;
  const int orow = tid & (BM - 1), ohalf = tid >> 7;  // list owned by this thread

  float lk[C];
  int li[C];
#pragma unroll
  for (int q = 0; q < C; ++q) {
    lk[q] = FINF;
    li[q] = -1;
  }
  thr[ohalf * BM + orow] = FINF;
  if (tid < 2) qcount[tid] = 0;

  // issue the direct-to-LDS loads of chunk g into its ring stage
  auto issue = [&](int64_t g) {
    const int64_t ct = g / nck;
    const int kc = (int)(g - ct * nck);
    char* st = ring + (int)(g % NST) * STB;
#pragma unroll
    for (int op = 0; op < NOPS; ++op) {
      const char* src = reinterpret_cast<const char*>((op & 1) ? ((op < 2) ? Xh : Xl)
                                                              : ((op < 2) ? QH : QL));
      const int64_t rbase = (op & 1) ? ct * BN : row0;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int blk = h * 4 + wave;          // 64 consecutive slots per instruction
        const int s = blk * 64 + lane;
        const int r = s >> 2, qs = s & 3;
        const int q = qs ^ ((r >> 2) & 3);
        const char* gp = src + (rbase + r) * rowb + kc * 64 + q * 16;
        __builtin_amdgcn_global_load_lds(gp, (lds_void*)(st + op * OPB + blk * 1024), 16, 0, 0);
      }
    }
  };

  f32x16 acc[2][2];
  float cs[2] = {0.0f, 0.0f};
  constexpr int PER_CHUNK = NOPS * 2;  // load instructions per thread per chunk
  for (int64_t g = g0; g < g0 + NST - 1 && g < total; ++g) issue(g);
  for (int64_t g = g0; g < total; ++g) {
    const int64_t ahead = std::min<int64_t>(NST - 2, total - 1 - g);
    if (ahead >= 2) wait_vm<2 * PER_CHUNK>();
    else if (ahead == 1) wait_vm<PER_CHUNK>();
    else wait_vm<0>();
    __syncthreads();
    if (g + NST - 1 < total) issue(g + NST - 1);
    const int64_t ct = g / nck;
    const int kc = (int)(g - ct * nck);
    if (kc == 0) {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (f32x16){};
#pragma unroll
      for (int tn = 0; tn < 2; ++tn) cs[tn] = colscale[ct * BN + wn * 64 + tn * 32 + (lane & 31)];
    }
    const char* st = ring + (int)(g % NST) * STB;
    if constexpr (F8) {
      // lane l holds 32 of the chunk's 64 values (pieces 2h, 2h + 1, h = l >> 5) of row
      // l & 31, for A and B alike, so each product pairs the same k
      const int q0 = 2 * (lane >> 5);
      v8i a[2], b[2];
#pragma unroll
      for (int tm = 0; tm < 2; ++tm) {
        const int r = wm * 64 + tm * 32 + (lane & 31);
        const i32x4 lo = *reinterpret_cast<const i32x4*>(st + 0 * OPB + slot_of(r, q0) * 16);
        const i32x4 hi = *reinterpret_cast<const i32x4*>(st + 0 * OPB + slot_of(r, q0 + 1) * 16);
        a[tm] = (v8i){lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
      }
#pragma unroll
      for (int tn = 0; tn < 2; ++tn) {
        const int r = wn * 64 + tn * 32 + (lane & 31);
        const i32x4 lo = *reinterpret_cast<const i32x4*>(st + 1 * OPB + slot_of(r, q0) * 16);
        const i32x4 hi = *reinterpret_cast<const i32x4*>(st + 1 * OPB + slot_of(r, q0 + 1) * 16);
        b[tn] = (v8i){lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
      }
      // formats 0/0: e4m3 for both operands; scales 127 (E8M0 for 2^0) leave them as is
#pragma unroll
      for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
          acc[tm][tn] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[tm], b[tn], acc[tm][tn], 0, 0,
                                                                       0, 127, 0, 127);
    }
#pragma unroll
    for (int kk = 0; kk < (F8 ? 0 : KC / 16); ++kk) {
      const int q = kk * 2 + (lane >> 5);
      bf16x8 a[2], b[2];
#pragma unroll
      for (int tm = 0; tm < 2; ++tm) {
        const int r = wm * 64 + tm * 32 + (lane & 31);
        a[tm] = *reinterpret_cast<const bf16x8*>(st + 0 * OPB + slot_of(r, q) * 16);
      }
#pragma unroll
      for (int tn = 0; tn < 2; ++tn) {
        const int r = wn * 64 + tn * 32 + (lane & 31);
        b[tn] = *reinterpret_cast<const bf16x8*>(st + 1 * OPB + slot_of(r, q) * 16);
      }
#pragma unroll
      for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
          acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[tm], b[tn], acc[tm][tn], 0, 0, 0);
      if constexpr (SPLIT) {
        bf16x8 al[2], bl[2];
#pragma unroll
        for (int tm = 0; tm < 2; ++tm) {
          const int r = wm * 64 + tm * 32 + (lane & 31);
          al[tm] = *reinterpret_cast<const bf16x8*>(st + 2 * OPB + slot_of(r, q) * 16);
        }
#pragma unroll
        for (int tn = 0; tn < 2; ++tn) {
          const int r = wn * 64 + tn * 32 + (lane & 31);
          bl[tn] = *reinterpret_cast<const bf16x8*>(st + 3 * OPB + slot_of(r, q) * 16);
        }
#pragma unroll
        for (int tm = 0; tm < 2; ++tm)
#pragma unroll
          for (int tn = 0; tn < 2; ++tn) {
            acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[tm], bl[tn], acc[tm][tn], 0, 0, 0);
            acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[tm], b[tn], acc[tm][tn], 0, 0, 0);
          }
      }
    }
    if (kc != nck - 1) continue;

    // ---------------- epilogue of column tile ct: filter keys against the row thresholds
    const int par = (int)(ct & 1);
    const bool diag = ct * BN <= qhi && ct * BN + BN > qlo;
    const bool tail = (ct + 1) * BN > n;
#define EVOC_KEY(TM, TN, J, I, OUT)                                                    \
  do {                                                                                   \
    const float s_ = acc[TM][TN][(J) * 4 + (I)];                                         \
    float k_ = metric == 1 ? -(s_ * cs[TN]) : (cs[TN] - 2.0f * s_);                      \
    if (diag || tail) {                                                                  \
      const int rl_ = wm * 64 + (TM) * 32 + (J) * 8 + (lane >> 5) * 4 + (I);             \
      const int64_t gc_ = ct * BN + wn * 64 + (TN) * 32 + (lane & 31);                   \
      if (gc_ >= n || gc_ == qg[rl_]) k_ = FINF;                                         \
    }                                                                                    \
    OUT = k_;                                                                            \
  } while (0)
    // thresholds of this lane's rows, four rows (one float4) at a time
    int mine = 0;
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float4 t4 = *reinterpret_cast<const float4*>(
            thr + wn * BM + wm * 64 + tm * 32 + j * 8 + (lane >> 5) * 4);
        const float tv[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            float key;
            EVOC_KEY(tm, tn, j, i, key);
            mine += key < tv[i];
          }
      }
    // wave-exclusive prefix of the pass counts, one LDS atomic per wave
    int incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    const int wtot = __shfl(incl, 63, 64);
    int base = 0;
    if (wtot > 0) {
      if (lane == 0) base = atomicAdd(&qcount[par], wtot);
      base = __shfl(base, 0, 64);
    }
    if (mine > 0) {
      int pos = base + incl - mine;
#pragma unroll
      for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float4 t4 = *reinterpret_cast<const float4*>(
              thr + wn * BM + wm * 64 + tm * 32 + j * 8 + (lane >> 5) * 4);
          const float tv[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
          for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              float key;
              EVOC_KEY(tm, tn, j, i, key);
              if (key < tv[i]) {
                if (pos < QCAP) {
                  const int rl = wm * 64 + tm * 32 + j * 8 + (lane >> 5) * 4 + i;
                  const int cl = wn * 64 + tn * 32 + (lane & 31);
                  queue[pos] = make_float2(key, __int_as_float((rl << 7) | cl));
                }
                ++pos;
              }
            }
        }
    }
    if (tid == 0) qcount[par ^ 1] = 0;  // counter of the next tile (its readers are done)
    __syncthreads();
    const int qn = qcount[par];
    const int cbase = (int)(ct * BN);
    if (qn <= QCAP) {
      for (int e = 0; e < qn; ++e) {
        const float2 en = queue[e];
        const int pk = __float_as_int(en.y);
        const int er = pk >> 7, ec = pk & 127;
        if (er == orow && (ec >> 6) == ohalf && en.x < lk[C - 1])
          list_insert<C>(lk, li, en.x, cbase + ec);
      }
    } else {
      // queue overflow (the first tiles): scan the whole key tile row by row
      __syncthreads();
#pragma unroll
      for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
          for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int rl = wm * 64 + tm * 32 + j * 8 + (lane >> 5) * 4 + i;
              const int cl = wn * 64 + tn * 32 + (lane & 31);
              float key;
              EVOC_KEY(tm, tn, j, i, key);
              sc[rl * SCS + cl] = key;
            }
      __syncthreads();
      const float* srow = sc + orow * SCS + ohalf * 64;
      const int cb = cbase + ohalf * 64;
      for (int q = 0; q < 64; q += 4) {
        const float4 v = *reinterpret_cast<const float4*>(srow + q);
        if (v.x < lk[C - 1]) list_insert<C>(lk, li, v.x, cb + q);
        if (v.y < lk[C - 1]) list_insert<C>(lk, li, v.y, cb + q + 1);
        if (v.z < lk[C - 1]) list_insert<C>(lk, li, v.z, cb + q + 2);
        if (v.w < lk[C - 1]) list_insert<C>(lk, li, v.w, cb + q + 3);
      }
    }
    thr[ohalf * BM + orow] = lk[C - 1];
  }
  // each (row, column-half) list is written as is: candidates per row = 2*C
  const int64_t grow = row0 + orow;
  if (grow < nrows) {
#pragma unroll
    for (int q = 0; q < C; ++q) {
      cand_key[(grow * 2 + ohalf) * C + q] = lk[q];
      cand_idx[(grow * 2 + ohalf) * C + q] = li[q];
    }
  }
}

// ------------------------------------------------------------------ 2b. resident-A top-C
// bf16-exact inputs (one MFMA per product). A 256-thread workgroup owns 256 query rows;
// wavefront w keeps the MFMA A-fragments of its 64 rows (all of d <= 512) in registers
// for the whole sweep, so the query block is read from memory once and only 32-column
// database tiles stream, in 128-wide k-chunks, through an 8-stage LDS ring fed by direct
// global->LDS loads; every B fragment read from LDS feeds two MFMAs. Lane l of
// wavefront w owns the candidate list of row w*64 + l, so the whole epilogue of a tile
// (keys, threshold filter, queue, list inserts) is private to the wavefront: the only
// workgroup barrier is the one per ring stage.
constexpr int RA_ROWS = 256;   // rows per workgroup
constexpr int RA_MAXD = 512;   // largest d the resident-query kernel takes
constexpr int RA_BN = 32;      // columns per tile
constexpr int RA_SCS = 36;     // padded key-tile row stride (floats)
constexpr int RA_BKT = 32;                               // bucket slots per row
constexpr int RA_DRAIN = 16;                             // drain once a row holds this many
constexpr int RA_WREG = 64 * RA_BKT * 8;                 // per-wavefront buckets / key tile
static_assert(RA_WREG >= 64 * RA_SCS * 4, "the key-tile dump must fit the bucket region");
constexpr int RA_EPI_BYTES = 4 * RA_WREG + RA_ROWS * 4;             // static: epilogue

// Ring geometry for a k-chunk of KC (128 for d <= 128, else 256): a stage holds the
// 32 x KC bf16 database chunk (16-byte pieces, XOR-swizzled) and one 256-byte copy of
// the tile's column scales per wavefront.
template <int KC>
struct RaRing {
  static constexpr int STB_B = RA_BN * KC * 2;            // 8 or 16 KB
  static constexpr int STB = STB_B + 4 * 256;
  static constexpr int NST = KC == 128 ? 8 : 4;           // stages (power of two)
  static constexpr int PIECES = KC / 8;                   // 16-byte pieces per column
  static constexpr int LOADS = STB_B / (256 * 16);        // 16-byte loads per thread
  static constexpr size_t SMEM = (size_t)NST * STB;       // dynamic LDS
  // slot of piece q of column c: 16 consecutive columns read with the same q hit 16
  // distinct slot positions mod 16 (conflict-free ds_read_b128)
  static __device__ __forceinline__ int slot(int c, int q) { return c * PIECES + (q ^ (c & 15)); }
};

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Sorted insert of (nk, ni) into an ascending key list, branch-free. Equal keys keep
// their arrival order (a new key goes after its equals), which is deterministic because
// each row receives its columns in a fixed order. What the certificate needs holds for
// any arrival order: every pair not in the list has a key >= the list's last key.
template <int C>
__device__ __forceinline__ void cand_insert(float (&lk)[C], int (&li)[C], float nk, int ni) {
#pragma unroll
  for (int q = 0; q < C; ++q) {
    const bool sw = nk < lk[q];
    const int ti = sw ? li[q] : ni;
    li[q] = sw ? ni : li[q];
    const float lo = fminf(nk, lk[q]), hi = fmaxf(nk, lk[q]);
    lk[q] = lo;
    nk = hi;
    ni = ti;
  }
}

// byte address of a shared-memory location, as LDS instructions take it
template <typename T>
__device__ __forceinline__ unsigned lds_addr(T* p) {
  return (unsigned)(uintptr_t)(__attribute__((address_space(3))) char*)(p);
}

// LDS accesses of the tile epilogue, written as asm so that the compiler, which cannot
// tell them apart from the ring's direct-to-LDS loads in flight, does not put a full
// vmcnt(0) before them. Each load waits for itself inside its own block; stores need no
// wait (LDS executes one wavefront's operations in order).
__device__ __forceinline__ void lds_st_b64(unsigned a, int2 v) {
  asm volatile("ds_write_b64 %0, %1" ::"v"(a), "v"(v) : "memory");
}
__device__ __forceinline__ void lds_st_b32(unsigned a, int v) {
  asm volatile("ds_write_b32 %0, %1" ::"v"(a), "v"(v) : "memory");
}
__device__ __forceinline__ int lds_ld_b32(unsigned a) {
  int v;
  asm volatile("ds_read_b32 %0, %1\n s_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(a) : "memory");
  return v;
}
__device__ __forceinline__ int2 lds_ld_b64(unsigned a) {
  int2 v;
  asm volatile("ds_read_b64 %0, %1\n s_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(a) : "memory");
  return v;
}

// C = list length, NKC = dpad / KC, COS = cosine keys (else squared-Euclidean)
template <int C, int NKC, int KC, bool COS>
__global__ __launch_bounds__(256, 1) void knn_topc_ra_kernel(
    const uint16_t* __restrict__ Xh, int64_t n, int64_t npad, int dpad,
    const float* __restrict__ colscale, int32_t* __restrict__ cand_idx,
    float* __restrict__ cand_key) {
  using R = RaRing<KC>;
  constexpr int NST = R::NST;
  constexpr int LOADS = R::LOADS;
  // The ring (written by direct-to-LDS loads) and the epilogue state are separate LDS
  // objects, so the compiler can see that epilogue LDS accesses never alias loads in
  // flight and does not drain the prefetch before them.
  extern __shared__ __attribute__((aligned(16))) char smem[];
  __shared__ __attribute__((aligned(16))) char epi[RA_EPI_BYTES];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  char* ring = smem;
  float* sc = reinterpret_cast<float*>(epi + wave * RA_WREG);
  int2* bucket = reinterpret_cast<int2*>(sc);  // [64 rows][RA_BKT] (key bits, column)
  float* thr = reinterpret_cast<float*>(epi + 4 * RA_WREG) + wave * 64;

  const int64_t row0 = (int64_t)blockIdx.x * RA_ROWS;
  const int64_t ntile = npad / RA_BN;
  const int64_t total = ntile * NKC;

  // resident A fragments: rows wave*64 + tm*32 + (lane & 31), k = f*16 + (lane >> 5)*8
  bf16x8 afr[2][NKC * KC / 16];
#pragma unroll
  for (int tm = 0; tm < 2; ++tm) {
    const uint16_t* ar = Xh + (row0 + wave * 64 + tm * 32 + (lane & 31)) * dpad + (lane >> 5) * 8;
#pragma unroll
    for (int f = 0; f < NKC * KC / 16; ++f) afr[tm][f] = *reinterpret_cast<const bf16x8*>(ar + f * 16);
  }
  wait_vm<0>();  // the sweep's vmcnt accounting starts from an empty queue
  float lk[C];  // this lane's row: best C (key, column) pairs, ascending
  int li[C];
#pragma unroll
  for (int q = 0; q < C; ++q) {
    lk[q] = FINF;
    li[q] = -1;
  }
  thr[lane] = FINF;
  int bcnt = 0;  // entries waiting in this lane's row bucket

  // Stage loads, issued in stage order from running pointers: LOADS 16-byte pieces of
  // the 32-column x KC database chunk per thread, and the tile's column scales (4 bytes
  // per lane, lanes 32..63 repeat lanes 0..31) into this wavefront's 256-byte slot.
  // Past the last stage the pointers stop, so the tail reloads the last stage into a
  // slot no later chunk reads: every chunk issues the same number of loads and all
  // vmcnt counts are static.
  const uint16_t* pb[LOADS];
#pragma unroll
  for (int h = 0; h < LOADS; ++h) {
    const int s = (h * 4 + wave) * 64 + lane;
    const int c = s / R::PIECES, qs = s % R::PIECES;
    pb[h] = Xh + (int64_t)c * dpad + (qs ^ (c & 15)) * 8;
  }
  const float* pc = colscale + (lane & 31);
  int64_t gi = 0;   // next stage to issue
  int kci = 0;      // its k-chunk
  int sloti = 0;    // its ring slot
  const int64_t tile_step = (int64_t)RA_BN * dpad - (int64_t)(NKC - 1) * KC;
  auto issue_next = [&]() {
    char* st = ring + sloti * R::STB;
#pragma unroll
    for (int h = 0; h < LOADS; ++h) {
      const uint16_t* gp = pb[h];  // the builtin rejects the captured element itself
      __builtin_amdgcn_global_load_lds(gp, (lds_void*)(st + (h * 4 + wave) * 1024), 16, 0, 0);
    }
    {
      const float* cp = pc;
      __builtin_amdgcn_global_load_lds(cp, (lds_void*)(st + R::STB_B + wave * 256), 4, 0, 0);
    }
    if (gi + 1 < total) {
      if (++kci == NKC) {
        kci = 0;
#pragma unroll
        for (int h = 0; h < LOADS; ++h) pb[h] += tile_step;
        pc += RA_BN;
      } else {
#pragma unroll
        for (int h = 0; h < LOADS; ++h) pb[h] += KC;
      }
    }
    ++gi;
    sloti = (sloti + 1) & (NST - 1);
  };

  for (int g = 0; g < NST - 1; ++g) issue_next();
  for (int64_t ct = 0; ct < ntile; ++ct) {
    f32x16 acc[2] = {(f32x16){}, (f32x16){}};
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) {
      const int64_t g = ct * NKC + kc;
      wait_vm<(LOADS + 1) * (NST - 2)>();  // this thread's loads of stage g landed
      ring_barrier();                       // ... and everyone else's
      const char* st = ring + (int)(g & (NST - 1)) * R::STB;
      bf16x8 bq[KC / 16];  // all B fragments of the stage in flight at once
#pragma unroll
      for (int kk = 0; kk < KC / 16; ++kk)
        bq[kk] = *reinterpret_cast<const bf16x8*>(st + R::slot(lane & 31, kk * 2 + (lane >> 5)) * 16);
#pragma unroll
      for (int kk = 0; kk < KC / 16; ++kk) {
#pragma unroll
        for (int tm = 0; tm < 2; ++tm)
          acc[tm] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afr[tm][kc * (KC / 16) + kk], bq[kk],
                                                           acc[tm], 0, 0, 0);
        if (kk == 0) issue_next();  // stage g + NST - 1, into the slot stage g - 1 left
      }
    }

    // ---------------- wavefront-private epilogue of column tile ct
    // element (tm, e) of this lane: local row tm*32 + (e>>2)*8 + (lane>>5)*4 + (e&3) of
    // the wavefront, column ct*32 + (lane & 31)
    const bool diag = (ct * RA_BN < row0 + RA_ROWS) && (ct * RA_BN + RA_BN > row0);
    const bool tail = (ct + 1) * RA_BN > n;
    // This lane's 32 row thresholds and its column scale, read in one block with its
    // own wait: the compiler, seeing LDS reads while direct-to-LDS loads are in flight,
    // would otherwise wait for the whole prefetch (vmcnt(0)) although these addresses
    // are never written by those loads. The scale comes from the last chunk's stage,
    // complete since that chunk's barrier and not reused before the next chunk.
    f32x4 tq[8];
    float cs;
    {
      const unsigned ta = lds_addr(thr + ((lane >> 5) * 4));
      const unsigned ca = lds_addr(ring + (int)((ct * NKC + NKC - 1) & (NST - 1)) * R::STB +
                                   R::STB_B + wave * 256 + lane * 4);
      asm volatile(
          "ds_read_b128 %0, %9\n"
          "ds_read_b128 %1, %9 offset:32\n"
          "ds_read_b128 %2, %9 offset:64\n"
          "ds_read_b128 %3, %9 offset:96\n"
          "ds_read_b128 %4, %9 offset:128\n"
          "ds_read_b128 %5, %9 offset:160\n"
          "ds_read_b128 %6, %9 offset:192\n"
          "ds_read_b128 %7, %9 offset:224\n"
          "ds_read_b32 %8, %10\n"
          "s_waitcnt lgkmcnt(0)"
          : "=v"(tq[0]), "=v"(tq[1]), "=v"(tq[2]), "=v"(tq[3]), "=v"(tq[4]), "=v"(tq[5]),
            "=v"(tq[6]), "=v"(tq[7]), "=v"(cs)
          : "v"(ta), "v"(ca)
          : "memory");
    }
    // per-lane bases made opaque once per tile, so the per-element indices below are
    // base + compile-time constant rather than values hoisted out of the sweep
    const int lrow = (lane >> 5) * 4;
    const int lcol = lane & 31;
    int dv = (int)(ct * RA_BN - row0) - wave * 64 + lcol - lrow;  // column - row, locally
    int rem = (int)std::min<int64_t>(n - ct * RA_BN, RA_BN) - lcol;
    int sb = lrow * RA_SCS + lcol;
    asm volatile("" : "+v"(dv), "+v"(rem), "+v"(sb));
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const float s_ = acc[tm][e];
        acc[tm][e] = COS ? -(s_ * cs) : (cs - 2.0f * s_);
      }
    if (__builtin_expect(diag || tail, 0)) {  // wave-uniform: the self pair and padding
#pragma unroll
      for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int e = 0; e < 16; ++e)
          if (rem <= 0 || dv == tm * 32 + (e >> 2) * 8 + (e & 3)) acc[tm][e] = FINF;
    }
    // Keys below their row's threshold are appended to that row's LDS bucket, which
    // keeps filling across tiles; the lists take the buckets in one parallel drain once
    // some row holds RA_DRAIN entries, so each drain inserts for many rows at once. Row r
    // of the wavefront sits in one register slot (tm, j, i) of one lane half, one column
    // per lane: a ballot gives each passing key its position after the row's waiting
    // entries (whose count the owner lane r holds). Thresholds only change at drains,
    // which keeps the filter valid: every key left out is >= its row's list end.
    const int cbase = (int)(ct * RA_BN);
    const bool lo_half = lane < 32;
    bool touched = false, overflow = false;
    int inc = 0;  // entries this tile adds to this lane's row
    // the lane id made opaque per tile, so that the compiler keeps no per-slot lane masks
    // across the walk (they would spill scalar registers)
    int lsl = lane;
    asm volatile("" : "+v"(lsl));
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const f32x4 t4 = tq[tm * 4 + j];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float key = acc[tm][j * 4 + i];
          const bool pass = key < t4[i];
          const unsigned long long m = __ballot(pass);
          if (m) {
            touched = true;
            const int r_lo = tm * 32 + j * 8 + i, r_hi = r_lo + 4;
            const unsigned mlo = (unsigned)m, mhi = (unsigned)(m >> 32);
            const int c_lo = __builtin_amdgcn_readlane(bcnt, r_lo);
            const int c_hi = __builtin_amdgcn_readlane(bcnt, r_hi);
            const int n_lo = __popc(mlo), n_hi = __popc(mhi);
            overflow |= (c_lo + n_lo > RA_BKT) | (c_hi + n_hi > RA_BKT);
            const int pos = lo_half ? c_lo + (int)__builtin_amdgcn_mbcnt_lo(mlo, 0u)
                                    : c_hi + (int)__builtin_amdgcn_mbcnt_hi(mhi, 0u);
            const int r = lrow + r_lo;
            if (pass && pos < RA_BKT)
              lds_st_b64(lds_addr(bucket + r * RA_BKT + pos),
                         make_int2(__float_as_int(key), cbase + lcol));  // absolute column
            inc += (lsl == r_lo) ? n_lo : 0;
            inc += (lsl == r_hi) ? n_hi : 0;
          }
        }
      }
    if (touched) {
      asm volatile("" ::: "memory");
      __builtin_amdgcn_wave_barrier();
      bool drain = overflow;
      if (!overflow) {
        bcnt += inc;
        drain = __ballot(bcnt >= RA_DRAIN) != 0;
      }
      if (drain) {
        // every lane inserts its own row's waiting entries: the cost is the largest
        // count, not the sum. Inserts are unconditional (a branch here would be
        // if-converted into a second live copy of the list); the next entry is read
        // while this one is inserted. After an overflow only the entries of earlier
        // tiles are taken here and the current tile is scanned whole below.
        const unsigned ba = lds_addr(bucket + lane * RA_BKT);
        int2 en = lds_ld_b64(ba);
#pragma unroll 1
        for (int e = 0; e < bcnt; ++e) {
          const int2 cur = en;
          en = lds_ld_b64(ba + (unsigned)(e + 1) * 8u);
          cand_insert<C>(lk, li, __int_as_float(cur.x), cur.y);
        }
        bcnt = 0;
        if (overflow) {
          // bucket overflow (the first tiles): dump the key tile over the buckets (all
          // reads above have completed), each lane scans its row
          asm volatile("" ::: "memory");
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int tm = 0; tm < 2; ++tm)
#pragma unroll
            for (int e = 0; e < 16; ++e)
              sc[sb + (tm * 32 + (e >> 2) * 8 + (e & 3)) * RA_SCS] = acc[tm][e];
          asm volatile("" ::: "memory");
          __builtin_amdgcn_wave_barrier();
          const float* srow = sc + lane * RA_SCS;
#pragma unroll 1
          for (int q = 0; q < RA_BN; ++q) cand_insert<C>(lk, li, srow[q], cbase + q);
        }
        asm volatile("" ::: "memory");
        __builtin_amdgcn_wave_barrier();
        lds_st_b32(lds_addr(thr + lane), __float_as_int(lk[C - 1]));
        asm volatile("" ::: "memory");
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  // entries still waiting after the last tile
  if (__ballot(bcnt > 0) != 0) {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_wave_barrier();
    const unsigned ba = lds_addr(bucket + lane * RA_BKT);
#pragma unroll 1
    for (int e = 0; e < bcnt; ++e) {
      const int2 cur = lds_ld_b64(ba + (unsigned)e * 8u);
      cand_insert<C>(lk, li, __int_as_float(cur.x), cur.y);
    }
  }
  wait_vm<0>();  // the tail reloads have landed before the wavefront ends
  const int64_t grow = row0 + tid;
  if (grow < n) {
#pragma unroll
    for (int q = 0; q < C; ++q) {
      cand_key[grow * C + q] = lk[q];
      cand_idx[grow * C + q] = li[q];
    }
  }
}

template __global__ void knn_topc_ra_kernel<32, 1, 128, false>(const uint16_t*, int64_t, int64_t, int,
                                                               const float*, int32_t*, float*);
template __global__ void knn_topc_ra_kernel<32, 1, 128, true>(const uint16_t*, int64_t, int64_t, int,
                                                               const float*, int32_t*, float*);
template __global__ void knn_topc_ra_kernel<32, 1, 256, false>(const uint16_t*, int64_t, int64_t, int,
                                                               const float*, int32_t*, float*);
template __global__ void knn_topc_ra_kernel<32, 1, 256, true>(const uint16_t*, int64_t, int64_t, int,
                                                               const float*, int32_t*, float*);
template __global__ void knn_topc_ra_kernel<32, 2, 256, false>(const uint16_t*, int64_t, int64_t, int,
                                                               const float*, int32_t*, float*);
template __global__ void knn_topc_ra_kernel<32, 2, 256, true>(const uint16_t*, int64_t, int64_t, int,
                                                               const float*, int32_t*, float*);

// ------------------------------------------------------------------ 2c. two waves per SIMD
// bf16-exact inputs, d <= 512, k <= 16, as knn_topc_ra_kernel but with eight 32-row
// wavefronts per 256-row workgroup (two per SIMD). A wavefront's resident A-fragments
// (32 rows x d, 128 registers at d = 512) then fit the VGPRs that the MFMAs read
// directly, and while one wavefront waits (LDS, barrier, epilogue) the other keeps the
// SIMD's matrix core busy. Each B fragment feeds one MFMA, so the ring's LDS reads are
// 1 KB per MFMA. Per tile a wavefront holds one 32 x 32 key tile: lane l has column
// l & 31 of rows j*8 + (l >> 5)*4 + i. Candidates are kept per (row, column half):
// lane l owns row l & 31, columns with bit 4 = l >> 5, in a sorted list of 16, so a
// row has 32 candidates and its bound is the smaller of its two halves' last keys.
constexpr int RB_ROWS = 256;
constexpr int RB_WAVES = 8;
constexpr int RB_BKT = 8;      // bucket slots per (row, half)
constexpr int RB_DRAIN = 4;    // drain once some (row, half) holds this many
#ifndef RB_LA
#define RB_LA 4        // B fragments read from LDS ahead of their MFMAs
#endif
// per-wavefront buckets, or the key tile dumped over them after a bucket overflow
constexpr int RB_WREG = 64 * RB_BKT * 8 > 32 * RA_SCS * 4 ? 64 * RB_BKT * 8 : 32 * RA_SCS * 4;
static_assert(RB_WREG >= 32 * RA_SCS * 4, "the key-tile dump must fit the bucket region");
constexpr int RB_EPI_BYTES = RB_WAVES * RB_WREG + RB_WAVES * 64 * 4 + RB_WAVES * 32 * 4;

// The column side of one launch. The plain sweep sets X = the query rows, self = 1 and
// covers every column tile. The symmetric sweep covers each unordered pair of 256-row
// blocks once: block I takes the cyclic range I, I+1, ... of half the blocks, keeps the
// forward keys (rows of I) in its lists and queues the transposed keys (rows of the
// other block) that fall below that row's threshold thr0, one queue per wavefront. The
// threshold pass runs the plain sweep against a sample of rows (self = 0).
struct RbCols {
  const uint16_t* X;    // column rows, bf16, dpad wide, ntile * 32 of them (zero padded)
  const float* scale;   // their key scales
  float* kt;            // symmetric sweep: (key scale, reverse-candidate threshold thr0)
                        // of every row, interleaved; the sweep reads scales from here
  int64_t ncol;         // valid columns
  int64_t ntile;        // 32-column tiles
  int self;             // columns are the query rows: mask the diagonal
  int32_t* qj;          // symmetric sweep: queued reverse candidates (row j, column r, dot)
  int32_t* qr;
  float* qk;
  int64_t qcap;         // entries per wavefront queue
  int32_t* qcount;      // entries written per wavefront
  int64_t group;        // workgroups resident at once (one per CU): walks start in step
  int block0;           // symmetric sweep launched in slices: first block of this launch
  int nblocks;          // and the whole sweep's block count (0: this launch's grid)
  // A slice with fewer blocks than CUs (the sweep's last) splits each block's walk into
  // `parts` consecutive pieces, one workgroup each (launch index = block * parts + part):
  // part q writes its lists to plist_idx/key + (q * prows + row) * 32 (the slice's rows
  // are 0 .. prows - 1, the lowest blocks) and its queue to slot block0 + launch index.
  int parts;
  int64_t prows;
  int32_t* plist_idx;
  float* plist_key;
};

template <int KC>
struct RbRing {
  static constexpr int STB_B = RA_BN * KC * 2;              // 16 KB at KC = 256
  static constexpr int STB = STB_B + RB_WAVES * 256;        // + column scales per wave
  static constexpr int NST = KC >= 512 ? 3 : 4;              // 32 KB stages: three fit
  static constexpr int PIECES = KC / 8;
  static constexpr int LOADS = STB_B / (64 * RB_WAVES * 16);  // 16-byte loads per thread
  static constexpr size_t SMEM = (size_t)NST * STB;
  static __device__ __forceinline__ int slot(int c, int q) { return c * PIECES + (q ^ (c & 15)); }
};
// symmetric sweep: each wavefront stages its reverse candidates (column, row, dot) in LDS
// and copies them to its global queue in batches, with no vector-memory operation of the
// ring in flight (stores mixed with the ring's loads would void its counted waits)
constexpr int RB_PASS_C = 8;  // candidates per (row, half) kept by the threshold pass
constexpr int RB_KC = 512, RB_NKC = 1;  // d <= 512: one stage holds all of d
// buffer descriptor word 3 for the ring's raw loads (gfx9: 32-bit data format, no swizzle)
constexpr int RB_RSRC_FLAGS = 0x00020000;
// Raw buffer loads straight into LDS (each lane's 16 or 4 bytes land at lds + lane * size):
// base is wave-uniform (the descriptor lives in scalar registers), bytes bounds the range.
__device__ __forceinline__ void buffer_load_lds16(const void* base, int bytes, void* lds, int voff,
                                                  int soff) {
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, RB_RSRC_FLAGS);
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void*)lds, 16, voff, soff, 0, 0);
}
__device__ __forceinline__ void buffer_load_lds4(const void* base, int bytes, void* lds, int voff) {
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, RB_RSRC_FLAGS);
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void*)lds, 4, voff, 0, 0, 0);
}
constexpr int RB_QL = 192;                                      // LDS queue entries per wavefront
constexpr size_t RB_QBYTES = (size_t)RB_WAVES * RB_QL * 12;     // after the ring

// C: candidates kept per (row, column half); the threshold pass needs only RB_THR_RANK
template <int NKC, int KC, bool COS, bool SYM, int C = 16>
__global__ __launch_bounds__(512, 1) void knn_topc_rb_kernel(
    const uint16_t* __restrict__ Xh, int64_t n, int64_t npad, int dpad,
    const float* __restrict__ colscale, int32_t* __restrict__ cand_idx,
    float* __restrict__ cand_key, const RbCols cols) {
  using R = RbRing<KC>;
  constexpr int NST = R::NST;
  constexpr int LOADS = R::LOADS;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  __shared__ __attribute__((aligned(16))) char epi[RB_EPI_BYTES];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  char* ring = smem;
  float* sc = reinterpret_cast<float*>(epi + wave * RB_WREG);
  int2* bucket = reinterpret_cast<int2*>(sc);  // [64 owners][RB_BKT]
  float* thr = reinterpret_cast<float*>(epi + RB_WAVES * RB_WREG) + wave * 64;  // [owner]
  // key scales of this wavefront's 32 rows (the transposed keys of the symmetric sweep)
  float* rsc = reinterpret_cast<float*>(epi + RB_WAVES * RB_WREG + RB_WAVES * 64 * 4) + wave * 32;

  // The symmetric sweep takes its blocks in descending order: the blocks a workgroup
  // queues transposed keys for come after its own, so most of them have finished by then
  // and replaced their rows' thresholds with the tighter largest forward key.
  // (tile and row indices fit 32 bits: n < 2^31; 64-bit products only for addresses)
  const int P = SYM && cols.parts > 1 ? cols.parts : 1;  // pieces of each block's walk
  const int gblk = (int)blockIdx.x / P + cols.block0;  // block of the whole sweep
  const int part = (int)blockIdx.x % P;
  const int qslot = (int)blockIdx.x + cols.block0;  // this workgroup's queues
  const int bid = SYM ? (cols.nblocks > 0 ? cols.nblocks : (int)gridDim.x) - 1 - gblk : gblk;
  const int row0 = bid * RB_ROWS;
  const int ntile = (int)cols.ntile;
  int tile0 = 0, nt = ntile, tshift = 0;
  if (SYM) {
    // Block I owns the cyclic range I, I+1, ... of nblk blocks. The workgroups resident
    // together (consecutive blockIdx, `group` of them: blocks I0 - group + 1 .. I0) start
    // their walks at I0, so that they stream the shared part of their ranges in step and
    // the L2 serves most column reads; each walk then wraps to its own start.
    const int m = (int)(npad / RB_ROWS);  // 256-row blocks; ntile == 8 m
    const int I = bid;
    const int nblk = (m & 1) ? (m + 1) / 2 : m / 2 + (I < m / 2 ? 1 : 0);
    tile0 = I * (RB_ROWS / RA_BN);
    nt = nblk * (RB_ROWS / RA_BN);
    const int g = cols.group > 0 ? (int)cols.group : 1;
    tshift = std::min(gblk % g, nblk - 1) * (RB_ROWS / RA_BN);
  }
  // absolute column tile of the ct-th tile of this walk
  auto tile_at = [&](int ct) {
    int r = tshift + ct;
    if (r >= nt) r -= nt;
    int a = tile0 + r;
    if (a >= ntile) a -= ntile;
    return a;
  };
  // this workgroup's piece [ct0, ct1) of the walk (all of it unless the walk is split)
  const int ct0 = P > 1 ? (int)((int64_t)nt * part / P) : 0;
  const int ct1 = P > 1 ? (int)((int64_t)nt * (part + 1) / P) : nt;
  const int total = (ct1 - ct0) * NKC;
  int qn = 0;  // this wavefront's reverse candidates in its global queue
  int lq = 0;  // and staged in LDS
  const int64_t qbase = SYM ? ((int64_t)qslot * RB_WAVES + wave) * cols.qcap : 0;
  int32_t* const qjw = SYM ? cols.qj + qbase : nullptr;
  int32_t* const qrw = SYM ? cols.qr + qbase : nullptr;
  float* const qkw = SYM ? cols.qk + qbase : nullptr;
  const int qcap = (int)std::min<int64_t>(cols.qcap, INT32_MAX / 2);
  // LDS queue: [RB_QL] columns, [RB_QL] rows, [RB_QL] keys per wavefront
  const unsigned lqa = lds_addr(smem + R::SMEM + (size_t)wave * RB_QL * 12);
  auto flush = [&]() {
    wait_vm<0>();
    for (int e = lane; e < lq; e += 64) {
      const int cj = lds_ld_b32(lqa + 4u * e);
      const int cr = lds_ld_b32(lqa + 4u * (RB_QL + e));
      const int ck = lds_ld_b32(lqa + 4u * (2 * RB_QL + e));
      const int pos = qn + e;
      if (pos < qcap) {
        qjw[pos] = cj;
        qrw[pos] = cr;
        qkw[pos] = __int_as_float(ck);
      }
    }
    wait_vm<0>();
    qn += lq;
    lq = 0;
  };

  bf16x8 afr[NKC * KC / 16];
  {
    const uint16_t* ar = Xh + (int64_t)(row0 + wave * 32 + (lane & 31)) * dpad + (lane >> 5) * 8;
#pragma unroll
    for (int f = 0; f < NKC * KC / 16; ++f) afr[f] = *reinterpret_cast<const bf16x8*>(ar + f * 16);
  }
  wait_vm<0>();
  float lk[C];  // owner list: row lane & 31 of this wavefront, column half lane >> 5
  int li[C];
#pragma unroll
  for (int q = 0; q < C; ++q) {
    lk[q] = FINF;
    li[q] = -1;
  }
  thr[lane] = FINF;
  // the smallest and largest row scale of this wavefront: a cheap bound that rules out
  // transposed candidates for most tiles before any per-row work
  float rs_min = 0.0f, rs_max = 0.0f;
  if (SYM) {
    const int r = row0 + wave * 32 + (lane & 31);
    const float v = r < n ? colscale[r] : 0.0f;
    if (lane < 32) rsc[lane] = v;
    rs_min = v;
    rs_max = v;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      rs_min = fminf(rs_min, __shfl_xor(rs_min, o, 64));
      rs_max = fmaxf(rs_max, __shfl_xor(rs_max, o, 64));
    }
    rs_min = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(rs_min)));
    rs_max = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(rs_max)));
  }
  int bcnt = 0;

  // Stage loads are buffer loads straight into LDS: a descriptor per column tile (its base
  // is wave-uniform, so it lives in scalar registers) and fixed per-lane byte offsets,
  // 32-bit since a tile spans 32 * dpad * 2 bytes. (The global-address form of the same
  // load is counted by the compiler as a possible LDS access outstanding in lgkmcnt,
  // which would turn every wait for a B fragment below into lgkmcnt(0).)
  int voff[LOADS];
#pragma unroll
  for (int h = 0; h < LOADS; ++h) {
    const int s = (h * RB_WAVES + wave) * 64 + lane;
    const int c = s / R::PIECES, qs = s % R::PIECES;
    voff[h] = (c * dpad + (qs ^ (c & 15)) * 8) * 2;
  }
  // per tile, the even lanes bring the 32 column scales and the odd lanes (symmetric
  // sweep) the columns' reverse thresholds into the ring, so column c's pair (scale,
  // threshold) sits in 8 consecutive bytes; without thresholds both lanes load the scale
  const int soff = SYM ? lane * 4 : (lane >> 1) * 4;
  const int tile_bytes = RA_BN * dpad * 2;
  int gi = 0, tcur = ct0;  // walk index of the tile being issued
  int tabs = tile_at(ct0);  // its absolute column tile
  int kci = 0, sloti = 0;
  auto issue_next = [&]() {
    char* st = ring + sloti * R::STB;
#pragma unroll
    for (int h = 0; h < LOADS; ++h)
      buffer_load_lds16(cols.X + (int64_t)tabs * (RA_BN * dpad), tile_bytes,
                        st + (h * RB_WAVES + wave) * 1024, voff[h], kci * KC * 2);
    buffer_load_lds4(SYM ? cols.kt + (int64_t)tabs * (2 * RA_BN) : cols.scale + (int64_t)tabs * RA_BN,
                     SYM ? 8 * RA_BN : 4 * RA_BN, st + R::STB_B + wave * 256, soff);
    // past the last stage the tile stays put: the tail reloads the last stage into a slot
    // no later chunk reads, so every chunk issues the same loads and the waits are static
    if (gi + 1 < total) {
      if (++kci == NKC) {
        kci = 0;
        ++tcur;
        tabs = SYM ? tile_at(tcur) : tcur;
      }
    }
    ++gi;
    sloti = sloti + 1 == NST ? 0 : sloti + 1;
  };

  const int lcol = lane & 31, half = lcol >> 4, grp = lane >> 4;
  // the (row, half) list this lane owns is counted in slot ((lane & 31) >> 3) * 4 + (lane & 3)
  // of the tile epilogue, in lane group (((lane & 31) >> 2) & 1) * 2 + (lane >> 5) of that
  // slot's ballot
  const int myslot = (lcol >> 3) * 4 + (lane & 3);
  const int gsh = 16 * (((lcol >> 2) & 1) * 2 + (lane >> 5));
  const int lrow = (lane >> 5) * 4;
  // B piece q = 2*kk + (lane >> 5) of column lcol sits at slot lcol*PIECES + (q ^ (lcol &
  // 15)) = lcol*PIECES + (2*kk ^ bm)
  const int bcol = lcol * R::PIECES * 16, bm = (lane >> 5) ^ (lcol & 15);
  // The ring's barrier sits after each stage's MFMAs, before the tile epilogue: a
  // wavefront passes it once its own loads of the next stage have landed and every read
  // of this stage has completed (each was waited for before the MFMA that used it). So
  // past the barrier the next stage is whole, the slot of the stage just read may be
  // refilled, and a wavefront still in its epilogue delays the others only when it is
  // more than a stage behind.
  for (int g = 0; g < NST - 1; ++g) issue_next();
  wait_vm<(LOADS + 1) * (NST - 2)>();  // this wavefront's loads of stage 0 have landed
  ring_barrier();
  int slotc = 0;  // ring slot of the stage being read
  for (int ct = ct0; ct < ct1; ++ct) {
    f32x16 acc = (f32x16){};
    float2 csv;  // the column's key scale and (symmetric sweep) reverse threshold
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) {
      const char* st = ring + slotc * R::STB + bcol;
      // RB_LA reads in flight ahead of the MFMAs: the first RB_LA fragments at once, then
      // each next one as soon as an MFMA frees its register
      bf16x8 bq[RB_LA];
#pragma unroll
      for (int u = 0; u < RB_LA; ++u) bq[u] = *reinterpret_cast<const bf16x8*>(st + ((2 * u) ^ bm) * 16);
#pragma unroll
      for (int kk = 0; kk < KC / 16; ++kk) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afr[kc * (KC / 16) + kk], bq[kk % RB_LA],
                                                      acc, 0, 0, 0);
        if (kk == 0) issue_next();  // stage g + NST - 1, into the slot of stage g - 1
        if (kk + RB_LA < KC / 16)
          bq[kk % RB_LA] = *reinterpret_cast<const bf16x8*>(st + ((2 * (kk + RB_LA)) ^ bm) * 16);
      }
      // order: the first reads, then MFMA and read alternating, then the last MFMAs
      __builtin_amdgcn_sched_group_barrier(0x100, RB_LA, 0);
#pragma unroll
      for (int kk = 0; kk + RB_LA < KC / 16; ++kk) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x008, RB_LA, 0);
      if (kc == NKC - 1) {
        // the column's scale and threshold, from the tile's last stage before its barrier
        // (past it the slot may be refilled); written as asm with its own wait, since the
        // compiler would wait for every load in flight before a plain read of the ring
        const unsigned ca = lds_addr(ring + slotc * R::STB + R::STB_B + wave * 256 + lcol * 8);
        asm volatile("ds_read_b64 %0, %1\ns_waitcnt lgkmcnt(0)" : "=v"(csv) : "v"(ca) : "memory");
      }
      wait_vm<(LOADS + 1) * (NST - 2)>();  // own loads of stage g + 1 landed
      ring_barrier();
      slotc = slotc + 1 == NST ? 0 : slotc + 1;
    }

    // ---------------- wavefront-private epilogue of column tile ct
    const int cta = SYM ? tile_at(ct) : ct;  // absolute column tile
    const bool diag = cols.self && (cta * RA_BN < row0 + RB_ROWS) && (cta * RA_BN + RA_BN > row0);
    const bool tail = (cta + 1) * RA_BN > cols.ncol;
    const float cs = csv.x;
    float t0c = csv.y;
    if (SYM && (cta >> 3) != bid) {
      // The largest of this lane's 16 dots (and, below, of each group of four rows). Keys fall as
      // dots rise (-(s c) for cosine with c >= 0, c - 2 s for squared Euclidean) and f32
      // rounding is monotone, so every key formed from a dot is at least the key formed
      // from the maximum.
      float smax;
      {
        const float m0 = fmaxf(fmaxf(acc[0], acc[1]), acc[2]), m1 = fmaxf(fmaxf(acc[3], acc[4]), acc[5]);
        const float m2 = fmaxf(fmaxf(acc[6], acc[7]), acc[8]), m3 = fmaxf(fmaxf(acc[9], acc[10]), acc[11]);
        const float m4 = fmaxf(fmaxf(acc[12], acc[13]), acc[14]);
        smax = fmaxf(fmaxf(m0, m1), fmaxf(fmaxf(m2, m3), fmaxf(m4, acc[15])));
      }
      // Transposed keys: column cta*32 + lcol is a row too, and this wavefront's rows are
      // its candidates; a (row, column) pair is queued when its key, formed with the row's
      // scale, beats the column's threshold. Bounds filter first: for dots s <= ms and any
      // row scale r in [rs_min, rs_max], -(s r) >= -max(ms rs_min, ms rs_max) and
      // r - 2 s >= rs_min - 2 ms (f32 rounding is monotone, so they hold as computed): the
      // whole tile with ms = smax, then, for cosine keys, each group of four rows. The
      // queue keeps the raw dot; the key is formed again, identically, when it is
      // scattered.
      const int cabs = (int)(cta * RA_BN) + lcol;
      if (!((int64_t)cabs < n)) t0c = -FINF;
      const float rkmin = COS ? -fmaxf(smax * rs_min, smax * rs_max) : (rs_min - 2.0f * smax);
      unsigned m16 = 0u;
      if (__ballot(rkmin < t0c) != 0) {
        m16 = 0xffffu;
        if (COS) {  // per row group of 4 slots: the bound at the group's largest dot
          m16 = 0;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float gm = fmaxf(fmaxf(acc[4 * j], acc[4 * j + 1]), fmaxf(acc[4 * j + 2], acc[4 * j + 3]));
            m16 |= (-fmaxf(gm * rs_min, gm * rs_max) < t0c ? 15u : 0u) << (4 * j);
          }
        }
        if (row0 + RB_ROWS > n) {  // the last block: rows past n are padding
#pragma unroll
          for (int e = 0; e < 16; ++e)
            if (!(row0 + wave * 32 + (e >> 2) * 8 + lrow + (e & 3) < n)) m16 &= ~(1u << e);
        }
      }
      if (__ballot(m16 != 0) != 0) {
        const int rbase = (int)(row0 + wave * 32) + lrow;
#pragma unroll 1
        for (int j = 0; j < 4; ++j) {
          if (__ballot(((m16 >> (4 * j)) & 15u) != 0) == 0) continue;
          // flagged slots of this row group: the exact key with the row's own scale
          f32x4 rs;
          {
            const unsigned ra_ = lds_addr(rsc + lrow + j * 8);
            asm volatile("ds_read_b128 %0, %1\ns_waitcnt lgkmcnt(0)" : "=v"(rs) : "v"(ra_) : "memory");
          }
          float sj[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) sj[i] = acc[j * 4 + i];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float rk = COS ? -(sj[i] * rs[i]) : (rs[i] - 2.0f * sj[i]);
            const bool rp = ((m16 >> (4 * j + i)) & 1u) && (rk < t0c);
            const unsigned long long rm = __ballot(rp);
            if (rm) {
              if (lq > RB_QL - 64) flush();
              const int below = (int)__builtin_amdgcn_mbcnt_hi(
                  (unsigned)(rm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)rm, 0u));
              const unsigned q = (unsigned)(lq + below);
              if (rp) {
                lds_st_b32(lqa + 4u * q, cabs);
                lds_st_b32(lqa + 4u * (RB_QL + q), rbase + j * 8 + i);
                lds_st_b32(lqa + 4u * (2 * RB_QL + q), __float_as_int(sj[i]));
              }
              lq += __popcll(rm);
            }
          }
        }
      }
    }
    // Forward keys, each against its own (row, half) threshold (a bound over several rows
    // lets almost every tile through on clustered data: some row of any 16 has a loose
    // threshold while another row's cluster-mates give large dots).
    f32x4 tq[4];  // thresholds of this lane's 16 rows in its column half (changed only at drains)
#pragma unroll
    for (int j = 0; j < 4; ++j) tq[j] = reinterpret_cast<const f32x4*>(thr + 32 * half + lrow)[2 * j];
    int dv = (int)(cta * RA_BN - row0) - wave * 32 + lcol - lrow;
    int rem = (int)std::min<int64_t>(cols.ncol - cta * RA_BN, RA_BN) - lcol;
    int sb = lrow * RA_SCS + lcol;
    asm volatile("" : "+v"(dv), "+v"(rem), "+v"(sb));
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const float s_ = acc[e];
      acc[e] = COS ? -(s_ * cs) : (cs - 2.0f * s_);
    }
    if (__builtin_expect(diag || tail, 0)) {
#pragma unroll
      for (int e = 0; e < 16; ++e)
        if (rem <= 0 || dv == (e >> 2) * 8 + (e & 3)) acc[e] = FINF;
    }
    {
      unsigned long long anyp = 0;
#pragma unroll
      for (int e = 0; e < 16; ++e) anyp |= __ballot(acc[e] < tq[e >> 2][e & 3]);
      if (anyp == 0) continue;  // nothing in this tile enters a list
    }
    const int cbase = (int)(cta * RA_BN);
    bool touched = false, overflow = false;
    // the ballot of the slot this lane's list is counted in; the slot id is made opaque per
    // tile, so that the compiler keeps no per-slot lane masks across the walk (they would
    // spill scalar registers)
    unsigned mlo = 0u, mhi = 0u;
    int lsl = myslot;
    asm volatile("" : "+v"(lsl));
    // owner of this lane's (row, half) group in slot (j, i): row j*8+i (+4 for lanes >= 32),
    // half (lane >> 4) & 1; its bucket is ob + (j*8+i) * RB_BKT * 8 bytes
    int ob = (int)lds_addr(bucket) + ((grp & 1) * 32 + (grp >> 1) * 4) * RB_BKT * 8;
    asm volatile("" : "+v"(ob));
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float key = acc[j * 4 + i];
        const bool pass = key < tq[j][i];
        const unsigned long long m = __ballot(pass);
        if (m) {
          touched = true;
          // the 4 lane groups of 16 are the (row, half) pairs (r0, 0), (r0, 1),
          // (r0 + 4, 0), (r0 + 4, 1); owner lane of a pair = row + 32 * half
          const int r0 = j * 8 + i;
          const int o0 = r0, o1 = r0 + 32, o2 = r0 + 4, o3 = r0 + 36;
          const int c0 = __builtin_amdgcn_readlane(bcnt, o0);
          const int c1 = __builtin_amdgcn_readlane(bcnt, o1);
          const int c2 = __builtin_amdgcn_readlane(bcnt, o2);
          const int c3 = __builtin_amdgcn_readlane(bcnt, o3);
          const int n0 = __popcll(m & 0xffffull), n1 = __popcll(m & 0xffff0000ull);
          const int n2 = __popcll(m & 0xffff00000000ull), n3 = __popcll(m & 0xffff000000000000ull);
          overflow |= (c0 + n0 > RB_BKT) | (c1 + n1 > RB_BKT) | (c2 + n2 > RB_BKT) |
                      (c3 + n3 > RB_BKT);
          const unsigned long long gm = m & (0xffffull << (16 * grp));
          const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(gm >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((unsigned)gm, 0u));
          const int base = grp == 0 ? c0 : (grp == 1 ? c1 : (grp == 2 ? c2 : c3));
          const int pos = base + below;
          if (pass && pos < RB_BKT)
            lds_st_b64((unsigned)(ob + r0 * RB_BKT * 8 + pos * 8),
                       make_int2(__float_as_int(key), cbase + lcol));
          if (lsl == j * 4 + i) {
            mlo = (unsigned)m;
            mhi = (unsigned)(m >> 32);
          }
        }
      }
    if (touched) {
      asm volatile("" ::: "memory");
      __builtin_amdgcn_wave_barrier();
      bool drain = overflow;
      if (!overflow) {
        const unsigned long long mine = ((unsigned long long)mhi << 32) | mlo;
        bcnt += (int)__popcll((mine >> gsh) & 0xffffull);
        drain = __ballot(bcnt >= RB_DRAIN) != 0;
      }
      if (drain) {
        const unsigned ba = lds_addr(bucket + lane * RB_BKT);
        int2 en = lds_ld_b64(ba);
#pragma unroll 1
        for (int e = 0; e < bcnt; ++e) {
          const int2 cur = en;
          en = lds_ld_b64(ba + (unsigned)(e + 1) * 8u);
          cand_insert<C>(lk, li, __int_as_float(cur.x), cur.y);
        }
        bcnt = 0;
        if (overflow) {
          asm volatile("" ::: "memory");
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int e = 0; e < 16; ++e) sc[sb + ((e >> 2) * 8 + (e & 3)) * RA_SCS] = acc[e];
          asm volatile("" ::: "memory");
          __builtin_amdgcn_wave_barrier();
          // owner lane: row lane & 31, columns 16 * (lane >> 5) .. + 15
          const float* srow = sc + (lane & 31) * RA_SCS + 16 * (lane >> 5);
#pragma unroll 1
          for (int q = 0; q < 16; ++q) cand_insert<C>(lk, li, srow[q], cbase + 16 * (lane >> 5) + q);
        }
        asm volatile("" ::: "memory");
        __builtin_amdgcn_wave_barrier();
        lds_st_b32(lds_addr(thr + lane), __float_as_int(lk[C - 1]));
        asm volatile("" ::: "memory");
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  if (__ballot(bcnt > 0) != 0) {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_wave_barrier();
    const unsigned ba = lds_addr(bucket + lane * RB_BKT);
#pragma unroll 1
    for (int e = 0; e < bcnt; ++e) {
      const int2 cur = lds_ld_b64(ba + (unsigned)e * 8u);
      cand_insert<C>(lk, li, __int_as_float(cur.x), cur.y);
    }
  }
  wait_vm<0>();
  // list of (row, half) -> candidates (row * 2 + half) * 16
  const int64_t grow = row0 + wave * 32 + (lane & 31);
  if (SYM) {
    if (lq > 0) flush();
    if (lane == 0) cols.qcount[(int64_t)qslot * RB_WAVES + wave] = qn;
    // Publish the row's largest forward key as its threshold when it is smaller. Any
    // transposed key at or above it is beaten by all 32 forward candidates; workgroups
    // still to come read the smaller value (or, from a stale cache line, the old one:
    // every value a threshold takes is an upper bound the merge accounts for).
    const float fwdmax = fmaxf(lk[C - 1], __shfl_xor(lk[C - 1], 32, 64));
    if (lane < 32 && grow < n && fwdmax < cols.kt[2 * grow + 1]) cols.kt[2 * grow + 1] = fwdmax;
  }
  if (grow < n) {
    // a piece of a split walk writes its partial lists (the same layout, C = 16) to its
    // part's rows of the piece lists
    float* const okey = P > 1 ? cols.plist_key + part * cols.prows * 2 * C : cand_key;
    int32_t* const oidx = P > 1 ? cols.plist_idx + part * cols.prows * 2 * C : cand_idx;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      okey[(grow * 2 + (lane >> 5)) * C + q] = lk[q];
      oidx[(grow * 2 + (lane >> 5)) * C + q] = li[q];
    }
  }
}

template __global__ void knn_topc_rb_kernel<RB_NKC, RB_KC, true, false, RB_PASS_C>(
    const uint16_t*, int64_t, int64_t, int, const float*, int32_t*, float*, const RbCols);
template __global__ void knn_topc_rb_kernel<RB_NKC, RB_KC, false, false, RB_PASS_C>(
    const uint16_t*, int64_t, int64_t, int, const float*, int32_t*, float*, const RbCols);
template __global__ void knn_topc_rb_kernel<RB_NKC, RB_KC, true, false>(const uint16_t*, int64_t, int64_t,
                                                                 int, const float*, int32_t*,
                                                                 float*, const RbCols);
template __global__ void knn_topc_rb_kernel<RB_NKC, RB_KC, false, false>(const uint16_t*, int64_t, int64_t,
                                                                  int, const float*, int32_t*,
                                                                  float*, const RbCols);
template __global__ void knn_topc_rb_kernel<RB_NKC, RB_KC, true, true>(const uint16_t*, int64_t, int64_t,
                                                                int, const float*, int32_t*,
                                                                float*, const RbCols);
template __global__ void knn_topc_rb_kernel<RB_NKC, RB_KC, false, true>(const uint16_t*, int64_t, int64_t,
                                                                 int, const float*, int32_t*,
                                                                 float*, const RbCols);

// ------------------------------------------------------------------ 2c. resident, 1/SIMD
// One wavefront of 32 query rows per SIMD (four per 128-row workgroup), with the rows'
// A-fragments for all of d resident in the unified register file (up to 256 registers of
// the 512 a single wavefront may hold): split inputs (hi*hi + hi*lo + lo*hi of x = hi + lo,
// bf16 each, as the streaming kernel forms them; d <= 512, hi and lo both resident) and
// bf16 rows with 512 < d <= 1024. Each B fragment read from LDS feeds three MFMAs (split)
// or the tile holds 48-64 k-steps (d > 512): the matrix work per epilogue key is 1.5-3
// times that of the two-wavefront-per-SIMD kernel, which lets a single wavefront keep the
// matrix core busy through its own epilogue. Column tiles of 32 stream through a ring of
// four stages of 256 dimensions; the tile epilogue (per-value filters, buckets, sorted
// lists of C per row and column half) is the bf16 kernel's plain-sweep epilogue. Query
// rows are the column rows themselves (qmap null: self-join, the diagonal masked) or
// gathered rows with their global indices in qmap (ascending; the re-search of rows whose
// certificate failed), over the column tiles [tc0, tc1).
constexpr int RS_WAVES = 4;
constexpr int RS_ROWS = RS_WAVES * 32;
constexpr int RS_KC = 256;
constexpr int RS_LA = 4;  // B fragments (pairs, split) read ahead of their MFMAs
template <bool SPL>
struct RsRing {
  static constexpr int STB_H = RA_BN * RS_KC * 2;            // 16 KB: hi pieces of a stage
  static constexpr int STB_B = SPL ? 2 * STB_H : STB_H;       // hi and lo
  static constexpr int STB = STB_B + RS_WAVES * 256;          // + column scales per wave
  static constexpr int NST = 4;
  static constexpr int PIECES = RS_KC / 8;
  static constexpr int HLOADS = STB_H / (64 * RS_WAVES * 16);  // 16-byte loads per thread, hi
  static constexpr size_t SMEM = (size_t)NST * STB;
};
static_assert(RsRing<true>::HLOADS == 4, "four hi (and four lo) loads per thread and stage");
// buckets and key-tile dumps, thresholds, and (symmetric sweep) the rows' key scales
constexpr int RS_EPI_BYTES = RS_WAVES * RB_WREG + RS_WAVES * 64 * 4 + RS_WAVES * 32 * 4;
// symmetric sweep: LDS queue entries (column, row, dot) per wavefront, after the ring
constexpr int RS_QL = 160;
constexpr size_t RS_QBYTES = (size_t)RS_WAVES * RS_QL * 12;
static_assert(RsRing<true>::SMEM + RS_QBYTES + RS_EPI_BYTES <= 160 * 1024,
              "the split symmetric sweep's LDS fits one CU");

// SYM: the symmetric sweep of a self-join (qmap null, column tiles [0, tc1) = all rows):
// 128-row block I takes the cyclic range I, I+1, ... of half the blocks, as the
// two-wavefront kernel's sweep does with 256-row blocks (see RbCols: thresholds in kt,
// transposed keys queued per wavefront, the largest forward key published at the end).
// self = 0 (qmap null): the columns are other rows than the queries (the threshold pass
// against a sample); nothing but the column padding is masked.
template <int NKC, bool COS, bool SPL, int C, bool SYM = false>
__global__ __launch_bounds__(64 * RS_WAVES, 1) void knn_topc_rs_kernel(
    const uint16_t* __restrict__ Xh, const uint16_t* __restrict__ Xl,
    const uint16_t* __restrict__ Qh, const uint16_t* __restrict__ Ql,
    const int32_t* __restrict__ qmap, int64_t nq, int64_t n, int dpad,
    const float* __restrict__ colscale, int tc0, int tc1, int self,
    int32_t* __restrict__ cand_idx, float* __restrict__ cand_key, const RbCols cols) {
  using R = RsRing<SPL>;
  constexpr int NST = R::NST;
  constexpr int HL = R::HLOADS;
  constexpr int LOADS = (SPL ? 2 : 1) * HL;  // vector-memory operations of a stage: + 1 scale
  constexpr int KS = RS_KC / 16;             // MFMA k-steps per stage
  constexpr int TPB = RS_ROWS / RA_BN;       // column tiles per 128-row block
  static_assert(!SYM || C == 16, "the symmetric sweep's merge takes half-lists of 16");
  extern __shared__ __attribute__((aligned(16))) char smem[];
  __shared__ __attribute__((aligned(16))) char epi[RS_EPI_BYTES];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  char* ring = smem;
  float* sc = reinterpret_cast<float*>(epi + wave * RB_WREG);
  int2* bucket = reinterpret_cast<int2*>(sc);  // [64 owners][RB_BKT]
  float* thr = reinterpret_cast<float*>(epi + RS_WAVES * RB_WREG) + wave * 64;  // [owner]
  float* rsc = reinterpret_cast<float*>(epi + RS_WAVES * RB_WREG + RS_WAVES * 64 * 4) + wave * 32;
  // blocks in descending order in the symmetric sweep (the blocks a workgroup queues
  // transposed keys for come after its own and have mostly tightened their thresholds)
  const int P = SYM && cols.parts > 1 ? cols.parts : 1;  // pieces of each block's walk
  const int gblk = (int)blockIdx.x / P + (SYM ? cols.block0 : 0);
  const int part = (int)blockIdx.x % P;
  const int qslot = (int)blockIdx.x + (SYM ? cols.block0 : 0);  // this workgroup's queues
  const int bid = SYM ? (cols.nblocks > 0 ? cols.nblocks : (int)gridDim.x) - 1 - gblk : gblk;
  const int row0 = bid * RS_ROWS;  // first query (list) row of the workgroup
  if (!SYM && gridDim.y > 1) {
    // column shards: workgroup row y takes the y-th share of [tc0, tc1) and writes its
    // lists at y * nq * 2C
    const int all = tc1 - tc0, sh = (int)blockIdx.y, ns = (int)gridDim.y;
    const int a = tc0 + (int)((int64_t)all * sh / ns), b = tc0 + (int)((int64_t)all * (sh + 1) / ns);
    tc0 = a;
    tc1 = b;
    cand_idx += (int64_t)sh * nq * 2 * C;
    cand_key += (int64_t)sh * nq * 2 * C;
  }
  const int ntile = tc1;
  int tile0 = tc0, nt = tc1 - tc0, tshift = 0;
  if (SYM) {
    // the workgroups resident together (`group` consecutive launch indices) start their
    // walks at the same block, so that the L2 serves most of their column reads
    const int m = ntile / TPB;
    const int nblk = (m & 1) ? (m + 1) / 2 : m / 2 + (bid < m / 2 ? 1 : 0);
    tile0 = bid * TPB;
    nt = nblk * TPB;
    const int g = cols.group > 0 ? (int)cols.group : 1;
    tshift = std::min(gblk % g, nblk - 1) * TPB;
  }
  // absolute column tile of the ct-th tile of this walk
  auto tile_at = [&](int ct) {
    if (!SYM) return tc0 + ct;
    int r = tshift + ct;
    if (r >= nt) r -= nt;
    int a = tile0 + r;
    if (a >= ntile) a -= ntile;
    return a;
  };
  // this workgroup's piece [ct0, ct1) of the walk (all of it unless the walk is split)
  const int ct0 = P > 1 ? (int)((int64_t)nt * part / P) : 0;
  const int ct1 = P > 1 ? (int)((int64_t)nt * (part + 1) / P) : nt;
  const int total = (ct1 - ct0) * NKC;
  int qn = 0;  // this wavefront's reverse candidates in its global queue
  int lq = 0;  // and staged in LDS
  const int64_t qbase = SYM ? ((int64_t)qslot * RS_WAVES + wave) * cols.qcap : 0;
  int32_t* const qjw = SYM ? cols.qj + qbase : nullptr;
  int32_t* const qrw = SYM ? cols.qr + qbase : nullptr;
  float* const qkw = SYM ? cols.qk + qbase : nullptr;
  const int qcap = SYM ? (int)std::min<int64_t>(cols.qcap, INT32_MAX / 2) : 0;
  // LDS queue: [RS_QL] columns, [RS_QL] rows, [RS_QL] dots per wavefront; copied to the
  // global queue with every load of the ring landed and the stores waited for, so that the
  // ring's counted waits see ring loads only
  const unsigned lqa = lds_addr(smem + R::SMEM + (size_t)wave * RS_QL * 12);
  auto flush = [&]() {
    wait_vm<0>();
    for (int e = lane; e < lq; e += 64) {
      const int cj = lds_ld_b32(lqa + 4u * e);
      const int cr = lds_ld_b32(lqa + 4u * (RS_QL + e));
      const int ck = lds_ld_b32(lqa + 4u * (2 * RS_QL + e));
      const int pos = qn + e;
      if (pos < qcap) {
        qjw[pos] = cj;
        qrw[pos] = cr;
        qkw[pos] = __int_as_float(ck);
      }
    }
    wait_vm<0>();
    qn += lq;
    lq = 0;
  };

  bf16x8 afh[NKC * KS];
  bf16x8 afl[SPL ? NKC * KS : 1];
  {
    const int64_t ro = (int64_t)(row0 + wave * 32 + (lane & 31)) * dpad + (lane >> 5) * 8;
#pragma unroll
    for (int f = 0; f < NKC * KS; ++f) {
      afh[f] = *reinterpret_cast<const bf16x8*>(Qh + ro + f * 16);
      if (SPL) afl[SPL ? f : 0] = *reinterpret_cast<const bf16x8*>(Ql + ro + f * 16);
    }
  }
  // gathered rows: the global indices of the 16 rows whose values this lane holds (the
  // diagonal mask) and the wavefront's range of them
  int qg[16];
  int qlo = 0, qhi = -1;
  if (qmap) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int64_t r = row0 + wave * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
      qg[e] = r < nq ? qmap[r] : -1;
    }
    const int64_t rf = row0 + wave * 32, rl = std::min<int64_t>(rf + 31, nq - 1);
    if (rf < nq) {
      qlo = qmap[rf];
      qhi = qmap[rl];
    }
  }
  wait_vm<0>();
  float lk[C];
  int li[C];
#pragma unroll
  for (int q = 0; q < C; ++q) {
    lk[q] = FINF;
    li[q] = -1;
  }
  thr[lane] = FINF;
  // the smallest and largest row scale of this wavefront: a bound that rules out the
  // transposed keys of most tiles before any per-row work
  float rs_min = 0.0f, rs_max = 0.0f;
  if (SYM) {
    const int64_t r = row0 + wave * 32 + (lane & 31);
    const float v = r < n ? colscale[r] : 0.0f;
    if (lane < 32) rsc[lane] = v;
    rs_min = v;
    rs_max = v;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      rs_min = fminf(rs_min, __shfl_xor(rs_min, o, 64));
      rs_max = fmaxf(rs_max, __shfl_xor(rs_max, o, 64));
    }
    rs_min = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(rs_min)));
    rs_max = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(rs_max)));
  }
  int bcnt = 0;

  // buffer loads straight into LDS: loads 0..HL-1 of a thread bring hi pieces, HL..2HL-1
  // (split) the lo pieces at the same offsets; even and odd lanes both bring their
  // column's scale, or (symmetric sweep) the even lanes its scale and the odd lanes its
  // reverse threshold
  int voff[HL];
#pragma unroll
  for (int h = 0; h < HL; ++h) {
    const int s_ = (h * RS_WAVES + wave) * 64 + lane;
    const int c = s_ / R::PIECES, qs = s_ % R::PIECES;
    voff[h] = (c * dpad + (qs ^ (c & 15)) * 8) * 2;
  }
  const int soff = SYM ? lane * 4 : (lane >> 1) * 4;
  const int tile_bytes = RA_BN * dpad * 2;
  int gi = 0, tcur = ct0, kci = 0, sloti = 0;
  int tabs = tile_at(ct0);  // absolute column tile being issued
  auto issue_next = [&]() {
    char* st = ring + sloti * R::STB;
    const int64_t tb = (int64_t)tabs * (RA_BN * dpad);
#pragma unroll
    for (int h = 0; h < HL; ++h) {
      buffer_load_lds16(Xh + tb, tile_bytes, st + (h * RS_WAVES + wave) * 1024, voff[h],
                        kci * RS_KC * 2);
      if (SPL)
        buffer_load_lds16(Xl + tb, tile_bytes, st + R::STB_H + (h * RS_WAVES + wave) * 1024,
                          voff[h], kci * RS_KC * 2);
    }
    if (SYM)
      buffer_load_lds4(cols.kt + (int64_t)tabs * (2 * RA_BN), 8 * RA_BN, st + R::STB_B + wave * 256,
                       soff);
    else
      buffer_load_lds4(colscale + (int64_t)tabs * RA_BN, 4 * RA_BN, st + R::STB_B + wave * 256,
                       soff);
    // past the last stage the tile stays put: the tail reloads the last stage into a slot
    // no later chunk reads, so every chunk issues the same loads and the waits are static
    if (gi + 1 < total) {
      if (++kci == NKC) {
        kci = 0;
        ++tcur;
        tabs = tile_at(tcur);
      }
    }
    ++gi;
    sloti = sloti + 1 == NST ? 0 : sloti + 1;
  };

  const int lcol = lane & 31, half = lcol >> 4, grp = lane >> 4;
  // the (row, half) list this lane owns is counted in slot ((lane & 31) >> 3) * 4 + (lane & 3)
  // of the tile epilogue, in lane group (((lane & 31) >> 2) & 1) * 2 + (lane >> 5) of that
  // slot's ballot
  const int myslot = (lcol >> 3) * 4 + (lane & 3);
  const int gsh = 16 * (((lcol >> 2) & 1) * 2 + (lane >> 5));
  const int lrow = (lane >> 5) * 4;
  const int bcol = lcol * R::PIECES * 16, bm = (lane >> 5) ^ (lcol & 15);
  // an empty column range (a shard with no tile) issues no load: its lists stay empty
  if (ct1 > ct0) {
    for (int g = 0; g < NST - 1; ++g) issue_next();
    wait_vm<(LOADS + 1) * (NST - 2)>();  // this wavefront's loads of stage 0 have landed
    ring_barrier();
  }
  int slotc = 0;
  for (int ct = ct0; ct < ct1; ++ct) {
    f32x16 acc = (f32x16){};
    float2 csv;
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) {
      const char* st = ring + slotc * R::STB + bcol;
      bf16x8 bh[RS_LA], bl[SPL ? RS_LA : 1];
#pragma unroll
      for (int u = 0; u < RS_LA; ++u) {
        bh[u] = *reinterpret_cast<const bf16x8*>(st + ((2 * u) ^ bm) * 16);
        if (SPL) bl[SPL ? u : 0] = *reinterpret_cast<const bf16x8*>(st + R::STB_H + ((2 * u) ^ bm) * 16);
      }
#pragma unroll
      for (int kk = 0; kk < KS; ++kk) {
        const int f = kc * KS + kk;
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afh[f], bh[kk % RS_LA], acc, 0, 0, 0);
        if (SPL) {
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afh[f], bl[SPL ? kk % RS_LA : 0], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afl[SPL ? f : 0], bh[kk % RS_LA], acc, 0, 0, 0);
        }
        if (kk == 0) issue_next();  // stage g + NST - 1, into the slot of stage g - 1
        if (kk + RS_LA < KS) {
          bh[kk % RS_LA] = *reinterpret_cast<const bf16x8*>(st + ((2 * (kk + RS_LA)) ^ bm) * 16);
          if (SPL)
            bl[SPL ? kk % RS_LA : 0] =
                *reinterpret_cast<const bf16x8*>(st + R::STB_H + ((2 * (kk + RS_LA)) ^ bm) * 16);
        }
      }
      if (kc == NKC - 1) {
        const unsigned ca = lds_addr(ring + slotc * R::STB + R::STB_B + wave * 256 + lcol * 8);
        asm volatile("ds_read_b64 %0, %1\ns_waitcnt lgkmcnt(0)" : "=v"(csv) : "v"(ca) : "memory");
      }
      wait_vm<(LOADS + 1) * (NST - 2)>();  // own loads of stage g + 1 landed
      ring_barrier();
      slotc = slotc + 1 == NST ? 0 : slotc + 1;
    }

    // ---------------- wavefront-private epilogue of the walk's ct-th column tile
    const int cta = tile_at(ct);
    const int cbase = cta * RA_BN;
    const bool tail = (int64_t)(cta + 1) * RA_BN > n;
    const float cs = csv.x;
    if (SYM && cta / TPB != bid) {
      // Transposed keys: column cbase + lcol is a row too and this wavefront's rows are its
      // candidates; a (row, column) pair is queued when its key, formed with the row's scale,
      // beats the column's threshold csv.y. Keys fall as dots rise (-(s r) for cosine with
      // r >= 0, r - 2 s for squared Euclidean) and f32 rounding is monotone, so for dots
      // s <= smax and row scales r in [rs_min, rs_max] every key is at least
      // -max(smax rs_min, smax rs_max), resp. rs_min - 2 smax: the whole tile is tested
      // against that bound first, then (cosine) each group of four rows against its own.
      // The queue keeps the raw dot; the scatter forms the key again, identically.
      float smax;
      {
        const float m0 = fmaxf(fmaxf(acc[0], acc[1]), acc[2]), m1 = fmaxf(fmaxf(acc[3], acc[4]), acc[5]);
        const float m2 = fmaxf(fmaxf(acc[6], acc[7]), acc[8]), m3 = fmaxf(fmaxf(acc[9], acc[10]), acc[11]);
        const float m4 = fmaxf(fmaxf(acc[12], acc[13]), acc[14]);
        smax = fmaxf(fmaxf(m0, m1), fmaxf(fmaxf(m2, m3), fmaxf(m4, acc[15])));
      }
      const int cabs = cbase + lcol;
      float t0c = csv.y;
      if (!((int64_t)cabs < n)) t0c = -FINF;
      const float rkmin = COS ? -fmaxf(smax * rs_min, smax * rs_max) : (rs_min - 2.0f * smax);
      unsigned m16 = 0u;
      if (__ballot(rkmin < t0c) != 0) {
        m16 = 0xffffu;
        if (COS) {
          m16 = 0;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float gm = fmaxf(fmaxf(acc[4 * j], acc[4 * j + 1]), fmaxf(acc[4 * j + 2], acc[4 * j + 3]));
            m16 |= (-fmaxf(gm * rs_min, gm * rs_max) < t0c ? 15u : 0u) << (4 * j);
          }
        }
        if (row0 + RS_ROWS > n) {  // the last block: rows past n are padding
#pragma unroll
          for (int e = 0; e < 16; ++e)
            if (!((int64_t)(row0 + wave * 32 + (e >> 2) * 8 + lrow + (e & 3)) < n)) m16 &= ~(1u << e);
        }
      }
      if (__ballot(m16 != 0) != 0) {
        const int rbase = row0 + wave * 32 + lrow;
#pragma unroll 1
        for (int j = 0; j < 4; ++j) {
          if (__ballot(((m16 >> (4 * j)) & 15u) != 0) == 0) continue;
          f32x4 rs;  // the scales of rows lrow + 8 j .. + 3
          {
            const unsigned ra_ = lds_addr(rsc + lrow + j * 8);
            asm volatile("ds_read_b128 %0, %1\ns_waitcnt lgkmcnt(0)" : "=v"(rs) : "v"(ra_) : "memory");
          }
          float sj[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) sj[i] = acc[j * 4 + i];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float rk = COS ? -(sj[i] * rs[i]) : (rs[i] - 2.0f * sj[i]);
            const bool rp = ((m16 >> (4 * j + i)) & 1u) && (rk < t0c);
            const unsigned long long rm = __ballot(rp);
            if (rm) {
              if (lq > RS_QL - 64) flush();
              const int below = (int)__builtin_amdgcn_mbcnt_hi(
                  (unsigned)(rm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)rm, 0u));
              const unsigned q = (unsigned)(lq + below);
              if (rp) {
                lds_st_b32(lqa + 4u * q, cabs);
                lds_st_b32(lqa + 4u * (RS_QL + q), rbase + j * 8 + i);
                lds_st_b32(lqa + 4u * (2 * RS_QL + q), __float_as_int(sj[i]));
              }
              lq += __popcll(rm);
            }
          }
        }
      }
    }
    f32x4 tq[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) tq[j] = reinterpret_cast<const f32x4*>(thr + 32 * half + lrow)[2 * j];
    int rem = (int)std::min<int64_t>(n - (int64_t)cbase, RA_BN) - lcol;
    int sb = lrow * RA_SCS + lcol;
    asm volatile("" : "+v"(rem), "+v"(sb));
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const float s_ = acc[e];
      acc[e] = COS ? -(s_ * cs) : (cs - 2.0f * s_);
    }
    if (!qmap) {
      const bool diag = self && (cbase < row0 + RS_ROWS) && (cbase + RA_BN > row0);
      int dv = cbase - row0 - wave * 32 + lcol - lrow;
      asm volatile("" : "+v"(dv));
      if (__builtin_expect(diag || tail, 0)) {
#pragma unroll
        for (int e = 0; e < 16; ++e)
          if (rem <= 0 || dv == (e >> 2) * 8 + (e & 3)) acc[e] = FINF;
      }
    } else if (__builtin_expect(tail || (cbase + RA_BN > qlo && cbase <= qhi), 0)) {
      const int cg = cbase + lcol;
#pragma unroll
      for (int e = 0; e < 16; ++e)
        if (rem <= 0 || cg == qg[e]) acc[e] = FINF;
    }
    {
      unsigned long long anyp = 0;
#pragma unroll
      for (int e = 0; e < 16; ++e) anyp |= __ballot(acc[e] < tq[e >> 2][e & 3]);
      if (anyp == 0) continue;  // nothing in this tile enters a list
    }
    bool touched = false, overflow = false;
    // the ballot of the slot this lane's list is counted in; the slot id is made opaque per
    // tile, so that the compiler keeps no per-slot lane masks across the walk (they would
    // spill scalar registers)
    unsigned mlo = 0u, mhi = 0u;
    int lsl = myslot;
    asm volatile("" : "+v"(lsl));
    int ob = (int)lds_addr(bucket) + ((grp & 1) * 32 + (grp >> 1) * 4) * RB_BKT * 8;
    asm volatile("" : "+v"(ob));
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float key = acc[j * 4 + i];
        const bool pass = key < tq[j][i];
        const unsigned long long m = __ballot(pass);
        if (m) {
          touched = true;
          const int r0 = j * 8 + i;
          const int o0 = r0, o1 = r0 + 32, o2 = r0 + 4, o3 = r0 + 36;
          const int c0 = __builtin_amdgcn_readlane(bcnt, o0);
          const int c1 = __builtin_amdgcn_readlane(bcnt, o1);
          const int c2 = __builtin_amdgcn_readlane(bcnt, o2);
          const int c3 = __builtin_amdgcn_readlane(bcnt, o3);
          const int n0 = __popcll(m & 0xffffull), n1 = __popcll(m & 0xffff0000ull);
          const int n2 = __popcll(m & 0xffff00000000ull), n3 = __popcll(m & 0xffff000000000000ull);
          overflow |= (c0 + n0 > RB_BKT) | (c1 + n1 > RB_BKT) | (c2 + n2 > RB_BKT) |
                      (c3 + n3 > RB_BKT);
          const unsigned long long gm = m & (0xffffull << (16 * grp));
          const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(gm >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((unsigned)gm, 0u));
          const int base = grp == 0 ? c0 : (grp == 1 ? c1 : (grp == 2 ? c2 : c3));
          const int pos = base + below;
          if (pass && pos < RB_BKT)
            lds_st_b64((unsigned)(ob + r0 * RB_BKT * 8 + pos * 8),
                       make_int2(__float_as_int(key), cbase + lcol));
          if (lsl == j * 4 + i) {
            mlo = (unsigned)m;
            mhi = (unsigned)(m >> 32);
          }
        }
      }
    if (touched) {
      asm volatile("" ::: "memory");
      __builtin_amdgcn_wave_barrier();
      bool drain = overflow;
      if (!overflow) {
        const unsigned long long mine = ((unsigned long long)mhi << 32) | mlo;
        bcnt += (int)__popcll((mine >> gsh) & 0xffffull);
        drain = __ballot(bcnt >= RB_DRAIN) != 0;
      }
      if (drain) {
        const unsigned ba = lds_addr(bucket + lane * RB_BKT);
        int2 en = lds_ld_b64(ba);
#pragma unroll 1
        for (int e = 0; e < bcnt; ++e) {
          const int2 cur = en;
          en = lds_ld_b64(ba + (unsigned)(e + 1) * 8u);
          cand_insert<C>(lk, li, __int_as_float(cur.x), cur.y);
        }
        bcnt = 0;
        if (overflow) {
          asm volatile("" ::: "memory");
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int e = 0; e < 16; ++e) sc[sb + ((e >> 2) * 8 + (e & 3)) * RA_SCS] = acc[e];
          asm volatile("" ::: "memory");
          __builtin_amdgcn_wave_barrier();
          const float* srow = sc + (lane & 31) * RA_SCS + 16 * (lane >> 5);
#pragma unroll 1
          for (int q = 0; q < 16; ++q) cand_insert<C>(lk, li, srow[q], cbase + 16 * (lane >> 5) + q);
        }
        asm volatile("" ::: "memory");
        __builtin_amdgcn_wave_barrier();
        lds_st_b32(lds_addr(thr + lane), __float_as_int(lk[C - 1]));
        asm volatile("" ::: "memory");
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  if (__ballot(bcnt > 0) != 0) {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_wave_barrier();
    const unsigned ba = lds_addr(bucket + lane * RB_BKT);
#pragma unroll 1
    for (int e = 0; e < bcnt; ++e) {
      const int2 cur = lds_ld_b64(ba + (unsigned)e * 8u);
      cand_insert<C>(lk, li, __int_as_float(cur.x), cur.y);
    }
  }
  wait_vm<0>();
  const int64_t lr = row0 + wave * 32 + (lane & 31);  // list row
  if (SYM) {
    if (lq > 0) flush();
    if (lane == 0) cols.qcount[(int64_t)qslot * RS_WAVES + wave] = qn;
    // the row's largest forward key becomes its threshold when smaller: a transposed key
    // at or above it is beaten by all 32 forward candidates (workgroups still to come read
    // the new value or, from a stale cache line, the old one: either is an upper bound)
    const float fwdmax = fmaxf(lk[C - 1], __shfl_xor(lk[C - 1], 32, 64));
    if (lane < 32 && lr < n && fwdmax < cols.kt[2 * lr + 1]) cols.kt[2 * lr + 1] = fwdmax;
  }
  if (lr < nq) {
    // a piece of a split walk writes its partial lists to its part's rows of the piece lists
    float* const okey = P > 1 ? cols.plist_key + part * cols.prows * 2 * C : cand_key;
    int32_t* const oidx = P > 1 ? cols.plist_idx + part * cols.prows * 2 * C : cand_idx;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      okey[(lr * 2 + (lane >> 5)) * C + q] = lk[q];
      oidx[(lr * 2 + (lane >> 5)) * C + q] = li[q];
    }
  }
}

// ------------------------------------------------------------------ 2d. symmetric sweep
constexpr int64_t KNN_SYM_MIN_N = 65536;  // automatic use of the symmetric sweep from this n
// The sample of the threshold pass: rows q * n / S, zero padded to a multiple of 32.
__global__ void rb_sample_kernel(const uint16_t* __restrict__ Xh, const float* __restrict__ scale,
                                 int64_t n, int dpad, int64_t S, int64_t Spad,
                                 uint16_t* __restrict__ Xs, float* __restrict__ ss) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int per = dpad / 8;  // 16-byte pieces per row
  if (t >= Spad * per) return;
  const int64_t q = t / per;
  const int c = (int)(t % per);
  const int64_t src = q < S ? q * n / S : -1;
  int4 v = make_int4(0, 0, 0, 0);
  if (src >= 0) v = reinterpret_cast<const int4*>(Xh + src * dpad)[c];
  reinterpret_cast<int4*>(Xs + q * dpad)[c] = v;
  if (c == 0) ss[q] = src >= 0 ? scale[src] : 0.0f;
}
// A row's reverse threshold: the RB_THR_RANK-th smallest key of its two sorted half-lists
// against the sample (any threshold is correct; this one queues about RB_THR_RANK / S of
// the transposed keys, S the sample size).
constexpr int RB_THR_RANK = 6;
static_assert(RB_THR_RANK <= RB_PASS_C, "the pass lists must hold the threshold rank");
template <int L>  // candidates per half-list of the pass
__global__ void rb_thr0_kernel(const float* __restrict__ ckey, const float* __restrict__ colscale,
                               int64_t n, int64_t npad, float* __restrict__ kt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npad) return;
  kt[2 * r] = colscale[r];
  if (r >= n) {
    kt[2 * r + 1] = -FINF;
    return;
  }
  const float* a = ckey + (r * 2) * L;
  const float* b = ckey + (r * 2 + 1) * L;
  int ia = 0, ib = 0;
  float v = FINF;
  for (int t = 0; t < RB_THR_RANK; ++t) {
    const float x = ia < L ? a[ia] : FINF, y = ib < L ? b[ib] : FINF;
    if (x <= y) {
      v = x;
      ++ia;
    } else {
      v = y;
      ++ib;
    }
  }
  kt[2 * r + 1] = v;
}
// Largest forward key of every row (+inf when its lists are not full): a reverse
// candidate above it is beaten by all 32 forward candidates and cannot be kept.
__global__ void rb_fwdmax_kernel(const float* __restrict__ ckey, const int32_t* __restrict__ cidx,
                                 int64_t n, float* __restrict__ fmax, int32_t* __restrict__ dropped) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float m = -FINF;
  for (int e = 0; e < 32; ++e) m = cidx[r * 32 + e] < 0 ? FINF : fmaxf(m, ckey[r * 32 + e]);
  fmax[r] = m;
  dropped[r] = 0;
}
// queued reverse candidates into per-row buckets (order within a bucket is free: the
// merge ranks by (key, index)); the key of row j's candidate r is formed here from the
// dot and r's scale, as the tile kernel forms forward keys
template <bool COS>
__global__ __launch_bounds__(256) void rb_scatter_kernel(const int32_t* __restrict__ qj,
                                                         const int32_t* __restrict__ qr,
                                                         const float* __restrict__ qs,
                                                         const int32_t* __restrict__ qcount,
                                                         int64_t qcap, int rcap,
                                                         const float* __restrict__ colscale,
                                                         const float* __restrict__ fmax,
                                                         int32_t* __restrict__ dropped,
                                                         int32_t* __restrict__ rcnt,
                                                         float* __restrict__ rkey,
                                                         int32_t* __restrict__ ridx) {
  const int64_t w = blockIdx.x;
  const int64_t cnt = min<int64_t>(qcount[w], qcap);
  for (int64_t e = threadIdx.x; e < cnt; e += 256) {
    const int64_t q = w * qcap + e;
    const int32_t j = qj[q];
    const int32_t r = qr[q];
    const float s_ = qs[q];
    const float cs = colscale[r];
    const float key = COS ? -(s_ * cs) : (cs - 2.0f * s_);
    if (key > fmax[j]) {  // ranked out by the row's forward candidates
      dropped[j] = 1;
      continue;
    }
    const int slot = atomicAdd(&rcnt[j], 1);
    if (slot < rcap) {
      rkey[(int64_t)j * rcap + slot] = key;
      ridx[(int64_t)j * rcap + slot] = r;
    }
  }
}
// Per row: the 32 forward candidates (two half-lists) and its reverse bucket ranked by
// (key, index); the 32 best are kept, sorted. Every column not kept has a key >= the
// returned bound: forward non-candidates >= their half's last key, reverse keys never
// queued >= thr0 (the smallest value the threshold took during the sweep), entries ranked out (here, or dropped at the scatter above the largest
// forward key) >= the 32nd kept. A full bucket voids the bound. Rows below fb_rows (a
// split walk's, their pieces merged beforehand) take the forward bound from fbound.
constexpr int RB_RCAP = 768;
__global__ __launch_bounds__(256) void rb_merge_kernel(
    const float* __restrict__ ckey, const int32_t* __restrict__ cidx, const int32_t* __restrict__ rcnt,
    const float* __restrict__ rkey, const int32_t* __restrict__ ridx, const float* __restrict__ kt,
    const int32_t* __restrict__ dropped, int64_t n, float* __restrict__ mkey,
    int32_t* __restrict__ midx, float* __restrict__ bound, const float* __restrict__ fbound,
    int64_t fb_rows) {
  __shared__ float sk[4][32 + RB_RCAP];
  __shared__ int32_t si[4][32 + RB_RCAP];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + wave;
  if (row >= n) return;
  const int cnt = rcnt[row];
  const int m = min(cnt, RB_RCAP);
  const int K = 32 + m;
  for (int e = lane; e < K; e += 64) {
    float key;
    int32_t idx;
    if (e < 32) {
      key = ckey[row * 32 + e];
      idx = cidx[row * 32 + e];
    } else {
      key = rkey[row * RB_RCAP + (e - 32)];
      idx = ridx[row * RB_RCAP + (e - 32)];
    }
    if (idx < 0) key = FINF;
    sk[wave][e] = key;
    si[wave][e] = idx;
  }
  __builtin_amdgcn_wave_barrier();
  __threadfence_block();
  int nvalid = 0;
  for (int e = lane; e < K; e += 64) nvalid += si[wave][e] >= 0 ? 1 : 0;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) nvalid += __shfl_xor(nvalid, o, 64);
  float key31 = FINF;  // the 32nd kept key
  for (int e = lane; e < K; e += 64) {
    const float key = sk[wave][e];
    const int32_t idx = si[wave][e];
    if (idx < 0) continue;
    int rank = 0;
    for (int f = 0; f < K; ++f) {
      const float k2 = sk[wave][f];
      const int32_t i2 = si[wave][f];
      rank += (i2 >= 0 && (k2 < key || (k2 == key && i2 < idx))) ? 1 : 0;
    }
    if (rank < 32) {
      mkey[row * 32 + rank] = key;
      midx[row * 32 + rank] = idx;
    }
    if (rank == 31) key31 = key;
  }
  for (int q = nvalid + lane; q < 32; q += 64) {
    mkey[row * 32 + q] = FINF;
    midx[row * 32 + q] = -1;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) key31 = fminf(key31, __shfl_xor(key31, o, 64));
  if (lane == 0) {
    const float fb = row < fb_rows ? fbound[row]
                                   : fminf(ckey[(row * 2) * 16 + 15], ckey[(row * 2 + 1) * 16 + 15]);
    float b = fminf(fb, kt[2 * row + 1]);
    if (nvalid > 32 || dropped[row]) b = fminf(b, key31);  // some candidate was ranked out
    if (cnt > RB_RCAP) b = -FINF;
    bound[row] = b;
  }
}

// Column shards (the certificate-rate sample's) or pieces of split walks (the symmetric
// sweep's last slice) merged per row, one wavefront a row: the 32 best of the G lists
// (two sorted half-lists of 16 each, list t's at (t * stride + row) * 32) ranked by
// (key, index), and a bound at or below every left-out
// column's key: each half-list bounds the columns of its half and shard by its 16th key
// (+inf when not full), and a candidate ranked out lies at or above the 32nd kept.
constexpr int KNN_SAMPLE_SHARDS_MAX = 8;
__global__ __launch_bounds__(256) void knn_shard_merge_kernel(
    const float* __restrict__ ckey, const int32_t* __restrict__ cidx, int G, int64_t nrows,
    int64_t stride, float* __restrict__ mkey, int32_t* __restrict__ midx, float* __restrict__ bound) {
  __shared__ float sk[4][32 * KNN_SAMPLE_SHARDS_MAX];
  __shared__ int32_t si[4][32 * KNN_SAMPLE_SHARDS_MAX];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + wave;
  if (row >= nrows) return;
  const int K = 32 * G;
  float b = FINF;
  for (int e = lane; e < K; e += 64) {
    const int64_t at = ((int64_t)(e >> 5) * stride + row) * 32 + (e & 31);
    const int32_t idx = cidx[at];
    const float key = idx >= 0 ? ckey[at] : FINF;
    sk[wave][e] = key;
    si[wave][e] = idx;
    if ((e & 15) == 15) b = fminf(b, key);  // a half-list's 16th key
  }
  __builtin_amdgcn_wave_barrier();
  __threadfence_block();
  int nvalid = 0;
  for (int e = lane; e < K; e += 64) nvalid += si[wave][e] >= 0 ? 1 : 0;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    nvalid += __shfl_xor(nvalid, o, 64);
    b = fminf(b, __shfl_xor(b, o, 64));
  }
  float key31 = FINF;
  for (int e = lane; e < K; e += 64) {
    const float key = sk[wave][e];
    const int32_t idx = si[wave][e];
    if (idx < 0) continue;
    int rank = 0;
    for (int f = 0; f < K; ++f) {
      const float k2 = sk[wave][f];
      const int32_t i2 = si[wave][f];
      rank += (i2 >= 0 && (k2 < key || (k2 == key && i2 < idx))) ? 1 : 0;
    }
    if (rank < 32) {
      mkey[row * 32 + rank] = key;
      midx[row * 32 + rank] = idx;
    }
    if (rank == 31) key31 = key;
  }
  for (int q = nvalid + lane; q < 32; q += 64) {
    mkey[row * 32 + q] = FINF;
    midx[row * 32 + q] = -1;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) key31 = fminf(key31, __shfl_xor(key31, o, 64));
  if (lane == 0) bound[row] = nvalid > 32 ? fminf(b, key31) : b;
}

// ------------------------------------------------------------------ 3. re-rank
// The re-rank and exact kernels hold a query row in registers, T floats per lane: T = 16
// covers d <= 1024, T = 64 covers d <= 4096 (KNN_MAXD). Lane l sums dimensions l, l + 64,
// ... in order whatever T is, so both forms give the same canonical distance.
constexpr int KNN_MAXD = 4096;

template <int MAXT>
__global__ __launch_bounds__(256) void knn_rerank_kernel(
    const float* __restrict__ X, int64_t n, int d, int64_t ld, const float* __restrict__ nrm,
    const float* __restrict__ sqn, const float* __restrict__ colscale,
    const unsigned* __restrict__ flags, const int32_t* __restrict__ cand_idx,
    const float* __restrict__ cand_key, int C, int halves, const float* __restrict__ row_bound,
    int k, int metric, float gamma, int32_t* __restrict__ out_idx, float* __restrict__ out_dist,
    int* __restrict__ fail_count, int32_t* __restrict__ fail_rows,
    const int32_t* __restrict__ rowmap, int64_t nrows) {
  // the lists of list row `lr` belong to data row `row` (rowmap, or the same row)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t lr = (int64_t)blockIdx.x * 4 + wave;
  if (lr >= (rowmap ? nrows : n)) return;
  const int64_t row = rowmap ? rowmap[lr] : lr;
  const int nt = (d + 63) / 64;
  float q[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t) {
    const int c = t * 64 + lane;
    q[t] = (t < nt && c < d) ? X[row * ld + c] : 0.0f;
  }
  // Candidate c sits in lane c. Its approximate key bounds its exact distance from below
  // (the same bound the certificate applies to non-candidates), so only candidates
  // whose bound does not clear the exact k-th distance of the k smallest-key
  // candidates need an exact distance: the rest are strictly farther than the k-th.
  const float u = 5.9604645e-08f;  // 2^-24
  const float dt = (float)(d / 64 + 12);
  float E = 0.0f, eps = 0.0f;
  if (metric == 1) {
    eps = 1.5f * gamma + 64.0f * u;
  } else {
    const float nmax = __uint_as_float(flags[1]);
    E = 1.5f * (dt * u * (sqn[row] + nmax * nmax) + 2.0f * (gamma + 2.0f * u) * nrm[row] * nmax);
  }
  const int cj = lane < C ? cand_idx[lr * C + lane] : -1;
  const float ckey = cj >= 0 ? cand_key[lr * C + lane] : FINF;
  const bool valid = cj >= 0;
  const int nvalid = __popcll(__ballot(valid));
  int krank = 0;
  for (int c = 0; c < C; ++c) {
    const float ok = __shfl(ckey, c, 64);
    const int oj = __shfl(cj, c, 64);
    if (oj >= 0 && (ok < ckey || (ok == ckey && c < lane))) ++krank;
  }
  float my_d = FINF, my_sq = FINF;
  int my_j = -1;
  // Two candidates are evaluated together so that their loads are in flight at once;
  // each one's sum is formed exactly as alone. (Reading the rows from the bf16 copy of
  // bf16-exact inputs instead halves the bytes but measured slower: 15.4 against 8.1 ms
  // at 1M rows.)
  auto load = [&](int j, int cc) -> float { return X[(int64_t)j * ld + cc]; };
  auto term = [&](float acc, float qt, float v) {
    if (metric == 1) {
      const float p = qt * v;
      return acc + p;
    }
    const float df = qt - v;
    const float p = df * df;
    return acc + p;
  };
  auto finish = [&](float acc, int c, int j) {
    acc = tree64(acc);
    acc = __shfl(acc, 0, 64);
    const float dist = metric == 1 ? cos_dist(acc, nrm[row], nrm[j]) : sqrtf(acc);
    if (lane == c) {
      my_d = dist;
      my_sq = acc;
      my_j = j;
    }
  };
  auto evaluate = [&](unsigned long long mask) {
    while (mask) {
      const int c1 = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      const int j1 = __shfl(cj, c1, 64);
      if (mask) {
        const int c2 = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const int j2 = __shfl(cj, c2, 64);
        float v1[MAXT], v2[MAXT];
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
          const int cc = t * 64 + lane;
          const bool in = t < nt && cc < d;
          v1[t] = in ? load(j1, cc) : 0.0f;
          v2[t] = in ? load(j2, cc) : 0.0f;
        }
        float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
          const int cc = t * 64 + lane;
          if (t < nt && cc < d) {
            a1 = term(a1, q[t], v1[t]);
            a2 = term(a2, q[t], v2[t]);
          }
        }
        finish(a1, c1, j1);
        finish(a2, c2, j2);
      } else {
        float a1 = 0.0f;
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
          const int cc = t * 64 + lane;
          if (t < nt && cc < d) a1 = term(a1, q[t], load(j1, cc));
        }
        finish(a1, c1, j1);
      }
    }
  };
  auto exact_rank = [&]() {
    int r = 0;
    for (int c = 0; c < C; ++c) {
      const float od = __shfl(my_d, c, 64);
      const int oj = __shfl(my_j, c, 64);
      if (oj >= 0 && (od < my_d || (od == my_d && oj < my_j))) ++r;
    }
    return r;
  };
  evaluate(__ballot(valid && krank < k));
  if (nvalid > k) {
    const int r0 = exact_rank();
    const unsigned long long kb0 = __ballot(my_j >= 0 && r0 == k - 1);
    const int src0 = __ffsll((long long)kb0) - 1;
    const float dk0 = __shfl(my_d, src0, 64);
    const float dk20 = __shfl(my_sq, src0, 64);
    bool need;
    if (metric == 1) need = !((1.0f + ckey * colscale[row]) - eps > dk0);
    else need = !(sqn[row] + ckey - E > dk20 * (1.0f + (dt + 4.0f) * u) + 1e-30f);
    evaluate(__ballot(valid && krank >= k && need));
  }
  const int rank = exact_rank();
  const bool kept = my_j >= 0;
  if (kept && rank < k) {
    out_idx[row * k + rank] = my_j;
    out_dist[row * k + rank] = my_d;
  }
  for (int s = lane; s < k; s += 64)
    if (s >= nvalid) {
      out_idx[row * k + s] = -1;
      out_dist[row * k + s] = FINF;
    }
  // certificate: every non-candidate's approximate key is >= the C-th kept key
  bool pass;
  if ((int64_t)nvalid >= n - 1) {
    pass = true;
  } else if (nvalid < k) {
    pass = false;
  } else {
    const unsigned long long kb = __ballot(kept && rank == k - 1);
    const int src = __ffsll((long long)kb) - 1;
    const float dk = __shfl(my_d, src, 64);
    const float dk2 = __shfl(my_sq, src, 64);
    // one sorted list of C keys, or two column halves that each kept their C/2 best, or
    // (symmetric sweep) the bound its merge computed
    const float kc = row_bound ? row_bound[lr]
                     : halves == 2 ? fminf(cand_key[lr * C + (C / 2 - 1)], cand_key[lr * C + (C - 1)])
                                   : cand_key[lr * C + (C - 1)];
    if (metric == 1) {
      const float lb = 1.0f + kc * colscale[row];
      pass = (lb - eps) > dk;
    } else {
      const float lb2 = sqn[row] + kc - E;
      pass = lb2 > dk2 * (1.0f + (dt + 4.0f) * u) + 1e-30f;
    }
  }
  if (!pass && lane == 0) {
    const int pos = atomicAdd(fail_count, 1);
    fail_rows[pos] = (int32_t)row;
  }
}

// ------------------------------------------------------------------ 3b. escalated rows
// Re-rank of a row searched again over S column shards, each of which kept two column
// halves of 32 candidates (list row lr of shard t at cand[(t * nrows + lr) * 64]). Every
// column a shard left out has a key at or above that shard's smaller last key, so the
// bound of the row is the smallest of those over the shards, which sits near rank 64 S of
// the row rather than 64. Every valid candidate (S per lane) gets its exact distance, the
// best k by (distance, index) are written, and the certificate is the one of
// knn_rerank_kernel against that bound; a failing row is listed for the exact scan.
template <int S, int MAXT>
__global__ __launch_bounds__(256) void knn_rerank_all_kernel(
    const float* __restrict__ X, int64_t n, int d, int64_t ld, const float* __restrict__ nrm,
    const float* __restrict__ sqn, const float* __restrict__ colscale,
    const unsigned* __restrict__ flags, const int32_t* __restrict__ cand_idx,
    const float* __restrict__ cand_key, int k, int metric, float gamma,
    int32_t* __restrict__ out_idx, float* __restrict__ out_dist, int* __restrict__ fail_count,
    int32_t* __restrict__ fail_rows, const int32_t* __restrict__ rowmap, int64_t nrows) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t lr = (int64_t)blockIdx.x * 4 + wave;
  if (lr >= nrows) return;
  const int64_t row = rowmap[lr];
  const int nt = (d + 63) / 64;
  float q[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t) {
    const int c = t * 64 + lane;
    q[t] = (t < nt && c < d) ? X[row * ld + c] : 0.0f;
  }
  const float u = 5.9604645e-08f;  // 2^-24
  const float dt = (float)(d / 64 + 12);
  float E = 0.0f, eps = 0.0f;
  if (metric == 1) {
    eps = 1.5f * gamma + 64.0f * u;
  } else {
    const float nmax = __uint_as_float(flags[1]);
    E = 1.5f * (dt * u * (sqn[row] + nmax * nmax) + 2.0f * (gamma + 2.0f * u) * nrm[row] * nmax);
  }
  int cj[S];
  float bound = FINF;
  int nvalid = 0;
#pragma unroll
  for (int t = 0; t < S; ++t) {
    const int64_t base = ((int64_t)t * nrows + lr) * 64;
    cj[t] = cand_idx[base + lane];
    const float ck = cand_key[base + lane];
    bound = fminf(bound, fminf(__shfl(ck, 31, 64), __shfl(ck, 63, 64)));
    nvalid += __popcll(__ballot(cj[t] >= 0));
  }
  // exact canonical distances of every valid candidate, two rows' loads in flight at once
  float my_d[S], my_sq[S];
#pragma unroll
  for (int t = 0; t < S; ++t) {
    my_d[t] = FINF;
    my_sq[t] = FINF;
  }
  auto finish = [&](float acc, int c, int j, int t) {
    acc = tree64(acc);
    acc = __shfl(acc, 0, 64);
    const float dist = metric == 1 ? cos_dist(acc, nrm[row], nrm[j]) : sqrtf(acc);
    if (lane == c) {
      my_d[t] = dist;
      my_sq[t] = acc;
    }
  };
#pragma unroll
  for (int t = 0; t < S; ++t) {
    unsigned long long mask = __ballot(cj[t] >= 0);
    while (mask) {
      const int c1 = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      const int j1 = __shfl(cj[t], c1, 64);
      const int c2 = mask ? __ffsll((long long)mask) - 1 : -1;
      if (mask) mask &= mask - 1;
      const int j2 = c2 >= 0 ? __shfl(cj[t], c2, 64) : -1;
      float v1[MAXT], v2[MAXT];
#pragma unroll
      for (int tt = 0; tt < MAXT; ++tt) {
        const int cc = tt * 64 + lane;
        const bool in = tt < nt && cc < d;
        v1[tt] = in ? X[(int64_t)j1 * ld + cc] : 0.0f;
        v2[tt] = (in && j2 >= 0) ? X[(int64_t)j2 * ld + cc] : 0.0f;
      }
      float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
      for (int tt = 0; tt < MAXT; ++tt) {
        const int cc = tt * 64 + lane;
        if (tt < nt && cc < d) {
          if (metric == 1) {
            const float p1 = q[tt] * v1[tt], p2 = q[tt] * v2[tt];
            a1 = a1 + p1;
            a2 = a2 + p2;
          } else {
            const float f1 = q[tt] - v1[tt], f2 = q[tt] - v2[tt];
            const float p1 = f1 * f1, p2 = f2 * f2;
            a1 = a1 + p1;
            a2 = a2 + p2;
          }
        }
      }
      finish(a1, c1, j1, t);
      if (c2 >= 0) finish(a2, c2, j2, t);
    }
  }
  // rank of each candidate among all of them by (distance, index)
  int rank[S];
#pragma unroll
  for (int t = 0; t < S; ++t) rank[t] = 0;
#pragma unroll
  for (int t2 = 0; t2 < S; ++t2)
    for (int c = 0; c < 64; ++c) {
      const float od = __shfl(my_d[t2], c, 64);
      const int oj = __shfl(cj[t2], c, 64);
      if (oj < 0) continue;
#pragma unroll
      for (int t = 0; t < S; ++t)
        rank[t] += (od < my_d[t] || (od == my_d[t] && oj < cj[t])) ? 1 : 0;
    }
  float dk = FINF, dk2 = FINF;
#pragma unroll
  for (int t = 0; t < S; ++t) {
    if (cj[t] >= 0 && rank[t] < k) {
      out_idx[row * k + rank[t]] = cj[t];
      out_dist[row * k + rank[t]] = my_d[t];
    }
    const bool kth = cj[t] >= 0 && rank[t] == k - 1;
    const unsigned long long kb = __ballot(kth);
    if (kb) {
      const int src = __ffsll((long long)kb) - 1;
      dk = __shfl(my_d[t], src, 64);
      dk2 = __shfl(my_sq[t], src, 64);
    }
  }
  for (int s = lane; s < k; s += 64)
    if (s >= nvalid) {
      out_idx[row * k + s] = -1;
      out_dist[row * k + s] = FINF;
    }
  bool pass;
  if ((int64_t)nvalid >= n - 1) {
    pass = true;
  } else if (nvalid < k) {
    pass = false;
  } else if (metric == 1) {
    pass = (1.0f + bound * colscale[row]) - eps > dk;
  } else {
    pass = sqn[row] + bound - E > dk2 * (1.0f + (dt + 4.0f) * u) + 1e-30f;
  }
  if (!pass && lane == 0) {
    const int pos = atomicAdd(fail_count, 1);
    fail_rows[pos] = (int32_t)row;
  }
}

// ------------------------------------------------------------------ 4. exact rows
// list capacities of the exact scan: 32 for k <= 32, 64 up to KMAX
constexpr int KMAX = 64;

// block-wide merge of 256 per-thread sorted lists (k entries each, in LDS) into list 0
template <int KMAX>
__device__ void exact_block_merge(float* sk, int* si, int k) {
  const int tid = threadIdx.x;
  for (int stride = 1; stride < 256; stride <<= 1) {
    if ((tid & (2 * stride - 1)) == 0) {
      float* ak = sk + tid * KMAX;
      int* ai = si + tid * KMAX;
      const float* bk = sk + (tid + stride) * KMAX;
      const int* bi = si + (tid + stride) * KMAX;
      float mk[KMAX];
      int mi[KMAX];
      int pa = 0, pb = 0;
      for (int q = 0; q < k; ++q) {
        const bool ta = (ak[pa] < bk[pb]) || (ak[pa] == bk[pb] && (unsigned)ai[pa] < (unsigned)bi[pb]);
        mk[q] = ta ? ak[pa] : bk[pb];
        mi[q] = ta ? ai[pa] : bi[pb];
        pa += ta;
        pb += !ta;
      }
      for (int q = 0; q < k; ++q) {
        ak[q] = mk[q];
        ai[q] = mi[q];
      }
    }
    __syncthreads();
  }
}

// Exact rows, phase 1: block (f, part) scans columns [part*span, (part+1)*span) of
// failing row f in the canonical order and keeps the block's best k by (dist, idx).
// Exact scan for rows whose certificate failed: up to EX_ROWS failing rows per launch,
// held in registers, against the columns of one part per workgroup. One wavefront per
// column at a time: lanes hold the column's dimensions (the canonical strided partial
// sums and halving tree of the rerank), and lane f keeps the sorted list of failing
// row f. Each wavefront writes its lists; the merge kernel combines them.
template <int MAXT, int EX_ROWS, int KMAX>
__global__ __launch_bounds__(256) void knn_exact_kernel(
    const float* __restrict__ X, int64_t n, int d, int64_t ld, const float* __restrict__ nrm,
    int k, int metric, const int32_t* __restrict__ fail_rows, int nrows, int parts, int64_t span,
    float* __restrict__ part_key, int32_t* __restrict__ part_idx) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int part = blockIdx.x;
  const int nt = (d + 63) / 64;
  float q[EX_ROWS][MAXT];
  int64_t rows[EX_ROWS];
#pragma unroll
  for (int f = 0; f < EX_ROWS; ++f) {
    rows[f] = f < nrows ? fail_rows[f] : -1;
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      const int c = t * 64 + lane;
      q[f][t] = (f < nrows && t < nt && c < d) ? X[rows[f] * ld + c] : 0.0f;
    }
  }
  float lk[KMAX];
  int li[KMAX];
#pragma unroll
  for (int r = 0; r < KMAX; ++r) {
    lk[r] = FINF;
    li[r] = -1;
  }
  const float my_nrm = lane < nrows ? nrm[fail_rows[lane]] : 0.0f;
  const int64_t my_row = lane < nrows ? fail_rows[lane] : -1;
  const int64_t j0 = (int64_t)part * span, j1 = min<int64_t>(n, j0 + span);
  for (int64_t j = j0 + wave; j < j1; j += 4) {
    float v[MAXT];
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      const int c = t * 64 + lane;
      v[t] = (t < nt && c < d) ? X[j * ld + c] : 0.0f;
    }
    const float nj = nrm[j];
#pragma unroll
    for (int f = 0; f < EX_ROWS; ++f) {
      if (f >= nrows) break;
      float acc = 0.0f;
#pragma unroll
      for (int t = 0; t < MAXT; ++t) {
        const int c = t * 64 + lane;
        if (t < nt && c < d) {
          if (metric == 1) {
            const float p = q[f][t] * v[t];
            acc = acc + p;
          } else {
            const float df = q[f][t] - v[t];
            const float p = df * df;
            acc = acc + p;
          }
        }
      }
      acc = tree64(acc);
      acc = __shfl(acc, 0, 64);
      if (lane == f && j != my_row) {
        const float dist = metric == 1 ? cos_dist(acc, my_nrm, nj) : sqrtf(acc);
        if (dist < lk[k - 1] || (dist == lk[k - 1] && (int)j < li[k - 1]))
          list_insert<KMAX>(lk, li, dist, (int)j);
      }
    }
  }
  if (lane < nrows) {
    const int64_t base = (((int64_t)lane * parts + part) * 4 + wave) * k;
    for (int r = 0; r < k; ++r) {
      part_key[base + r] = lk[r];
      part_idx[base + r] = li[r];
    }
  }
}
template <int KMAX>
__global__ __launch_bounds__(256) void knn_exact_merge_kernel(
    int k, const int32_t* __restrict__ fail_rows, int parts, const float* __restrict__ part_key,
    const int32_t* __restrict__ part_idx, int32_t* __restrict__ out_idx,
    float* __restrict__ out_dist) {
  extern __shared__ __attribute__((aligned(16))) char xsm[];
  float* sk = reinterpret_cast<float*>(xsm);
  int* si = reinterpret_cast<int*>(xsm + 256 * KMAX * sizeof(float));
  const int tid = threadIdx.x;
  const int f = blockIdx.x;
  const int64_t row = fail_rows[f];
  float lk[KMAX];
  int li[KMAX];
#pragma unroll
  for (int q = 0; q < KMAX; ++q) {
    lk[q] = FINF;
    li[q] = -1;
  }
  const int64_t m = (int64_t)parts * k;
  for (int64_t e = tid; e < m; e += 256) {
    const float dist = part_key[(int64_t)f * m + e];
    const int j = part_idx[(int64_t)f * m + e];
    if (j < 0) continue;
    if (dist < lk[k - 1] || (dist == lk[k - 1] && j < li[k - 1])) list_insert<KMAX>(lk, li, dist, j);
  }
#pragma unroll
  for (int q = 0; q < KMAX; ++q) {
    sk[tid * KMAX + q] = lk[q];
    si[tid * KMAX + q] = li[q];
  }
  __syncthreads();
  exact_block_merge<KMAX>(sk, si, k);
  for (int q = tid; q < k; q += 256) {
    out_idx[row * k + q] = si[q];
    out_dist[row * k + q] = sk[q];
  }
}

}  // namespace

// ------------------------------------------------------------------ launcher
// compute units of the current device: the symmetric sweep runs one workgroup per CU
static int64_t cu_count() {
  int dev = 0, cus = 0;
  EVOC_HIP(hipGetDevice(&dev));
  EVOC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return cus > 0 ? cus : 1;
}

// Rows whose certificate failed, gathered (16-byte pieces; rows past nq zero) for the
// re-search with longer lists.
__global__ void knn_gather_rows_kernel(const uint16_t* __restrict__ Xh,
                                       const uint16_t* __restrict__ Xl,
                                       const int32_t* __restrict__ qmap, int64_t nq,
                                       int64_t nqpad, int dpad, uint16_t* __restrict__ Qh,
                                       uint16_t* __restrict__ Ql) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int per = dpad / 8;
  if (t >= nqpad * per) return;
  const int64_t r = t / per;
  const int c = (int)(t - r * per);
  int4 h = make_int4(0, 0, 0, 0), l = make_int4(0, 0, 0, 0);
  if (r < nq) {
    const int64_t src = qmap[r];
    h = reinterpret_cast<const int4*>(Xh + src * dpad)[c];
    if (Xl) l = reinterpret_cast<const int4*>(Xl + src * dpad)[c];
  }
  reinterpret_cast<int4*>(Qh + r * dpad)[c] = h;
  if (Ql) reinterpret_cast<int4*>(Ql + r * dpad)[c] = l;
}
// at least this many failed certificates: re-search them with 64 candidates per row
// (fewer go straight to the exact scan, which is cheaper for a handful of rows)
constexpr int KNN_ESCALATE_MIN = 1024;
constexpr int KNN_ESC_SHARDS = 4;  // column shards of the escalation search
// rows sampled to estimate the first pass's certificate rate (streaming kernel, n >= 2^18)
constexpr int64_t KNN_SAMPLE_ROWS = 4096;
constexpr int64_t KNN_SAMPLE_MIN_N = 262144;

__global__ void knn_iota_kernel(int64_t n, int32_t* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int32_t)i;
}

// The resident one-wavefront-per-SIMD kernel's forms: split rows of 256 or 512 dimensions,
// bf16 rows of 256 to 1024 (16 candidates per column half only above 512: below, the
// two-wavefront kernel takes the first pass), cosine or squared-Euclidean keys.
#define EVOC_RS_FORMS(X)                                                                      \
  X(1, true, 16) X(2, true, 16) X(1, true, 32) X(2, true, 32) X(3, false, 16)              \
      X(4, false, 16) X(1, false, 32) X(2, false, 32) X(3, false, 32) X(4, false, 32)
// the symmetric sweep's forms: the first pass of split rows and of bf16 rows above 512
#define EVOC_RS_SYM_FORMS(X) X(1, true) X(2, true) X(3, false) X(4, false)
static void rs_set_attributes() {
#define EVOC_RS_ATTR(NKC_, SPL_, C_)                                                          \
  EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_rs_kernel<NKC_, true, SPL_, C_>,         \
                               hipFuncAttributeMaxDynamicSharedMemorySize,                    \
                               (int)RsRing<SPL_>::SMEM));                                     \
  EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_rs_kernel<NKC_, false, SPL_, C_>,        \
                               hipFuncAttributeMaxDynamicSharedMemorySize,                    \
                               (int)RsRing<SPL_>::SMEM));
  EVOC_RS_FORMS(EVOC_RS_ATTR)
#undef EVOC_RS_ATTR
#define EVOC_RS_SYM_ATTR(NKC_, SPL_)                                                          \
  EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_rs_kernel<NKC_, true, SPL_, 16, true>,   \
                               hipFuncAttributeMaxDynamicSharedMemorySize,                    \
                               (int)(RsRing<SPL_>::SMEM + RS_QBYTES)));                       \
  EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_rs_kernel<NKC_, false, SPL_, 16, true>,  \
                               hipFuncAttributeMaxDynamicSharedMemorySize,                    \
                               (int)(RsRing<SPL_>::SMEM + RS_QBYTES)));
  EVOC_RS_SYM_FORMS(EVOC_RS_SYM_ATTR)
#undef EVOC_RS_SYM_ATTR
}
static bool rs_supported(bool spl, int C, int dpad) {
  if (dpad % RS_KC != 0) return false;
  const int nkc = dpad / RS_KC;
#define EVOC_RS_HAS(NKC_, SPL_, C_) if (nkc == NKC_ && spl == SPL_ && C == C_) return true;
  EVOC_RS_FORMS(EVOC_RS_HAS)
#undef EVOC_RS_HAS
  return false;
}
static bool rs_sym_supported(bool spl, int dpad) {
  if (dpad % RS_KC != 0) return false;
  const int nkc = dpad / RS_KC;
#define EVOC_RS_SYM_HAS(NKC_, SPL_) if (nkc == NKC_ && spl == SPL_) return true;
  EVOC_RS_SYM_FORMS(EVOC_RS_SYM_HAS)
#undef EVOC_RS_SYM_HAS
  return false;
}
// the query rows Qh/Ql (nqpad of them, a multiple of RS_ROWS) against column tiles
// [tc0, tc1) of 32 of the column rows Xh/Xl (n valid); self: the columns are the query
// rows (qmap null), the diagonal is masked
static void rs_launch(bool spl, int C, int dpad, bool cos, const uint16_t* Xh, const uint16_t* Xl,
                      const uint16_t* Qh, const uint16_t* Ql, const int32_t* qmap, int64_t nq,
                      int64_t nqpad, int64_t n, const float* colscale, int tc0, int tc1, int self,
                      int32_t* cidx, float* ckey, hipStream_t s, int shards = 1) {
  const int nkc = dpad / RS_KC;
  const dim3 grid((unsigned)(nqpad / RS_ROWS), (unsigned)shards);
  const RbCols none{};
#define EVOC_RS_GO(NKC_, SPL_, C_)                                                            \
  if (nkc == NKC_ && spl == SPL_ && C == C_) {                                               \
    if (cos)                                                                                 \
      knn_topc_rs_kernel<NKC_, true, SPL_, C_><<<grid, 64 * RS_WAVES, RsRing<SPL_>::SMEM, s>>>( \
          Xh, Xl, Qh, Ql, qmap, nq, n, dpad, colscale, tc0, tc1, self, cidx, ckey, none);   \
    else                                                                                     \
      knn_topc_rs_kernel<NKC_, false, SPL_, C_><<<grid, 64 * RS_WAVES, RsRing<SPL_>::SMEM, s>>>( \
          Xh, Xl, Qh, Ql, qmap, nq, n, dpad, colscale, tc0, tc1, self, cidx, ckey, none);   \
    return;                                                                                  \
  }
  EVOC_RS_FORMS(EVOC_RS_GO)
#undef EVOC_RS_GO
  EVOC_CHECK(false, "knn: no resident-query kernel for this shape");
}
// one slice of the symmetric sweep: workgroups cols.block0 .. + nwg of the npad / RS_ROWS
// blocks (cols.nblocks), each over its half of the column tiles
static void rs_sym_launch(bool spl, int dpad, bool cos, const uint16_t* Xh, const uint16_t* Xl,
                          int64_t n, int64_t npad, const float* colscale, unsigned nwg,
                          const RbCols& cols, int32_t* cidx, float* ckey, hipStream_t s) {
  const int nkc = dpad / RS_KC;
  const int ntile = (int)(npad / RA_BN);
#define EVOC_RS_SYM_GO(NKC_, SPL_)                                                            \
  if (nkc == NKC_ && spl == SPL_) {                                                          \
    if (cos)                                                                                 \
      knn_topc_rs_kernel<NKC_, true, SPL_, 16, true>                                         \
          <<<nwg, 64 * RS_WAVES, RsRing<SPL_>::SMEM + RS_QBYTES, s>>>(                       \
              Xh, Xl, Xh, Xl, nullptr, n, n, dpad, colscale, 0, ntile, 1, cidx, ckey, cols); \
    else                                                                                     \
      knn_topc_rs_kernel<NKC_, false, SPL_, 16, true>                                        \
          <<<nwg, 64 * RS_WAVES, RsRing<SPL_>::SMEM + RS_QBYTES, s>>>(                       \
              Xh, Xl, Xh, Xl, nullptr, n, n, dpad, colscale, 0, ntile, 1, cidx, ckey, cols); \
    return;                                                                                  \
  }
  EVOC_RS_SYM_FORMS(EVOC_RS_SYM_GO)
#undef EVOC_RS_SYM_GO
  EVOC_CHECK(false, "knn: no symmetric resident-query kernel for this shape");
}

void knn_device(const float* dX, int64_t n, int d, int64_t ld, int metric, int k, int precision,
                int32_t* d_idx, float* d_dist, KnnStats* stats, hipStream_t s, DevPool& pool,
                int verbose, KnnCapture* capture) {
  EVOC_CHECK(n >= 1 && d >= 1, "knn: empty input");
  EVOC_CHECK(precision >= KNN_AUTO && precision <= KNN_FP8, "knn: unknown precision");
  const bool fp8 = precision == KNN_FP8;
  EVOC_CHECK(!fp8 || metric == 1, "knn: the fp8 precision supports the cosine metric only");
  EVOC_CHECK(k >= 1 && k <= KMAX, "knn: k must be in [1, 64]");
  EVOC_CHECK(d <= KNN_MAXD, "knn: d must be <= 4096");
  const bool wide = d > 1024;  // the re-rank and exact kernels' register-resident rows
  static bool attr_set = false;
  if (!attr_set) {
    // the streaming kernel: at most 134 KB of ring, key tile and thresholds (beside its
    // 512-byte static array of query indices)
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_kernel<true, 16>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_kernel<true, 32>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_kernel<false, 16>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_kernel<false, 32>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_kernel<false, 16, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_kernel<false, 32, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_ra_kernel<32, 1, 128, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)RaRing<128>::SMEM));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_ra_kernel<32, 1, 128, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)RaRing<128>::SMEM));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_ra_kernel<32, 1, 256, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)RaRing<256>::SMEM));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_ra_kernel<32, 1, 256, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)RaRing<256>::SMEM));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_ra_kernel<32, 2, 256, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)RaRing<256>::SMEM));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_ra_kernel<32, 2, 256, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)RaRing<256>::SMEM));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_rb_kernel<RB_NKC, RB_KC, true, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)RbRing<RB_KC>::SMEM));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_rb_kernel<RB_NKC, RB_KC, true, false, RB_PASS_C>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)RbRing<RB_KC>::SMEM));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_rb_kernel<RB_NKC, RB_KC, false, false, RB_PASS_C>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)RbRing<RB_KC>::SMEM));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_rb_kernel<RB_NKC, RB_KC, false, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)RbRing<RB_KC>::SMEM));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_rb_kernel<RB_NKC, RB_KC, true, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(RbRing<RB_KC>::SMEM + RB_QBYTES)));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_topc_rb_kernel<RB_NKC, RB_KC, false, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(RbRing<RB_KC>::SMEM + RB_QBYTES)));
    EVOC_HIP(hipFuncSetAttribute((const void*)knn_exact_merge_kernel<64>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 256 * 64 * 8));
    rs_set_attributes();
    attr_set = true;
  }
  // Two tile kernels: the resident-query kernel for bf16-exact inputs (d <= 512, k <= 16,
  // one list of 32 per row), else the 128x128 streaming kernel (split or not, two column
  // halves of CH per row). The prep layout depends on the choice; an AUTO guess of
  // bf16-exact that the prep flags refute is prepared again in the split layout.
  const int CH = (k <= 16) ? 16 : 32;
  const bool ra_possible = (precision != KNN_SPLIT) && !fp8 && d <= RA_MAXD && k <= 16;
  Buf<float> nrm, sqn, colscale;
  Buf<unsigned> flags(pool, 4);
  Buf<uint16_t> Xh, Xl;
  unsigned hflags[2] = {0, 0};
  int64_t npad = 0;
  int dpad = 0;
  bool ra = false, split = false;
  // the resident one-wavefront-per-SIMD kernel takes rows padded to a multiple of 256
  // dimensions (128 < d <= 1024): the first pass of split rows (d <= 512) and of bf16 rows
  // above 512, and the escalation's re-search
  const bool rs_layout = d > 128 && d <= 1024;
  auto prep = [&](bool use_ra, bool want_lo) {
    npad = use_ra ? ceil_div(n, RA_ROWS) * RA_ROWS : ceil_div(n, BM) * BM;
    dpad = use_ra ? (d <= 128 ? 128 : d <= 256 ? 256 : 512)
                  : rs_layout ? (int)(ceil_div(d, RS_KC) * RS_KC) : (int)(ceil_div(d, KC) * KC);
    Xh.alloc(pool, npad * dpad);
    if (want_lo) Xl.alloc(pool, npad * dpad);
    nrm.alloc(pool, npad);
    sqn.alloc(pool, npad);
    colscale.alloc(pool, npad);
    EVOC_HIP(hipMemsetAsync(flags.p, 0, 4 * sizeof(unsigned), s));
    knn_prep_kernel<<<(unsigned)std::min<int64_t>(ceil_div(npad, 4), 4096), 256, 0, s>>>(
        dX, n, d, ld, npad, dpad, Xh.p, want_lo ? Xl.p : nullptr, nrm.p, sqn.p, colscale.p,
        metric, flags.p);
    EVOC_HIP(hipGetLastError());
    EVOC_HIP(hipMemcpyAsync(hflags, flags.p, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
  };
  if (ra_possible) {
    prep(true, false);
    if (precision == KNN_BF16 || hflags[0] == 0) {
      ra = true;
    } else {  // AUTO and the input is not bf16-representable
      Xh.reset();
      nrm.reset();
      sqn.reset();
      colscale.reset();
      prep(false, true);
      split = true;
    }
  } else {
    // (the FP8 first pass is followed, for rows it cannot certify, by a bf16 search in
    // the AUTO form: one product when the input is bf16-representable, else the split)
    const bool want_lo = precision != KNN_BF16;
    prep(false, want_lo);
    split = want_lo && !((precision == KNN_AUTO || fp8) && hflags[0] == 0);
  }
  // FP8 copy, its key scales and the worst row's quantisation error
  int dpad8 = 0;
  Buf<uint8_t> Xq;
  Buf<float> scale8;
  float e_in8 = 0.0f;
  if (fp8) {
    dpad8 = (int)(ceil_div(d, 64) * 64);
    Xq.alloc(pool, npad * dpad8);
    scale8.alloc(pool, npad);
    knn_prep_fp8_kernel<<<(unsigned)std::min<int64_t>(ceil_div(npad, 4), 4096), 256, 0, s>>>(
        dX, n, d, ld, npad, dpad8, nrm.p, Xq.p, scale8.p, flags.p + 2);
    EVOC_HIP(hipGetLastError());
    unsigned rb = 0;
    EVOC_HIP(hipMemcpyAsync(&rb, flags.p + 2, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
    const float rho = __builtin_bit_cast(float, rb);
    // |x.y - s_x s_y q_x.q_y| <= ||x|| ||e_y|| + ||e_x|| ||y|| + ||e_x|| ||e_y||
    e_in8 = (2.0f * rho + rho * rho) * (1.0f + 1e-5f);
  }
  const int C = ra ? 32 : 2 * CH;
  Buf<int32_t> cidx(pool, n * C), frows(pool, n);
  Buf<float> ckey(pool, n * C);
  Buf<int> fcount(pool, 1);
  EVOC_HIP(hipMemsetAsync(fcount.p, 0, sizeof(int), s));
  const float u = 5.9604645e-08f;
  const float e_in = split ? 4.0f * 1.5258789e-05f : (hflags[0] ? 2.0f * 0.0078125f : 0.0f);
  const float gamma = 2.0f * (float)dpad * u + e_in;
  // the FP8 pass's: its MFMA sums the 64 products of an instruction with an error that
  // measured up to ~160 u of the sum of |products| (above the 2 d u of sequential f32
  // adds), hence the 512 u allowance
  const float gamma8 = (2.0f * (float)dpad8 + 512.0f) * u + e_in8;
  // The symmetric sweep (each pair of row blocks once) at large n, by the two-wavefront
  // kernel (bf16 rows, d in (256, 512]) or by the resident one (its first-pass forms: split
  // rows, bf16 rows above 512); EVOC_HIP_KNN_SYMMETRIC=1 / 0 forces it on / off.
  const bool rs_first = !ra && !fp8 && k <= 16 && rs_supported(split, 16, dpad);
  bool sym = false;
  if ((ra && dpad == 512) || (rs_first && rs_sym_supported(split, dpad))) {
    const char* env = std::getenv("EVOC_HIP_KNN_SYMMETRIC");
    sym = env ? (env[0] == '1') : (n >= KNN_SYM_MIN_N);
  }
  int64_t sym_S = 0, sym_Spad = 0, sym_qcap = 0;
  Buf<uint16_t> Xs, Xsl;
  Buf<float> ss, kt, qk;
  Buf<int32_t> qj, qr, qcount;
  // the sweep's slices (one workgroup per CU, EVOC_HIP_KNN_SLICED=0: one launch), its
  // queues (one per wavefront and launched workgroup) and the pieces of the last slice
  int64_t sym_blocks = 0, sym_gsz = 0, sym_rem = 0, sym_nq = 0, sym_prows = 0;
  int sym_parts = 1;
  Buf<int32_t> pidx;
  Buf<float> pkey, fbnd;
  // The first pass's certificate rate, estimated on evenly spaced sample rows searched
  // with the streaming kernel (bf16 or split, as the escalation would search them; the
  // FP8 pass certifies fewer rows than that). Where more than half of them fail (neighbour
  // distances concentrated below the error term, as in the dense leaves of a hierarchical
  // mixture), the full first pass would certify little: it is skipped and every row goes
  // straight to the sharded escalation. EVOC_HIP_KNN_SAMPLE=0 turns the test off.
  bool skip_first = false;
  if (!ra && !capture && n >= KNN_SAMPLE_MIN_N && C < 64) {
    const char* senv = std::getenv("EVOC_HIP_KNN_SAMPLE");
    if (!(senv && senv[0] == '0')) {
      const int64_t S = KNN_SAMPLE_ROWS, Spad = ceil_div(S, BM) * BM;
      std::vector<int32_t> hs(S);
      for (int64_t i = 0; i < S; ++i) hs[i] = (int32_t)(i * n / S);  // increasing: n >= S
      Buf<int32_t> sq(pool, S), scidx(pool, S * 32), sfr(pool, S);
      Buf<float> sckey(pool, S * 32);
      Buf<int> sfc(pool, 1);
      Buf<uint16_t> SQh(pool, Spad * dpad), SQl;
      if (split) SQl.alloc(pool, Spad * dpad);
      EVOC_HIP(hipMemcpyAsync(sq.p, hs.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, s));
      EVOC_HIP(hipMemsetAsync(sfc.p, 0, sizeof(int), s));
      knn_gather_rows_kernel<<<(unsigned)ceil_div(Spad * (dpad / 8), 256), 256, 0, s>>>(
          Xh.p, split ? Xl.p : nullptr, sq.p, S, Spad, dpad, SQh.p, split ? SQl.p : nullptr);
      Buf<float> sbnd;
      if (rs_supported(split, 16, dpad)) {
        // the resident kernel over column shards (the sample's few row blocks alone would
        // leave most CUs idle), merged into 32 candidates and a bound per row
        const int G = (int)std::max<int64_t>(
            1, std::min<int64_t>(KNN_SAMPLE_SHARDS_MAX, cu_count() / (Spad / RS_ROWS)));
        Buf<int32_t> gidx(pool, (int64_t)G * S * 32);
        Buf<float> gkey(pool, (int64_t)G * S * 32);
        sbnd.alloc(pool, S);
        rs_launch(split, 16, dpad, metric == 1, Xh.p, split ? Xl.p : nullptr, SQh.p,
                  split ? SQl.p : nullptr, sq.p, S, Spad, n, colscale.p, 0, (int)(npad / RA_BN), 0,
                  gidx.p, gkey.p, s, G);
        knn_shard_merge_kernel<<<(unsigned)ceil_div(S, 4), 256, 0, s>>>(
            gkey.p, gidx.p, G, S, S, sckey.p, scidx.p, sbnd.p);
      } else {
        const int nops = split ? 4 : 2;
        const int nst = split ? 2 : 4;
        const size_t smem = (size_t)nst * nops * BM * 4 * 16 + BM * SCS * sizeof(float) +
                            2 * BM * sizeof(float) + 4 * sizeof(int);
        if (split)
          knn_topc_kernel<true, 16><<<(unsigned)(Spad / BM), 256, smem, s>>>(
              Xh.p, Xl.p, n, npad, dpad, colscale.p, metric, scidx.p, sckey.p, SQh.p, SQl.p, sq.p,
              S, 0, npad / BN);
        else
          knn_topc_kernel<false, 16><<<(unsigned)(Spad / BM), 256, smem, s>>>(
              Xh.p, nullptr, n, npad, dpad, colscale.p, metric, scidx.p, sckey.p, SQh.p, nullptr,
              sq.p, S, 0, npad / BN);
      }
      EVOC_HIP(hipGetLastError());
      // (the sampled rows' results land in the output and are rewritten by the search)
      (wide ? knn_rerank_kernel<64> : knn_rerank_kernel<16>)<<<(unsigned)ceil_div(S, 4), 256, 0, s>>>(
          dX, n, d, ld, nrm.p, sqn.p, colscale.p, flags.p, scidx.p, sckey.p, 32, 2, sbnd.p, k,
          metric, gamma, d_idx, d_dist, sfc.p, sfr.p, sq.p, S);
      EVOC_HIP(hipGetLastError());
      int sfail = 0;
      EVOC_HIP(hipMemcpyAsync(&sfail, sfc.p, sizeof(int), hipMemcpyDeviceToHost, s));
      EVOC_HIP(hipStreamSynchronize(s));
      skip_first = 2 * (int64_t)sfail > S;
      if (verbose) {
        std::printf("  kNN sample: %d of %lld rows failed the certificate%s\n", sfail,
                    (long long)S, skip_first ? "; first pass skipped, every row escalated" : "");
        std::fflush(stdout);
      }
    }
  }
  if (skip_first) sym = false;
  if (sym) {
    sym_S = std::min<int64_t>(65536, std::max<int64_t>(256, n / 64 / 256 * 256));
    sym_S = std::min<int64_t>(sym_S, n);
    sym_Spad = ceil_div(sym_S, RA_BN) * RA_BN;
    Xs.alloc(pool, sym_Spad * dpad);
    ss.alloc(pool, sym_Spad);
    kt.alloc(pool, 2 * npad);
    rb_sample_kernel<<<(unsigned)ceil_div(sym_Spad * (dpad / 8), 256), 256, 0, s>>>(
        Xh.p, colscale.p, n, dpad, sym_S, sym_Spad, Xs.p, ss.p);
    if (split) {  // the sample's lo parts (its scales written again, identically)
      Xsl.alloc(pool, sym_Spad * dpad);
      rb_sample_kernel<<<(unsigned)ceil_div(sym_Spad * (dpad / 8), 256), 256, 0, s>>>(
          Xl.p, colscale.p, n, dpad, sym_S, sym_Spad, Xsl.p, ss.p);
    }
    // about 32 rows x n/2 columns x RB_THR_RANK/S reverse keys per wavefront; three times that
    const int64_t expect = 16 * RB_THR_RANK * ceil_div(n, sym_S);
    sym_qcap = std::max<int64_t>(16384, ceil_div(3 * expect, 256) * 256);
    // Slices of one workgroup per CU; when the last holds fewer blocks than CUs, each of
    // its blocks' walks is split into `parts` pieces (at most 8: the piece merge's width),
    // one workgroup each, so that the slice fills the CUs it would leave idle.
    const int64_t brows = ra ? RB_ROWS : RS_ROWS, bwaves = ra ? RB_WAVES : RS_WAVES;
    sym_blocks = npad / brows;
    const char* slenv = std::getenv("EVOC_HIP_KNN_SLICED");
    const bool sliced = !(slenv && slenv[0] == '0');
    sym_gsz = sliced ? std::max<int64_t>(1, cu_count()) : sym_blocks;
    sym_rem = sym_blocks % sym_gsz;
    sym_parts = sym_rem > 0 ? (int)std::max<int64_t>(
                                  1, std::min<int64_t>(KNN_SAMPLE_SHARDS_MAX, sym_gsz / sym_rem))
                            : 1;
    sym_nq = (sym_blocks + sym_rem * (sym_parts - 1)) * bwaves;
    qj.alloc(pool, sym_nq * sym_qcap);
    qr.alloc(pool, sym_nq * sym_qcap);
    qk.alloc(pool, sym_nq * sym_qcap);
    qcount.alloc(pool, sym_nq);
    if (sym_parts > 1) {  // the last slice holds blocks 0 .. rem - 1 (descending order)
      sym_prows = sym_rem * brows;
      pidx.alloc(pool, sym_parts * sym_prows * 32);
      pkey.alloc(pool, sym_parts * sym_prows * 32);
      fbnd.alloc(pool, sym_prows);
    }
  }
  hipEvent_t e0, e1;
  EVOC_HIP(hipEventCreate(&e0));
  EVOC_HIP(hipEventCreate(&e1));
  EVOC_HIP(hipEventRecord(e0, s));
  if (skip_first) {
  } else if (ra) {
    const unsigned grid = (unsigned)(npad / RA_ROWS);
#define EVOC_RA_LAUNCH(NKC_, KC_)                                                              \
  do {                                                                                         \
    if (metric == 1)                                                                           \
      knn_topc_ra_kernel<32, NKC_, KC_, true><<<grid, 256, RaRing<KC_>::SMEM, s>>>(            \
          Xh.p, n, npad, dpad, colscale.p, cidx.p, ckey.p);                                    \
    else                                                                                       \
      knn_topc_ra_kernel<32, NKC_, KC_, false><<<grid, 256, RaRing<KC_>::SMEM, s>>>(           \
          Xh.p, n, npad, dpad, colscale.p, cidx.p, ckey.p);                                    \
  } while (0)
    if (dpad == 128) {
      EVOC_RA_LAUNCH(1, 128);
    } else if (dpad == 256) {
      EVOC_RA_LAUNCH(1, 256);
    } else {  // d in (256, 512]: eight 32-row wavefronts, two per SIMD
      RbCols cols{Xh.p, colscale.p, nullptr, n, npad / RA_BN, 1, nullptr, nullptr, nullptr, 0,
                  nullptr, 0, 0, 0};
      if (sym) {
        // threshold pass against a sample of rows, then the symmetric sweep
        RbCols sc{Xs.p, ss.p, nullptr, sym_S, sym_Spad / RA_BN, 0, nullptr, nullptr, nullptr, 0,
                  nullptr, 0, 0, 0};
        if (metric == 1)
          knn_topc_rb_kernel<RB_NKC, RB_KC, true, false, RB_PASS_C><<<grid, 512, RbRing<RB_KC>::SMEM, s>>>(
              Xh.p, n, npad, dpad, colscale.p, cidx.p, ckey.p, sc);
        else
          knn_topc_rb_kernel<RB_NKC, RB_KC, false, false, RB_PASS_C><<<grid, 512, RbRing<RB_KC>::SMEM, s>>>(
              Xh.p, n, npad, dpad, colscale.p, cidx.p, ckey.p, sc);
        rb_thr0_kernel<RB_PASS_C><<<(unsigned)ceil_div(npad, 256), 256, 0, s>>>(ckey.p, colscale.p, n, npad,
                                                                    kt.p);
        EVOC_HIP(hipEventRecord(e0, s));
        cols.kt = kt.p;
        cols.qj = qj.p;
        cols.qr = qr.p;
        cols.qk = qk.p;
        cols.qcap = sym_qcap;
        cols.qcount = qcount.p;
        cols.group = cu_count();
        // Launched in slices of one workgroup per CU, so that each slice's walks start
        // together and stream their shared column tiles in step through L2 (in one launch
        // a workgroup starts whenever a CU frees up). EVOC_HIP_KNN_SLICED=0: one launch.
        cols.nblocks = (int)grid;
        cols.prows = sym_prows;
        cols.plist_idx = pidx.p;
        cols.plist_key = pkey.p;
        for (int b0 = 0; b0 < (int)grid; b0 += (int)sym_gsz) {
          cols.block0 = b0;
          const int nb = (int)std::min<int64_t>(sym_gsz, (int64_t)grid - b0);
          cols.parts = nb < sym_gsz ? sym_parts : 1;
          const unsigned g1 = (unsigned)(nb * cols.parts);
          if (metric == 1)
            knn_topc_rb_kernel<RB_NKC, RB_KC, true, true><<<g1, 512, RbRing<RB_KC>::SMEM + RB_QBYTES, s>>>(
                Xh.p, n, npad, dpad, colscale.p, cidx.p, ckey.p, cols);
          else
            knn_topc_rb_kernel<RB_NKC, RB_KC, false, true><<<g1, 512, RbRing<RB_KC>::SMEM + RB_QBYTES, s>>>(
                Xh.p, n, npad, dpad, colscale.p, cidx.p, ckey.p, cols);
        }
      } else {
        if (metric == 1)
          knn_topc_rb_kernel<RB_NKC, RB_KC, true, false><<<grid, 512, RbRing<RB_KC>::SMEM, s>>>(
              Xh.p, n, npad, dpad, colscale.p, cidx.p, ckey.p, cols);
        else
          knn_topc_rb_kernel<RB_NKC, RB_KC, false, false><<<grid, 512, RbRing<RB_KC>::SMEM, s>>>(
              Xh.p, n, npad, dpad, colscale.p, cidx.p, ckey.p, cols);
      }
    }
#undef EVOC_RA_LAUNCH
  } else if (rs_first) {
    if (sym) {
      // threshold pass against the sample (16 per column half: RB_THR_RANK of them are
      // read), then the symmetric sweep, in slices of one workgroup per CU as above
      rs_launch(split, 16, dpad, metric == 1, Xs.p, split ? Xsl.p : nullptr, Xh.p, Xl.p, nullptr,
                n, npad, sym_S, ss.p, 0, (int)(sym_Spad / RA_BN), 0, cidx.p, ckey.p, s);
      rb_thr0_kernel<16><<<(unsigned)ceil_div(npad, 256), 256, 0, s>>>(ckey.p, colscale.p, n, npad,
                                                                     kt.p);
      EVOC_HIP(hipGetLastError());
      EVOC_HIP(hipEventRecord(e0, s));
      RbCols cols{};
      cols.kt = kt.p;
      cols.qj = qj.p;
      cols.qr = qr.p;
      cols.qk = qk.p;
      cols.qcap = sym_qcap;
      cols.qcount = qcount.p;
      cols.group = cu_count();
      const int nblk = (int)(npad / RS_ROWS);
      cols.nblocks = nblk;
      cols.prows = sym_prows;
      cols.plist_idx = pidx.p;
      cols.plist_key = pkey.p;
      for (int b0 = 0; b0 < nblk; b0 += (int)sym_gsz) {
        cols.block0 = b0;
        const int nb = (int)std::min<int64_t>(sym_gsz, (int64_t)nblk - b0);
        cols.parts = nb < sym_gsz ? sym_parts : 1;
        rs_sym_launch(split, dpad, metric == 1, Xh.p, Xl.p, n, npad, colscale.p,
                      (unsigned)(nb * cols.parts), cols, cidx.p, ckey.p, s);
      }
    } else {
      rs_launch(split, 16, dpad, metric == 1, Xh.p, Xl.p, Xh.p, Xl.p, nullptr, n, npad, n,
                colscale.p, 0, (int)(npad / RA_BN), 1, cidx.p, ckey.p, s);
    }
  } else if (fp8) {
    const size_t smem = (size_t)4 * 2 * BM * 4 * 16 + BM * SCS * sizeof(float) +
                        2 * BM * sizeof(float) + 4 * sizeof(int);
    const unsigned grid = (unsigned)(npad / BM);
    const uint16_t* xq = reinterpret_cast<const uint16_t*>(Xq.p);  // bytes; the kernel reads bytes
    if (CH == 16)
      knn_topc_kernel<false, 16, true><<<grid, 256, smem, s>>>(xq, nullptr, n, npad, dpad8, scale8.p,
                                                               metric, cidx.p, ckey.p, nullptr,
                                                               nullptr, nullptr, 0, 0, npad / BN);
    else
      knn_topc_kernel<false, 32, true><<<grid, 256, smem, s>>>(xq, nullptr, n, npad, dpad8, scale8.p,
                                                               metric, cidx.p, ckey.p, nullptr,
                                                               nullptr, nullptr, 0, 0, npad / BN);
  } else {
    const int nops = split ? 4 : 2;
    const int nst = split ? 2 : 4;
    const size_t smem = (size_t)nst * nops * BM * 4 * 16 + BM * SCS * sizeof(float) +
                        2 * BM * sizeof(float) + 4 * sizeof(int);
    const unsigned grid = (unsigned)(npad / BM);
    if (split) {
      if (CH == 16)
        knn_topc_kernel<true, 16><<<grid, 256, smem, s>>>(Xh.p, Xl.p, n, npad, dpad, colscale.p,
                                                          metric, cidx.p, ckey.p, nullptr,
                                                          nullptr, nullptr, 0, 0, npad / BN);
      else
        knn_topc_kernel<true, 32><<<grid, 256, smem, s>>>(Xh.p, Xl.p, n, npad, dpad, colscale.p,
                                                          metric, cidx.p, ckey.p, nullptr,
                                                          nullptr, nullptr, 0, 0, npad / BN);
    } else {
      if (CH == 16)
        knn_topc_kernel<false, 16><<<grid, 256, smem, s>>>(Xh.p, nullptr, n, npad, dpad,
                                                           colscale.p, metric, cidx.p, ckey.p,
                                                           nullptr, nullptr, nullptr, 0, 0, npad / BN);
      else
        knn_topc_kernel<false, 32><<<grid, 256, smem, s>>>(Xh.p, nullptr, n, npad, dpad,
                                                           colscale.p, metric, cidx.p, ckey.p,
                                                           nullptr, nullptr, nullptr, 0, 0, npad / BN);
    }
  }
  EVOC_HIP(hipGetLastError());
  EVOC_HIP(hipEventRecord(e1, s));
  Buf<float> mkey, bnd;
  Buf<int32_t> midx;
  if (sym) {
    const int64_t nwaves = sym_nq;
    std::vector<int32_t> hq(nwaves);
    EVOC_HIP(hipMemcpyAsync(hq.data(), qcount.p, nwaves * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
    int64_t qmax = 0;
    for (auto q : hq) qmax = std::max<int64_t>(qmax, q);
    if (qmax > sym_qcap) {  // a queue overflowed: sweep every pair instead
      sym = false;
      if (ra) {
        RbCols cols{Xh.p, colscale.p, nullptr, n, npad / RA_BN, 1, nullptr, nullptr, nullptr, 0,
                    nullptr, 0, 0, 0};
        const unsigned grid = (unsigned)(npad / RA_ROWS);
        if (metric == 1)
          knn_topc_rb_kernel<RB_NKC, RB_KC, true, false><<<grid, 512, RbRing<RB_KC>::SMEM, s>>>(
              Xh.p, n, npad, dpad, colscale.p, cidx.p, ckey.p, cols);
        else
          knn_topc_rb_kernel<RB_NKC, RB_KC, false, false><<<grid, 512, RbRing<RB_KC>::SMEM, s>>>(
              Xh.p, n, npad, dpad, colscale.p, cidx.p, ckey.p, cols);
      } else {
        rs_launch(split, 16, dpad, metric == 1, Xh.p, Xl.p, Xh.p, Xl.p, nullptr, n, npad, n,
                  colscale.p, 0, (int)(npad / RA_BN), 1, cidx.p, ckey.p, s);
      }
      EVOC_HIP(hipGetLastError());
    } else {
      Buf<int32_t> rcnt(pool, n), ridx(pool, n * RB_RCAP), dropped(pool, n);
      Buf<float> rkey(pool, n * RB_RCAP), fmax(pool, n);
      EVOC_HIP(hipMemsetAsync(rcnt.p, 0, n * sizeof(int32_t), s));
      // the last slice's split walks: their pieces' lists merged into the forward lists
      const int64_t fb_rows = sym_parts > 1 ? std::min<int64_t>(n, sym_prows) : 0;
      if (fb_rows > 0)
        knn_shard_merge_kernel<<<(unsigned)ceil_div(fb_rows, 4), 256, 0, s>>>(
            pkey.p, pidx.p, sym_parts, fb_rows, sym_prows, ckey.p, cidx.p, fbnd.p);
      rb_fwdmax_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(ckey.p, cidx.p, n, fmax.p,
                                                                 dropped.p);
      if (metric == 1)
        rb_scatter_kernel<true><<<(unsigned)nwaves, 256, 0, s>>>(
            qj.p, qr.p, qk.p, qcount.p, sym_qcap, RB_RCAP, colscale.p, fmax.p, dropped.p, rcnt.p,
            rkey.p, ridx.p);
      else
        rb_scatter_kernel<false><<<(unsigned)nwaves, 256, 0, s>>>(
            qj.p, qr.p, qk.p, qcount.p, sym_qcap, RB_RCAP, colscale.p, fmax.p, dropped.p, rcnt.p,
            rkey.p, ridx.p);
      mkey.alloc(pool, n * 32);
      midx.alloc(pool, n * 32);
      bnd.alloc(pool, n);
      rb_merge_kernel<<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(ckey.p, cidx.p, rcnt.p, rkey.p,
                                                              ridx.p, kt.p, dropped.p, n,
                                                              mkey.p, midx.p, bnd.p, fbnd.p,
                                                              fb_rows);
      EVOC_HIP(hipGetLastError());
    }
  }
  if (capture) {  // the first pass's raw lists, for tests of the tile kernels
    capture->C = C;
    capture->idx.resize((size_t)n * C);
    capture->key.resize((size_t)n * C);
    capture->scale.resize((size_t)n);
    capture->gamma = fp8 ? gamma8 : gamma;
    EVOC_HIP(hipMemcpyAsync(capture->idx.data(), cidx.p, (size_t)n * C * sizeof(int32_t),
                            hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(capture->key.data(), ckey.p, (size_t)n * C * sizeof(float),
                            hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(capture->scale.data(), fp8 ? scale8.p : colscale.p, n * sizeof(float),
                            hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
  }
  int nfail = 0;
  if (skip_first) {  // every row is escalated
    knn_iota_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(n, frows.p);
    EVOC_HIP(hipGetLastError());
    nfail = (int)n;
  } else {
    // the FP8 pass's keys and error term are in its own units (row scales s / ||x||)
    (wide ? knn_rerank_kernel<64> : knn_rerank_kernel<16>)<<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(
        dX, n, d, ld, nrm.p, sqn.p, fp8 ? scale8.p : colscale.p, flags.p, sym ? midx.p : cidx.p,
        sym ? mkey.p : ckey.p, C, (ra && dpad < 512) ? 1 : 2, sym ? bnd.p : nullptr, k, metric,
        fp8 ? gamma8 : gamma, d_idx, d_dist, fcount.p, frows.p, nullptr, 0);
    EVOC_HIP(hipGetLastError());
    EVOC_HIP(hipMemcpyAsync(&nfail, fcount.p, sizeof(int), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
  }
  if (verbose && !skip_first) {
    float ms = 0.0f;
    EVOC_HIP(hipEventElapsedTime(&ms, e0, e1));
    std::printf("  kNN tile pass %.0f ms (%d candidates per row%s%s); %d of %lld rows failed the "
                "certificate\n", ms, C, sym ? ", symmetric sweep" : "", fp8 ? ", fp8" : "", nfail,
                (long long)n);
    std::fflush(stdout);
  }
  int escalated = 0;
  const int first_failed = nfail;
  // EVOC_HIP_KNN_ESCALATE=1 escalates any number of failing rows (tests of small inputs)
  const char* esc_env = std::getenv("EVOC_HIP_KNN_ESCALATE");
  const int esc_min = (esc_env && esc_env[0] == '1') ? 1 : KNN_ESCALATE_MIN;
  if (nfail >= esc_min && C < 64) {
    // Many certificates failed (neighbour distances concentrated, as in a dense leaf of a
    // hierarchical mixture): search those rows again, gathered, over KNN_ESC_SHARDS column
    // shards with 32 candidates per column half each, and re-rank every candidate; the
    // row's bound is then the smallest of the shards' bounds, near rank 64 * shards. Only
    // rows that fail again take the exact scan.
    escalated = nfail;
    std::vector<int32_t> hf(nfail);
    EVOC_HIP(hipMemcpyAsync(hf.data(), frows.p, nfail * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
    std::sort(hf.begin(), hf.end());
    const int64_t nq = nfail, nqpad = ceil_div(nq, BM) * BM;
    Buf<int32_t> qmap(pool, nq), ecidx(pool, KNN_ESC_SHARDS * nq * 64), frows2(pool, nq);
    Buf<float> eckey(pool, KNN_ESC_SHARDS * nq * 64);
    Buf<uint16_t> Qh(pool, nqpad * dpad), Ql;
    if (split) Ql.alloc(pool, nqpad * dpad);
    EVOC_HIP(hipMemcpyAsync(qmap.p, hf.data(), nq * sizeof(int32_t), hipMemcpyHostToDevice, s));
    const int64_t pieces = nqpad * (dpad / 8);
    knn_gather_rows_kernel<<<(unsigned)ceil_div(pieces, 256), 256, 0, s>>>(
        Xh.p, split ? Xl.p : nullptr, qmap.p, nq, nqpad, dpad, Qh.p, split ? Ql.p : nullptr);
    const int nops = split ? 4 : 2;
    const int nst = split ? 2 : 4;
    const size_t smem = (size_t)nst * nops * BM * 4 * 16 + BM * SCS * sizeof(float) +
                        2 * BM * sizeof(float) + 4 * sizeof(int);
    const unsigned grid = (unsigned)(nqpad / BM);
    const int64_t ntile = npad / BN;
    if (skip_first) EVOC_HIP(hipEventRecord(e0, s));  // then the escalation is the tile pass
    if (rs_supported(split, 32, dpad) && nqpad / RS_ROWS < cu_count()) {
      // few rows: the shards in one launch (grid rows), so that together they fill the CUs
      rs_launch(split, 32, dpad, metric == 1, Xh.p, split ? Xl.p : nullptr, Qh.p,
                split ? Ql.p : nullptr, qmap.p, nq, nqpad, n, colscale.p, 0, (int)(npad / RA_BN), 0,
                ecidx.p, eckey.p, s, KNN_ESC_SHARDS);
    } else if (rs_supported(split, 32, dpad)) {  // one launch per shard: each streams its columns
      for (int sh = 0; sh < KNN_ESC_SHARDS; ++sh) {
        const int64_t ct0 = ntile * sh / KNN_ESC_SHARDS, ct1 = ntile * (sh + 1) / KNN_ESC_SHARDS;
        rs_launch(split, 32, dpad, metric == 1, Xh.p, split ? Xl.p : nullptr, Qh.p,
                  split ? Ql.p : nullptr, qmap.p, nq, nqpad, n, colscale.p,
                  (int)(ct0 * (BN / RA_BN)), (int)(ct1 * (BN / RA_BN)), 0,
                  ecidx.p + (int64_t)sh * nq * 64, eckey.p + (int64_t)sh * nq * 64, s);
      }
    } else {  // the streaming kernel, one launch per shard
      for (int sh = 0; sh < KNN_ESC_SHARDS; ++sh) {
        const int64_t ct0 = ntile * sh / KNN_ESC_SHARDS, ct1 = ntile * (sh + 1) / KNN_ESC_SHARDS;
        int32_t* oi = ecidx.p + (int64_t)sh * nq * 64;
        float* ok = eckey.p + (int64_t)sh * nq * 64;
        if (split)
          knn_topc_kernel<true, 32><<<grid, 256, smem, s>>>(Xh.p, Xl.p, n, npad, dpad, colscale.p,
                                                            metric, oi, ok, Qh.p, Ql.p, qmap.p,
                                                            nq, ct0, ct1);
        else
          knn_topc_kernel<false, 32><<<grid, 256, smem, s>>>(Xh.p, nullptr, n, npad, dpad,
                                                             colscale.p, metric, oi, ok, Qh.p,
                                                             nullptr, qmap.p, nq, ct0, ct1);
      }
    }
    EVOC_HIP(hipGetLastError());
    if (skip_first) EVOC_HIP(hipEventRecord(e1, s));
    EVOC_HIP(hipMemsetAsync(fcount.p, 0, sizeof(int), s));
    (wide ? knn_rerank_all_kernel<KNN_ESC_SHARDS, 64>
          : knn_rerank_all_kernel<KNN_ESC_SHARDS, 16>)<<<(unsigned)ceil_div(nq, 4), 256, 0, s>>>(
        dX, n, d, ld, nrm.p, sqn.p, colscale.p, flags.p, ecidx.p, eckey.p, k, metric, gamma,
        d_idx, d_dist, fcount.p, frows2.p, qmap.p, nq);
    EVOC_HIP(hipGetLastError());
    EVOC_HIP(hipMemcpyAsync(&nfail, fcount.p, sizeof(int), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
    if (nfail > 0)
      EVOC_HIP(hipMemcpyAsync(frows.p, frows2.p, nfail * sizeof(int32_t),
                              hipMemcpyDeviceToDevice, s));
    if (verbose) {
      std::printf("  kNN escalation of %d rows over %d column shards done; %d rows go to the "
                  "exact search\n", escalated, KNN_ESC_SHARDS, nfail);
      std::fflush(stdout);
    }
  }
  if (nfail > 0) {
    // the columns are split over `parts` workgroups of four wavefronts, ex_rows failing
    // rows per launch (8 held in registers, 2 for wide rows); the per-wavefront lists of
    // each row are then merged
    const int parts = 1024;
    const int64_t span = ceil_div(n, (int64_t)parts);
    const int ex_rows = wide ? 2 : 8;
    const int nb = std::min(nfail, ex_rows);
    Buf<float> pk(pool, (size_t)nb * parts * 4 * k);
    Buf<int32_t> pi(pool, (size_t)nb * parts * 4 * k);
    for (int f0 = 0; f0 < nfail; f0 += ex_rows) {
      const int rows = std::min(ex_rows, nfail - f0);
      const bool k64 = k > 32;
      auto exact = wide ? (k64 ? knn_exact_kernel<64, 2, 64> : knn_exact_kernel<64, 2, 32>)
                        : (k64 ? knn_exact_kernel<16, 8, 64> : knn_exact_kernel<16, 8, 32>);
      exact<<<(unsigned)parts, 256, 0, s>>>(dX, n, d, ld, nrm.p, k, metric, frows.p + f0, rows,
                                            parts, span, pk.p, pi.p);
      if (k64)
        knn_exact_merge_kernel<64><<<(unsigned)rows, 256, 256 * 64 * 8, s>>>(
            k, frows.p + f0, parts * 4, pk.p, pi.p, d_idx, d_dist);
      else
        knn_exact_merge_kernel<32><<<(unsigned)rows, 256, 256 * 32 * 8, s>>>(
            k, frows.p + f0, parts * 4, pk.p, pi.p, d_idx, d_dist);
      if (verbose && f0 + rows < nfail && ((f0 / ex_rows) + 1) % 4096 == 0) {
        EVOC_HIP(hipStreamSynchronize(s));
        std::printf("  kNN exact search: %d of %d rows done\n", f0 + rows, nfail);
        std::fflush(stdout);
      }
    }
    EVOC_HIP(hipGetLastError());
  }
  float mfma_ms = 0.0f;
  EVOC_HIP(hipEventElapsedTime(&mfma_ms, e0, e1));
  EVOC_HIP(hipEventDestroy(e0));
  EVOC_HIP(hipEventDestroy(e1));
  KnnStats st;
  st.split = split ? 1 : 0;
  st.fallback_rows = nfail;
  st.escalated_rows = escalated;
  st.first_pass_failed = first_failed;
  st.candidates = C;
  st.mfma_ms = mfma_ms;
  double pairs = (double)npad * (double)npad;
  if (sym) {  // block I swept nblk(I) blocks of 256 (two-wavefront) or 128 (resident) columns
    const int64_t br = ra ? RB_ROWS : RS_ROWS, m = npad / br;
    int64_t blocks = 0;
    for (int64_t I = 0; I < m; ++I) blocks += (m & 1) ? (m + 1) / 2 : m / 2 + (I < m / 2 ? 1 : 0);
    pairs = (double)blocks * br * br;
  }
  st.flops = fp8 ? 2.0 * pairs * (double)dpad8 : 2.0 * pairs * (double)dpad * (split ? 3.0 : 1.0);
  if (skip_first)  // the escalation of every row over all columns
    st.flops = 2.0 * (double)ceil_div(n, BM) * BM * (double)npad * (double)dpad * (split ? 3.0 : 1.0);
  st.symmetric = sym ? 1 : 0;
  // a stage-level search (one that asks for its statistics) becomes the thread's last kNN;
  // internal searches, such as the warm start's subsample, do not replace it
  if (stats) {
    *stats = st;
    last_knn_stats() = st;
  }
}

KnnStats& last_knn_stats() {
  static thread_local KnnStats st;
  return st;
}

}  // namespace evoc
