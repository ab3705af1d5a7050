// EVoC node embedding (evoc_embedding, src/graph/embedding.rs:48-298) on gfx950.
//
// Every node and epoch reads the pre-epoch layout (Jacobi update, as the reference's
// parallel gradient pass does) and writes the next layout into a second buffer. The
// per-edge schedule (next sample / next negative epoch) is stored once per CSR slot: the
// two slots of an edge start equal and are advanced by the same f32 operations by their
// two owners, so they stay identical without a shared write and without a separate
// schedule pass (embedding.rs:275-286). Negative samples are drawn from Philox keyed by
// (seed, node, epoch, draw). Adam moments live in HBM next to the layout. All arithmetic
// is f32 in the reference's operation order.
//
// A node's lanes load a chunk of its row's slots at a time and compute every gradient
// term of the chunk (the pull of each sampled edge and each negative sample, with the
// per-node draw numbering) in parallel, stage the terms in LDS, and the lanes of the
// layout dimensions add them in slot order, the order of the reference's per-node loop.
// A node takes 8 lanes (at least dim; chunks of 8 slots), a wavefront with 64 or more
// edges (chunks of 64), or a workgroup of 16 wavefronts with 1024 or more (16 chunks of
// 64 staged at once); the three kernels give bit-identical layouts.
#include "api_internal.h"
#include "kernels.h"

namespace evoc {
namespace {

constexpr int MAXD = 16;
constexpr int GSEG = 8;  // lanes per node of the grouped epoch kernel (at least dim)
constexpr int WIDE_CAP = 256;  // gradient terms staged per wavefront at a time
constexpr int HUB_DEG = 1024;  // nodes with at least this many edges take a whole workgroup
constexpr int HUB_WAVES = 16;  // wavefronts per hub node: one 64-slot chunk each at a time
constexpr int HUB_CAP = 1024;  // gradient terms staged per hub workgroup at a time

__global__ void max_weight_kernel(const float* __restrict__ w, int64_t nnz,
                                  unsigned* __restrict__ out) {
  unsigned m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, __float_as_uint(fmaxf(w[i], 0.0f)));
  // wave max then one atomic per wave
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void schedule_init_kernel(const float* __restrict__ w, int64_t nnz,
                                     const unsigned* __restrict__ maxw_bits, float neg_factor,
                                     float* __restrict__ eps_s, float* __restrict__ next_s,
                                     float* __restrict__ next_n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const float maxw = __uint_as_float(*maxw_bits);
  const float norm = w[i] / maxw;
  const float e = norm > 0.0f ? 1.0f / norm : 1e8f;
  eps_s[i] = e;
  next_s[i] = e;
  next_n[i] = e * neg_factor;
}

__global__ void random_init_kernel(float* __restrict__ e, int64_t n, int dim, uint64_t seed) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * dim) return;
  const int64_t i = q / dim;
  const int c = (int)(q - i * dim);
  e[q] = (float)(-0.25 + 0.5 * Philox::unit(seed, (uint32_t)i, (uint32_t)c, 0, TAG_EINI));
}

// One wavefront per node with at least wide_deg edges, 64 slots at a time.
template <int D>
__global__ __launch_bounds__(256) void epoch_wide_kernel(
    int64_t n, int dim, const int64_t* __restrict__ indptr, const int32_t* __restrict__ other,
    float* __restrict__ next_s, const float* __restrict__ eps_s, float* __restrict__ next_n,
    float neg_factor, const float* __restrict__ e_old, float* __restrict__ e_new,
    float* __restrict__ m, float* __restrict__ v, float epoch_t, uint32_t epoch, float lr_ad,
    float epsc, float gamma, float noise, float one_m_b1, float one_m_b2, uint64_t seed,
    const int32_t* __restrict__ wide_nodes, int64_t nwide) {
  constexpr int DD = D > 0 ? D : MAXD;
  __shared__ float vbuf[4][WIDE_CAP * DD];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t w = (int64_t)blockIdx.x * 4 + wave;
  if (w >= nwide) return;  // whole wavefronts only: no workgroup barrier below
  const int64_t i = wide_nodes[w];
  const int nd = D > 0 ? D : dim;
  float ei[DD];
#pragma unroll
  for (int c = 0; c < DD; ++c) ei[c] = c < nd ? e_old[i * nd + c] : 0.0f;
  float gl = 0.0f;  // lane c < nd: gradient component c
  bool touched = false;
  uint32_t draw_carry = 0;
  const int64_t s0 = indptr[i], s1 = indptr[i + 1];
  float* vb = vbuf[wave];
  // Loads run ahead of the 64-slot chunks: a chunk's slot inputs two chunks ahead, its
  // positive endpoint row one chunk ahead, so inside a chunk only the negative samples'
  // rows are waited for. Values are read for lanes past the row end or not due this epoch
  // only where the slot exists, and used only for due slots.
  float ns_a = 0.0f, ps_a = 0.0f, nn_a = 0.0f, ns_b = 0.0f, ps_b = 0.0f, nn_b = 0.0f;
  int32_t j_a = 0, j_b = 0;  // _a: the current chunk, _b: the next one
  float eo_a[DD];
  if (s0 + lane < s1) {
    ns_a = next_s[s0 + lane];
    j_a = other[s0 + lane];
    ps_a = eps_s[s0 + lane];
    nn_a = next_n[s0 + lane];
  }
  if (s0 + 64 + lane < s1) {
    ns_b = next_s[s0 + 64 + lane];
    j_b = other[s0 + 64 + lane];
    ps_b = eps_s[s0 + 64 + lane];
    nn_b = next_n[s0 + 64 + lane];
  }
  {
    const bool due = s0 + lane < s1 && !(ns_a > epoch_t);
#pragma unroll
    for (int c = 0; c < DD; ++c) eo_a[c] = due && c < nd ? e_old[(int64_t)j_a * nd + c] : 0.0f;
  }
  for (int64_t sb = s0; sb < s1; sb += 64) {
    const int64_t s = sb + lane;
    const bool valid = s < s1;
    const float ns = valid ? ns_a : 0.0f;
    const bool active = valid && !(ns > epoch_t);
    int32_t j = 0;
    float ps = 0.0f, pn = 0.0f, nn = 0.0f;
    uint64_t cnt = 0;
    if (active) {
      j = j_a;
      ps = ps_a;
      pn = ps * neg_factor;
      nn = nn_a;
      const float f = floorf((epoch_t - nn) / pn);
      cnt = (f >= 0.0f && f < 1.8e19f) ? (uint64_t)f : 0;
    }
    float eo[DD];
#pragma unroll
    for (int c = 0; c < DD; ++c) eo[c] = eo_a[c];
    // the next chunk's positive rows and the slot inputs of the one after it
    {
      const bool due = s + 64 < s1 && !(ns_b > epoch_t);
#pragma unroll
      for (int c = 0; c < DD; ++c) eo_a[c] = due && c < nd ? e_old[(int64_t)j_b * nd + c] : 0.0f;
      ns_a = ns_b;
      j_a = j_b;
      ps_a = ps_b;
      nn_a = nn_b;
      if (s + 128 < s1) {
        ns_b = next_s[s + 128];
        j_b = other[s + 128];
        ps_b = eps_s[s + 128];
        nn_b = next_n[s + 128];
      }
    }
    touched |= __ballot(active) != 0;
    const uint32_t nneg = (active && (int64_t)j > i) ? (uint32_t)cnt : 0u;
    const uint32_t nvec = active ? 1u + nneg : 0u;
    // exclusive prefix sums over the lanes (slot order) of draws and of staged terms
    uint32_t dinc = nneg, vinc = nvec;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t td = __shfl_up(dinc, o, 64), tv = __shfl_up(vinc, o, 64);
      if (lane >= o) {
        dinc += td;
        vinc += tv;
      }
    }
    const uint32_t draw0 = draw_carry + dinc - nneg;
    const uint32_t voff = vinc - nvec;
    const uint32_t total = __shfl(vinc, 63, 64);
    draw_carry += __shfl(dinc, 63, 64);
    // this lane's positive term
    float coeff = 0.0f;
    bool pos_ok = false;
    if (active) {
      float dsq = 0.0f;
#pragma unroll
      for (int c = 0; c < DD; ++c)
        if (c < nd) {
          const float df = ei[c] - eo[c];
          const float p = df * df;
          dsq = dsq + p;
        }
      if (dsq >= 1e-8f) {
        const float dist = sqrtf(dsq);
        const float num = -((2.0f * noise) * dist) - 2.0f;
        const float den = ((2.0f * dsq) - (0.5f * dist)) + 1.0f;
        coeff = num / den;
        pos_ok = true;
      }
    }
    for (uint32_t w0 = 0; w0 < total; w0 += WIDE_CAP) {
      // terms of this lane that fall in the window [w0, w0 + WIDE_CAP)
      const uint32_t lo = voff > w0 ? voff : w0;
      const uint32_t hi = min(voff + nvec, w0 + (uint32_t)WIDE_CAP);
      for (uint32_t t = lo; t < hi; ++t) {
        const uint32_t q = t - voff;
        const int slot = (int)(t - w0);
        bool good = false;
        if (q == 0) {
          good = pos_ok;
          if (good) {
#pragma unroll
            for (int c = 0; c < DD; ++c)
              if (c < nd) vb[slot * DD + c] = coeff * (ei[c] - eo[c]);
          }
        } else {
          const uint32_t r = Philox::first(seed, draw0 + (q - 1), epoch, (uint32_t)i, TAG_NEG);
          const int64_t kk = Philox::range(r, (uint64_t)n);
          if (kk != i) {
            float ek[DD];
#pragma unroll
            for (int c = 0; c < DD; ++c) ek[c] = c < nd ? e_old[kk * nd + c] : 0.0f;
            float dk = 0.0f;
#pragma unroll
            for (int c = 0; c < DD; ++c)
              if (c < nd) {
                const float df = ei[c] - ek[c];
                const float p = df * df;
                dk = dk + p;
              }
            if (dk > 0.01f) {
              good = true;
              const float gc = (gamma * 4.0f) / ((1.0f + 0.25f * dk) * dk);
#pragma unroll
              for (int c = 0; c < DD; ++c)
                if (c < nd) {
                  const float diff = ei[c] - ek[c];
                  vb[slot * DD + c] = fminf(fmaxf(gc * diff, -4.0f), 4.0f);
                }
            }
          }
        }
        // a skipped term is staged as +0: the running sum starts at +0 and so is never -0,
        // and x + (+0) == x for every other x, so adding it changes no bit
        if (!good) {
#pragma unroll
          for (int c = 0; c < DD; ++c)
            if (c < nd) vb[slot * DD + c] = 0.0f;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (lane < nd) {
        const uint32_t cntw = min(total - w0, (uint32_t)WIDE_CAP);
        uint32_t t = 0;
        for (; t + 8 <= cntw; t += 8) {  // reads ahead of the dependent adds
          float x[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) x[u] = vb[(t + u) * DD + lane];
#pragma unroll
          for (int u = 0; u < 8; ++u) gl = gl + x[u];
        }
        for (; t < cntw; ++t) gl = gl + vb[t * DD + lane];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (active) {
      next_s[s] = ns + ps;
      next_n[s] = nn + (float)cnt * pn;
    }
  }
  if (lane < nd) {
    const int64_t qd = i * nd + lane;
    const float eic = e_old[qd];
    if (touched) {
      float mq = m[qd], vq = v[qd];
      mq = mq + one_m_b1 * (gl - mq);
      vq = vq + one_m_b2 * (gl * gl - vq);
      m[qd] = mq;
      v[qd] = vq;
      const float step = (lr_ad * mq) / (sqrtf(vq) + epsc);
      e_new[qd] = eic + step;
    } else {
      e_new[qd] = eic;
    }
  }
}

// SEG lanes per node (SEG >= dim), 64 / SEG nodes per wavefront, for nodes with fewer than
// wide_deg edges: the wavefront-per-node scheme on SEG-lane segments. A segment walks its
// node's row SEG slots at a time (slot loads coalesced, the neighbour and negative-sample rows
// gathered by all lanes at once), stages the chunk's gradient terms in LDS and its lanes
// c < dim add them in slot order.
template <int D, int SEG>
__global__ __launch_bounds__(256) void epoch_group_kernel(
    int64_t n, int dim, const int64_t* __restrict__ indptr, const int32_t* __restrict__ other,
    float* __restrict__ next_s, const float* __restrict__ eps_s, float* __restrict__ next_n,
    float neg_factor, const float* __restrict__ e_old, float* __restrict__ e_new,
    float* __restrict__ m, float* __restrict__ v, float epoch_t, uint32_t epoch, float lr_ad,
    float epsc, float gamma, float noise, float one_m_b1, float one_m_b2, uint64_t seed,
    int64_t wide_deg) {
  constexpr int DD = D > 0 ? D : MAXD;
  constexpr int NSEG = 64 / SEG;
  constexpr int GCAP = 32 * SEG / DD;  // terms staged per node at a time
  static_assert(SEG >= DD && 64 % SEG == 0, "a segment holds one lane per dimension");
  __shared__ float gbuf[4][NSEG][GCAP * DD];  // [wavefront][segment]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int seg = lane / SEG, sl = lane % SEG;
  const int64_t i = ((int64_t)blockIdx.x * 4 + wave) * NSEG + seg;
  const int nd = D > 0 ? D : dim;
  int64_t s0 = 0, s1 = 0;
  bool own = false;
  if (i < n) {
    s0 = indptr[i];
    s1 = indptr[i + 1];
    own = s1 - s0 < wide_deg;  // the others are done by the wide and hub kernels
    if (!own) s1 = s0;
  }
  float ei[DD];
#pragma unroll
  for (int c = 0; c < DD; ++c) ei[c] = own && c < nd ? e_old[i * nd + c] : 0.0f;
  float gl = 0.0f;  // lane sl < nd: gradient component sl
  bool touched = false;
  uint32_t draw_carry = 0;
  float* vb = gbuf[wave][seg];
  const uint64_t segmask = (SEG == 64 ? ~0ull : ((1ull << SEG) - 1)) << (seg * SEG);
  // every segment runs the wavefront's chunk count; finished ones have no valid slots
  // the next chunk's slot inputs and positive endpoint rows are loaded one chunk ahead
  float ns_a = 0.0f, ps_a = 0.0f, nn_a = 0.0f;
  int32_t j_a = 0;
  float eo_a[DD];
  if (s0 + sl < s1) {
    ns_a = next_s[s0 + sl];
    j_a = other[s0 + sl];
    ps_a = eps_s[s0 + sl];
    nn_a = next_n[s0 + sl];
  }
  {
    const bool due = s0 + sl < s1 && !(ns_a > epoch_t);
#pragma unroll
    for (int c = 0; c < DD; ++c) eo_a[c] = due && c < nd ? e_old[(int64_t)j_a * nd + c] : 0.0f;
  }
  for (int64_t sb = s0; __ballot(sb < s1) != 0; sb += SEG) {
    const int64_t s = sb + sl;
    const bool valid = s < s1;
    const float ns = valid ? ns_a : 0.0f;
    const bool active = valid && !(ns > epoch_t);
    int32_t j = 0;
    float ps = 0.0f, pn = 0.0f, nn = 0.0f;
    uint64_t cnt = 0;
    if (active) {
      j = j_a;
      ps = ps_a;
      pn = ps * neg_factor;
      nn = nn_a;
      const float f = floorf((epoch_t - nn) / pn);
      cnt = (f >= 0.0f && f < 1.8e19f) ? (uint64_t)f : 0;
    }
    float eo[DD];
#pragma unroll
    for (int c = 0; c < DD; ++c) eo[c] = eo_a[c];
    if (s + SEG < s1) {
      ns_a = next_s[s + SEG];
      j_a = other[s + SEG];
      ps_a = eps_s[s + SEG];
      nn_a = next_n[s + SEG];
      const bool due = !(ns_a > epoch_t);
#pragma unroll
      for (int c = 0; c < DD; ++c) eo_a[c] = due && c < nd ? e_old[(int64_t)j_a * nd + c] : 0.0f;
    }
    touched |= (__ballot(active) & segmask) != 0;
    const uint32_t nneg = (active && (int64_t)j > i) ? (uint32_t)cnt : 0u;
    const uint32_t nvec = active ? 1u + nneg : 0u;
    // inclusive prefix sums over the segment's lanes (slot order) of draws and terms
    uint32_t dinc = nneg, vinc = nvec;
#pragma unroll
    for (int o = 1; o < SEG; o <<= 1) {
      const uint32_t td = __shfl_up(dinc, o, SEG), tv = __shfl_up(vinc, o, SEG);
      if (sl >= o) {
        dinc += td;
        vinc += tv;
      }
    }
    const uint32_t draw0 = draw_carry + dinc - nneg;
    const uint32_t voff = vinc - nvec;
    const uint32_t total = __shfl(vinc, SEG - 1, SEG);
    draw_carry += __shfl(dinc, SEG - 1, SEG);
    uint32_t tmax = total;  // windows run to the largest segment total
#pragma unroll
    for (int o = SEG; o < 64; o <<= 1) tmax = max(tmax, (uint32_t)__shfl_xor((int)tmax, o, 64));
    // this lane's positive term
    float coeff = 0.0f;
    bool pos_ok = false;
    if (active) {
      float dsq = 0.0f;
#pragma unroll
      for (int c = 0; c < DD; ++c)
        if (c < nd) {
          const float df = ei[c] - eo[c];
          const float p = df * df;
          dsq = dsq + p;
        }
      if (dsq >= 1e-8f) {
        const float dist = sqrtf(dsq);
        const float num = -((2.0f * noise) * dist) - 2.0f;
        const float den = ((2.0f * dsq) - (0.5f * dist)) + 1.0f;
        coeff = num / den;
        pos_ok = true;
      }
    }
    for (uint32_t w0 = 0; w0 < tmax; w0 += GCAP) {
      // terms of this lane that fall in the window [w0, w0 + GCAP)
      const uint32_t lo = voff > w0 ? voff : w0;
      const uint32_t hi = min(voff + nvec, w0 + (uint32_t)GCAP);
      for (uint32_t t = lo; t < hi; ++t) {
        const uint32_t q = t - voff;
        float* tb = vb + (t - w0) * DD;
        bool good = false;
        if (q == 0) {
          good = pos_ok;
          if (good) {
#pragma unroll
            for (int c = 0; c < DD; ++c)
              if (c < nd) tb[c] = coeff * (ei[c] - eo[c]);
          }
        } else {
          const uint32_t r = Philox::first(seed, draw0 + (q - 1), epoch, (uint32_t)i, TAG_NEG);
          const int64_t kk = Philox::range(r, (uint64_t)n);
          if (kk != i) {
            float ek[DD];
#pragma unroll
            for (int c = 0; c < DD; ++c) ek[c] = c < nd ? e_old[kk * nd + c] : 0.0f;
            float dk = 0.0f;
#pragma unroll
            for (int c = 0; c < DD; ++c)
              if (c < nd) {
                const float df = ei[c] - ek[c];
                const float p = df * df;
                dk = dk + p;
              }
            if (dk > 0.01f) {
              good = true;
              const float gc = (gamma * 4.0f) / ((1.0f + 0.25f * dk) * dk);
#pragma unroll
              for (int c = 0; c < DD; ++c)
                if (c < nd) {
                  const float diff = ei[c] - ek[c];
                  tb[c] = fminf(fmaxf(gc * diff, -4.0f), 4.0f);
                }
            }
          }
        }
        // a skipped term is staged as +0, which leaves the running sum's bits unchanged
        if (!good) {
#pragma unroll
          for (int c = 0; c < DD; ++c)
            if (c < nd) tb[c] = 0.0f;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (sl < nd && total > w0) {
        const uint32_t cntw = min(total - w0, (uint32_t)GCAP);
        for (uint32_t t = 0; t < cntw; ++t) gl = gl + vb[t * DD + sl];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (active) {
      next_s[s] = ns + ps;
      next_n[s] = nn + (float)cnt * pn;
    }
  }
  if (own && sl < nd) {
    const int64_t qd = i * nd + sl;
    const float eic = e_old[qd];
    if (touched) {
      float mq = m[qd], vq = v[qd];
      mq = mq + one_m_b1 * (gl - mq);
      vq = vq + one_m_b2 * (gl * gl - vq);
      m[qd] = mq;
      v[qd] = vq;
      const float step = (lr_ad * mq) / (sqrtf(vq) + epsc);
      e_new[qd] = eic + step;
    } else {
      e_new[qd] = eic;
    }
  }
}

// One workgroup per node with at least HUB_DEG edges (a coarse graph's largest partition
// can touch every other one, thousands of 64-slot chunks in one row). The row is walked in groups of HUB_WAVES chunks: wavefront w
// takes the w-th chunk of a group, a scan over the wavefronts' totals gives each slot its
// draw numbers and the position of its terms, the terms of the group are staged in LDS
// (HUB_CAP at a time) and wavefront 0 adds them in slot order.
template <int D>
__global__ __launch_bounds__(1024) void epoch_hub_kernel(
    int64_t n, int dim, const int64_t* __restrict__ indptr, const int32_t* __restrict__ other,
    float* __restrict__ next_s, const float* __restrict__ eps_s, float* __restrict__ next_n,
    float neg_factor, const float* __restrict__ e_old, float* __restrict__ e_new,
    float* __restrict__ m, float* __restrict__ v, float epoch_t, uint32_t epoch, float lr_ad,
    float epsc, float gamma, float noise, float one_m_b1, float one_m_b2, uint64_t seed,
    const int32_t* __restrict__ hub_nodes) {
  constexpr int DD = D > 0 ? D : MAXD;
  __shared__ float tbuf[HUB_CAP * DD];
  __shared__ uint32_t wdraw[HUB_WAVES], wterm[HUB_WAVES];
  __shared__ int wtouch[HUB_WAVES];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t i = hub_nodes[blockIdx.x];  // the grid has one workgroup per listed node
  const int nd = D > 0 ? D : dim;
  float ei[DD];
#pragma unroll
  for (int c = 0; c < DD; ++c) ei[c] = c < nd ? e_old[i * nd + c] : 0.0f;
  float gl = 0.0f;  // wavefront 0, lane c < nd: gradient component c
  bool touched = false;
  uint32_t draw_carry = 0;
  const int64_t s0 = indptr[i], s1 = indptr[i + 1];
  for (int64_t gb = s0; gb < s1; gb += 64 * HUB_WAVES) {
    const int64_t s = gb + (int64_t)wave * 64 + lane;
    const bool valid = s < s1;
    const float ns = valid ? next_s[s] : 0.0f;
    const bool active = valid && !(ns > epoch_t);
    int32_t j = 0;
    float ps = 0.0f, pn = 0.0f, nn = 0.0f;
    uint64_t cnt = 0;
    if (active) {
      j = other[s];
      ps = eps_s[s];
      pn = ps * neg_factor;
      nn = next_n[s];
      const float f = floorf((epoch_t - nn) / pn);
      cnt = (f >= 0.0f && f < 1.8e19f) ? (uint64_t)f : 0;
    }
    touched |= __ballot(active) != 0;
    const uint32_t nneg = (active && (int64_t)j > i) ? (uint32_t)cnt : 0u;
    const uint32_t nvec = active ? 1u + nneg : 0u;
    // inclusive prefix sums over the lanes (slot order) of draws and of staged terms
    uint32_t dinc = nneg, vinc = nvec;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t td = __shfl_up(dinc, o, 64), tv = __shfl_up(vinc, o, 64);
      if (lane >= o) {
        dinc += td;
        vinc += tv;
      }
    }
    if (lane == 63) {
      wdraw[wave] = dinc;
      wterm[wave] = vinc;
    }
    __syncthreads();
    uint32_t dbase = draw_carry, vbase = 0, dtot = 0, vtot = 0;
    for (int w = 0; w < HUB_WAVES; ++w) {  // earlier chunks of the group, and the group totals
      const uint32_t a = wdraw[w], b = wterm[w];
      if (w < wave) {
        dbase += a;
        vbase += b;
      }
      dtot += a;
      vtot += b;
    }
    __syncthreads();  // every wavefront has read the totals before the next group's are written
    const uint32_t draw0 = dbase + dinc - nneg;
    const uint32_t voff = vbase + vinc - nvec;
    draw_carry += dtot;
    // this lane's positive term
    float eo[DD];
    float coeff = 0.0f;
    bool pos_ok = false;
    if (active) {
#pragma unroll
      for (int c = 0; c < DD; ++c) eo[c] = c < nd ? e_old[(int64_t)j * nd + c] : 0.0f;
      float dsq = 0.0f;
#pragma unroll
      for (int c = 0; c < DD; ++c)
        if (c < nd) {
          const float df = ei[c] - eo[c];
          const float p = df * df;
          dsq = dsq + p;
        }
      if (dsq >= 1e-8f) {
        const float dist = sqrtf(dsq);
        const float num = -((2.0f * noise) * dist) - 2.0f;
        const float den = ((2.0f * dsq) - (0.5f * dist)) + 1.0f;
        coeff = num / den;
        pos_ok = true;
      }
    }
    for (uint32_t w0 = 0; w0 < vtot; w0 += HUB_CAP) {
      // terms of this lane that fall in the window [w0, w0 + HUB_CAP)
      const uint32_t lo = voff > w0 ? voff : w0;
      const uint32_t hi = min(voff + nvec, w0 + (uint32_t)HUB_CAP);
      for (uint32_t t = lo; t < hi; ++t) {
        const uint32_t q = t - voff;
        float* tb = tbuf + (t - w0) * DD;
        bool good = false;
        if (q == 0) {
          good = pos_ok;
          if (good) {
#pragma unroll
            for (int c = 0; c < DD; ++c)
              if (c < nd) tb[c] = coeff * (ei[c] - eo[c]);
          }
        } else {
          const uint32_t r = Philox::first(seed, draw0 + (q - 1), epoch, (uint32_t)i, TAG_NEG);
          const int64_t kk = Philox::range(r, (uint64_t)n);
          if (kk != i) {
            float ek[DD];
#pragma unroll
            for (int c = 0; c < DD; ++c) ek[c] = c < nd ? e_old[kk * nd + c] : 0.0f;
            float dk = 0.0f;
#pragma unroll
            for (int c = 0; c < DD; ++c)
              if (c < nd) {
                const float df = ei[c] - ek[c];
                const float p = df * df;
                dk = dk + p;
              }
            if (dk > 0.01f) {
              good = true;
              const float gc = (gamma * 4.0f) / ((1.0f + 0.25f * dk) * dk);
#pragma unroll
              for (int c = 0; c < DD; ++c)
                if (c < nd) {
                  const float diff = ei[c] - ek[c];
                  tb[c] = fminf(fmaxf(gc * diff, -4.0f), 4.0f);
                }
            }
          }
        }
        // a skipped term is staged as +0, which leaves the running sum's bits unchanged
        if (!good) {
#pragma unroll
          for (int c = 0; c < DD; ++c)
            if (c < nd) tb[c] = 0.0f;
        }
      }
      __syncthreads();
      if (wave == 0 && lane < nd) {
        const uint32_t cntw = min(vtot - w0, (uint32_t)HUB_CAP);
        uint32_t t = 0;
        for (; t + 8 <= cntw; t += 8) {  // reads ahead of the dependent adds
          float x[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) x[u] = tbuf[(t + u) * DD + lane];
#pragma unroll
          for (int u = 0; u < 8; ++u) gl = gl + x[u];
        }
        for (; t < cntw; ++t) gl = gl + tbuf[t * DD + lane];
      }
      __syncthreads();
    }
    if (active) {
      next_s[s] = ns + ps;
      next_n[s] = nn + (float)cnt * pn;
    }
  }
  if (lane == 0) wtouch[wave] = touched ? 1 : 0;
  __syncthreads();
  if (wave == 0 && lane < nd) {
    bool any = false;
    for (int w = 0; w < HUB_WAVES; ++w) any |= wtouch[w] != 0;
    const int64_t qd = i * nd + lane;
    const float eic = e_old[qd];
    if (any) {
      float mq = m[qd], vq = v[qd];
      mq = mq + one_m_b1 * (gl - mq);
      vq = vq + one_m_b2 * (gl * gl - vq);
      m[qd] = mq;
      v[qd] = vq;
      const float step = (lr_ad * mq) / (sqrtf(vq) + epsc);
      e_new[qd] = eic + step;
    } else {
      e_new[qd] = eic;
    }
  }
}

// nodes with at least wide_deg edges, in any order: those below hub_deg into list, the
// others into hubs
__global__ void wide_nodes_kernel(const int64_t* __restrict__ indptr, int64_t n, int64_t wide_deg,
                                  int64_t hub_deg, int32_t* __restrict__ list,
                                  int32_t* __restrict__ hubs, int* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t deg = indptr[i + 1] - indptr[i];
  if (deg >= hub_deg) hubs[atomicAdd(count + 1, 1)] = (int32_t)i;
  else if (deg >= wide_deg) list[atomicAdd(count, 1)] = (int32_t)i;
}

template <int D>
void launch_epoch(hipStream_t s, int64_t n, int dim, const int64_t* indptr,
                  const int32_t* other, float* ns, const float* es, float* nn, float nf,
                  const float* eo, float* en, float* m, float* v, float et, uint32_t ep, float la,
                  float epsc, float gamma, float noise, float b1, float b2, uint64_t seed,
                  const int32_t* wide, int64_t nwide, int64_t wide_deg, const int32_t* hubs,
                  int64_t nhub) {
  constexpr int SEG = (D > 0 ? D : MAXD) > GSEG ? (D > 0 ? D : MAXD) : GSEG;
  epoch_group_kernel<D, SEG><<<(unsigned)ceil_div(n, 4 * (64 / SEG)), 256, 0, s>>>(
      n, dim, indptr, other, ns, es, nn, nf, eo, en, m, v, et, ep, la, epsc, gamma, noise, b1, b2,
      seed, nwide + nhub > 0 ? wide_deg : INT64_MAX);
  if (nwide > 0)
    epoch_wide_kernel<D><<<(unsigned)ceil_div(nwide, 4), 256, 0, s>>>(
        n, dim, indptr, other, ns, es, nn, nf, eo, en, m, v, et, ep, la, epsc, gamma, noise, b1,
        b2, seed, wide, nwide);
  if (nhub > 0)
    epoch_hub_kernel<D><<<(unsigned)nhub, 64 * HUB_WAVES, 0, s>>>(
        n, dim, indptr, other, ns, es, nn, nf, eo, en, m, v, et, ep, la, epsc, gamma, noise, b1,
        b2, seed, hubs);
}

}  // namespace

namespace {

__global__ void zero_words_kernel(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    p[q] = 0u;
}
void zero_words(void* p, int64_t words, hipStream_t s) {
  if (words <= 0) return;
  zero_words_kernel<<<(unsigned)std::min<int64_t>(ceil_div(words, 256), 4096), 256, 0, s>>>(
      static_cast<uint32_t*>(p), words);
}

// Device memory the epochs use besides the graph and the layout: the per-slot schedule,
// the Adam moments, the second layout buffer of the Jacobi update and the high-degree
// node lists (nnz, n * dim and n entries).
struct EmbedScratch {
  unsigned* maxw;  // 1
  float *eps_s, *next_s, *next_n;  // nnz each
  float *m, *v, *e2;  // n * dim each
  int32_t *wide, *hubs;  // n each
  int* nwide;  // 2
};

// The epochs of embed_device on a CSR of embedding slots (indptr/indices/weights),
// scratch given by the caller. No device allocation; kernel_ms (if not null) adds two
// events.
void embed_core(const int64_t* indptr, const int32_t* indices, const float* weights, int64_t n,
                int64_t nnz, int dim, const EmbedParams& p, float* d_emb, bool random_init,
                uint64_t seed, hipStream_t s, const EmbedScratch& w, float* kernel_ms) {
  if (n == 0) return;
  const unsigned gn = (unsigned)ceil_div(n * dim, 256);
  if (random_init) random_init_kernel<<<gn, 256, 0, s>>>(d_emb, n, dim, seed);
  if (nnz == 0 || p.n_epochs <= 0) return;
  zero_words(w.maxw, 1, s);
  zero_words(w.m, n * dim, s);
  zero_words(w.v, n * dim, s);
  zero_words(w.nwide, 2, s);
  max_weight_kernel<<<(unsigned)std::min<int64_t>(ceil_div(nnz, 256), 4096), 256, 0, s>>>(
      weights, nnz, w.maxw);
  const float neg_factor = 1.5f / p.negative_sample_rate;
  schedule_init_kernel<<<(unsigned)ceil_div(nnz, 256), 256, 0, s>>>(
      weights, nnz, w.maxw, neg_factor, w.eps_s, w.next_s, w.next_n);
  unsigned hmax = 0;
  EVOC_HIP(hipMemcpyAsync(&hmax, w.maxw, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  if (hmax == 0) return;  // every weight is zero: epochs_per_sample = 1/(0/0) -> reference NaN path skipped
  const float one_m_b1 = 1.0f - p.beta1, one_m_b2 = 1.0f - p.beta2;
  const float denom = (float)std::max(p.n_epochs - 1, 1);
  const float n_epochs_f = (float)p.n_epochs;
  const unsigned grid = (unsigned)ceil_div(n, 256);
  // high-degree nodes for the wavefront-per-node kernel
  EVOC_CHECK(p.wide_deg >= 1, "embedding: wide_deg must be >= 1");
  const int64_t hub_deg = std::max<int64_t>(p.wide_deg, HUB_DEG);
  wide_nodes_kernel<<<grid, 256, 0, s>>>(indptr, n, p.wide_deg, hub_deg, w.wide, w.hubs,
                                         w.nwide);
  int nwide_h[2] = {0, 0};
  EVOC_HIP(hipMemcpyAsync(nwide_h, w.nwide, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  const int64_t nwide = nwide_h[0], nhub = nwide_h[1];
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (kernel_ms) {
    EVOC_HIP(hipEventCreate(&e0));
    EVOC_HIP(hipEventCreate(&e1));
    EVOC_HIP(hipEventRecord(e0, s));
  }
  float* cur = d_emb;
  float* nxt = w.e2;
  for (int epoch = 0; epoch < p.n_epochs; ++epoch) {
    const float lr = p.initial_alpha * (1.0f - (float)epoch / n_epochs_f);
    const float epoch_t = (float)epoch;
    const float t = (float)(epoch + 1);
    const float b1t = std::pow(p.beta1, t), b2t = std::pow(p.beta2, t);
    const float sqrt_b2 = std::sqrt(1.0f - b2t);
    const float ad_scale = sqrt_b2 / (1.0f - b1t);
    const float epsc = sqrt_b2 * p.eps;
    const float gamma = 0.5f + (float)epoch / denom;
    const float lr_ad = lr * ad_scale;
#define EVOC_EPOCH(D)                                                                         \
  launch_epoch<D>(s, n, dim, indptr, indices, w.next_s, w.eps_s, w.next_n, neg_factor, cur, nxt, \
                  w.m, w.v, epoch_t, (uint32_t)epoch, lr_ad, epsc, gamma, p.noise_level,      \
                  one_m_b1, one_m_b2, seed, w.wide, nwide, p.wide_deg, w.hubs, nhub)
    switch (dim) {
      case 2: EVOC_EPOCH(2); break;
      case 4: EVOC_EPOCH(4); break;
      case 8: EVOC_EPOCH(8); break;
      case 16: EVOC_EPOCH(16); break;
      default: EVOC_EPOCH(0); break;
    }
#undef EVOC_EPOCH
    std::swap(cur, nxt);
  }
  EVOC_HIP(hipGetLastError());
  if (kernel_ms) EVOC_HIP(hipEventRecord(e1, s));
  if (cur != d_emb)
    EVOC_HIP(hipMemcpyAsync(d_emb, cur, n * dim * sizeof(float), hipMemcpyDeviceToDevice, s));
  EVOC_HIP(hipStreamSynchronize(s));
  if (kernel_ms) {
    float ms = 0.0f;
    EVOC_HIP(hipEventElapsedTime(&ms, e0, e1));
    *kernel_ms = ms;
    EVOC_HIP(hipEventDestroy(e0));
    EVOC_HIP(hipEventDestroy(e1));
  }
}

}  // namespace

void embed_device(const Graph& g, int dim, const EmbedParams& p, float* d_emb, bool random_init,
                  uint64_t seed, hipStream_t s, DevPool& pool, float* kernel_ms) {
  const int64_t n = g.n;
  EVOC_CHECK(dim >= 1 && dim <= MAXD, "embedding: dim must be in [1, 16]");
  if (n == 0) return;
  const int64_t nnz = g.nnz;
  const bool epochs = nnz > 0 && p.n_epochs > 0;
  Buf<unsigned> maxw(pool, 1);
  Buf<float> eps_s(pool, epochs ? nnz : 1), next_s(pool, epochs ? nnz : 1),
      next_n(pool, epochs ? nnz : 1);
  Buf<float> m(pool, epochs ? n * dim : 1), v(pool, epochs ? n * dim : 1),
      e2(pool, epochs ? n * dim : 1);
  Buf<int32_t> wide(pool, epochs ? n : 1), hubs(pool, epochs ? n : 1);
  Buf<int> nwide(pool, 2);
  EmbedScratch w{maxw.p, eps_s.p, next_s.p, next_n.p, m.p, v.p, e2.p, wide.p, hubs.p, nwide.p};
  embed_core(g.indptr.p, g.indices.p, g.weights.p, n, nnz, dim, p, d_emb, random_init, seed, s, w,
             kernel_ms);
}

}  // namespace evoc

// ------------------------------------------------------------------ allocation-free variant
// The epochs' scratch in one static device array per GPU, for callers that cannot allocate
// (code under HIP graph capture, serving loops). Limits: n <= EMBED_NOALLOC_MAX_N nodes,
// nnz <= EMBED_NOALLOC_MAX_NNZ slots, dim <= 16: 36.5 MB, held by each GPU the library
// runs on until the process ends.
namespace evoc {
namespace {
constexpr int64_t EMBED_NOALLOC_MAX_N = 65536;
constexpr int64_t EMBED_NOALLOC_MAX_NNZ = 2097152;
constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t SC_HEAD = 512;                                     // maxw, nwide
constexpr size_t SC_SCHED = al256(EMBED_NOALLOC_MAX_NNZ * 4);       // eps_s, next_s, next_n
constexpr size_t SC_STATE = al256(EMBED_NOALLOC_MAX_N * MAXD * 4);  // m, v, e2
constexpr size_t SC_NODES = al256(EMBED_NOALLOC_MAX_N * 4);         // wide, hubs
constexpr size_t SCRATCH_BYTES = SC_HEAD + 3 * SC_SCHED + 3 * SC_STATE + 2 * SC_NODES;
constexpr int MAX_DEVICES = 64;
}  // namespace
}  // namespace evoc

static __device__ __align__(256) unsigned char scratch[evoc::SCRATCH_BYTES];

namespace evoc {
namespace {
// device address of `scratch` on each GPU, looked up once per device with it current
unsigned char* g_scratch[MAX_DEVICES] = {};

unsigned char* scratch_on(int device) {
  EVOC_CHECK(device >= 0 && device < MAX_DEVICES, "device_id out of range");
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    EVOC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch[device];
}
}  // namespace
}  // namespace evoc

extern "C" int evoc_hip_embed_noalloc(const int64_t* d_indptr, const int32_t* d_indices,
                                      const float* d_weights, int64_t n, int64_t nnz, int32_t dim,
                                      int32_t n_epochs, float noise_level,
                                      float negative_sample_rate, float initial_alpha,
                                      float beta1, float beta2, float eps, int32_t random_init,
                                      uint64_t seed, int32_t device_id, float* d_emb) {
  using namespace evoc;
  return guarded([&] {
    EVOC_CHECK(d_indptr && d_emb && (nnz == 0 || (d_indices && d_weights)),
               "embed_noalloc: NULL buffer");
    EVOC_CHECK(n >= 0 && nnz >= 0 && dim >= 1 && dim <= MAXD,
               "embed_noalloc: need n >= 0, nnz >= 0 and dim in [1, 16]");
    EVOC_CHECK(n <= EMBED_NOALLOC_MAX_N, "embed_noalloc: n exceeds 65536 (the static scratch)");
    EVOC_CHECK(nnz <= EMBED_NOALLOC_MAX_NNZ,
               "embed_noalloc: nnz exceeds 2097152 (the static scratch)");
    EVOC_CHECK(n_epochs >= 0, "embed_noalloc: n_epochs must be >= 0");
    Ctx& c = context(device_id);
    unsigned char* base = scratch_on(device_id);
    EmbedScratch w;
    size_t off = 0;
    auto take = [&](size_t bytes) {
      unsigned char* q = base + off;
      off += al256(bytes);
      return q;
    };
    w.maxw = reinterpret_cast<unsigned*>(take(256));
    w.nwide = reinterpret_cast<int*>(take(256));
    w.eps_s = reinterpret_cast<float*>(take(SC_SCHED));
    w.next_s = reinterpret_cast<float*>(take(SC_SCHED));
    w.next_n = reinterpret_cast<float*>(take(SC_SCHED));
    w.m = reinterpret_cast<float*>(take(SC_STATE));
    w.v = reinterpret_cast<float*>(take(SC_STATE));
    w.e2 = reinterpret_cast<float*>(take(SC_STATE));
    w.wide = reinterpret_cast<int32_t*>(take(SC_NODES));
    w.hubs = reinterpret_cast<int32_t*>(take(SC_NODES));
    EVOC_CHECK(off <= SCRATCH_BYTES, "embed_noalloc: scratch layout");
    EmbedParams ep;
    ep.n_epochs = n_epochs;
    ep.noise_level = noise_level;
    ep.negative_sample_rate = negative_sample_rate;
    ep.initial_alpha = initial_alpha;
    ep.beta1 = beta1;
    ep.beta2 = beta2;
    ep.eps = eps;
    embed_core(d_indptr, d_indices, d_weights, n, nnz, dim, ep, d_emb, random_init != 0, seed,
               c.stream, w, nullptr);
  });
}
