// Shared device/host helpers for the evoc_hip library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

namespace evoc {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

#define EVOC_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::evoc::HipError(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                             __FILE__ + ":" + std::to_string(__LINE__) + " (" #call ")"); \
  } while (0)

#define EVOC_CHECK(cond, msg)                                   \
  do {                                                          \
    if (!(cond)) throw std::invalid_argument(std::string(msg)); \
  } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------------------ random numbers
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based, so every
// (seed, stream, indices) tuple gets its own number with no state to carry.
struct Philox {
  static __host__ __device__ inline void round10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
      uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
      uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
      uint32_t n1 = (uint32_t)p1;
      uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
      uint32_t n3 = (uint32_t)p0;
      c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
  }
  static __host__ __device__ inline uint32_t first(uint64_t seed, uint32_t c0, uint32_t c1,
                                                   uint32_t c2, uint32_t c3) {
    uint32_t c[4] = {c0, c1, c2, c3};
    round10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return c[0];
  }
  static __host__ __device__ inline double unit(uint64_t seed, uint32_t c0, uint32_t c1,
                                                uint32_t c2, uint32_t c3) {
    uint32_t c[4] = {c0, c1, c2, c3};
    round10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return (double)(((uint64_t)(c[0] >> 5) << 26) | (uint64_t)(c[1] >> 6)) *
           (1.0 / 9007199254740992.0);
  }
  // uniform integer in [0, n) by multiply-high
  static __host__ __device__ inline uint32_t range(uint32_t r, uint64_t n) {
    return (uint32_t)(((uint64_t)r * n) >> 32);
  }
};

// stream tags, shared with every component that draws random numbers
constexpr uint32_t TAG_NEG = 0x454E4547u;   // embedding negative samples
constexpr uint32_t TAG_EINI = 0x45494E49u;  // embedding random init
constexpr uint32_t TAG_LPSD = 0x4C505344u;  // label-prop seeds
constexpr uint32_t TAG_LPOL = 0x4C504F4Cu;  // label-prop outliers
constexpr uint32_t TAG_RINI = 0x52494E49u;  // label-prop random init
constexpr uint32_t TAG_MLPW = 0x4D4C5057u;  // MLP weight init
constexpr uint32_t TAG_MLPS = 0x4D4C5053u;  // MLP subsample / shuffles

}  // namespace evoc
