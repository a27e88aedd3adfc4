// Shared device helpers of libspecann: error reporting, counter-based hashing, wavefront and
// workgroup reductions. Written for gfx950 (CDNA4): 64-wide wavefronts, 256-thread workgroups.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/specann.h"

#define SA_BLOCK 256
#define SA_WAVES_PER_BLOCK (SA_BLOCK / 64)

namespace sa {

// ---------------------------------------------------------------------------------------------
// Error reporting: every entry point returns a status and leaves a message for sa_last_error.
// ---------------------------------------------------------------------------------------------
void set_error(const std::string& msg);
int check_launch(const char* what);

#define SA_REQUIRE(cond, msg)                                 \
  do {                                                        \
    if (!(cond)) {                                            \
      ::sa::set_error(std::string(__func__) + ": " + (msg));  \
      return SA_E_ARG;                                        \
    }                                                         \
  } while (0)

// ---------------------------------------------------------------------------------------------
// Counter-based hashing (SplitMix64 finaliser). The host oracle restates these bit for bit.
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t mix3(uint64_t a, uint64_t b, uint64_t c) {
  return splitmix64(splitmix64(splitmix64(a) ^ b) ^ c);
}

// 2^-53 * (53 random bits): uniform double in [0, 1).
__host__ __device__ inline double u53(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// ---------------------------------------------------------------------------------------------
// Reductions. Wave reductions use cross-lane shuffles; the workgroup reduction combines the four
// wave results in a fixed order, so every sum here is deterministic for a fixed launch shape.
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  return v;
}

// Sum over a 256-thread workgroup; the result is valid in thread 0. `red` holds 4 values.
template <typename T>
__device__ inline T block_sum(T v, T* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) r = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return r;
}

// 16-byte vector of T for the basis-streaming kernels (4 floats or 2 doubles per lane).
template <typename T>
struct vec16;
template <>
struct vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int L = 4;
};
template <>
struct vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int L = 2;
};

template <typename T>
__device__ inline T stream_load(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Rows [lo, hi) owned by this workgroup in the basis-streaming kernels: one contiguous range per
// workgroup (a multiple of 64 rows), so that the waves of a workgroup walk through the same pages of
// every basis vector instead of striding the whole vector.
__device__ inline void block_rows(int64_t n, int64_t& lo, int64_t& hi) {
  const int64_t per = (((n + gridDim.x - 1) / gridDim.x) + 63) & ~(int64_t)63;
  lo = (int64_t)blockIdx.x * per;
  hi = lo + per < n ? lo + per : n;
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

inline int grid_for(int64_t work, int per_block, int cap) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

}  // namespace sa
