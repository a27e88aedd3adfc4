// Deterministic grid-wide reductions over the caller's work buffer.
//
// A reduction kernel runs at most SA_RED_GRID workgroups; each workgroup reduces its share to one
// value per column, stores it in the partials area, and the last workgroup to arrive sums the
// partials in block order. No floating-point atomics are used, so results are bitwise
// reproducible for a fixed (n, grid).
#pragma once

#include "sa_common.hpp"

namespace sa {

inline size_t state_bytes_padded() { return (sizeof(sa_state) + 255) & ~(size_t)255; }

__device__ inline sa_state* state_of(void* work) { return reinterpret_cast<sa_state*>(work); }

__device__ inline double* partials_of(void* work) {
  return reinterpret_cast<double*>(reinterpret_cast<char*>(work) + ((sizeof(sa_state) + 255) & ~(size_t)255));
}

// Number of doubles in the partials area.
constexpr int64_t kPartials = (int64_t)2048 * (SA_MAX_NCV + 8);

// Thick-restart scratch after the partials: the kept Ritz coefficient vectors (SA_MAX_KEEP x
// SA_MAX_NCV) and per-thread inverse-iteration scratch (7 vectors and a pivot array each).
constexpr int64_t kRestart = (int64_t)SA_MAX_KEEP * SA_MAX_NCV * 9;

__device__ inline double* restart_of(void* work) { return partials_of(work) + kPartials; }

enum RedOp { RED_SUM = 0, RED_MAX = 1 };

// Reduce C per-thread values over the whole grid. Every thread of every block must call this.
// Results land in out[0..C) (device memory, written by the last block only). Returns true in the
// last block (after the results are written), so callers may run a serial epilogue there.
template <int C, int OP>
__device__ bool grid_reduce(double (&v)[C], void* work, double* out) {
  __shared__ double red[SA_WAVES_PER_BLOCK];
  __shared__ bool last;
  sa_state* st = state_of(work);
  double* part = partials_of(work);
#pragma unroll
  for (int c = 0; c < C; ++c) {
    double x = v[c];
    if (OP == RED_SUM) {
      x = block_sum(x, red);
    } else {
      x = wave_max(x);
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      if (lane == 0) red[wave] = x;
      __syncthreads();
      if (threadIdx.x == 0) x = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * C + c] = x;
  }
  if (!last_block_arrive(&st->counters[0], gridDim.x, &last)) return false;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    double x = (OP == RED_SUM) ? 0.0 : -INFINITY;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
      double p = load_coherent(&part[(int64_t)b * C + c]);
      x = (OP == RED_SUM) ? x + p : fmax(x, p);
    }
    if (OP == RED_SUM) {
      x = block_sum(x, red);
    } else {
      x = wave_max(x);
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      if (lane == 0) red[wave] = x;
      __syncthreads();
      if (threadIdx.x == 0) x = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = x;
  }
  __threadfence();
  __syncthreads();
  return true;
}

// Same for wrapping uint64 sums (signatures) and exact int64 counts.
__device__ inline bool grid_reduce_u64(uint64_t v, void* work, uint64_t* out) {
  __shared__ uint64_t red[SA_WAVES_PER_BLOCK];
  __shared__ bool last;
  sa_state* st = state_of(work);
  uint64_t* part = reinterpret_cast<uint64_t*>(partials_of(work));
  uint64_t x = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = x;
  if (!last_block_arrive(&st->counters[0], gridDim.x, &last)) return false;
  uint64_t acc = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) acc += load_coherent(&part[b]);
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) *out = acc;
  __threadfence();
  __syncthreads();
  return true;
}

}  // namespace sa
