// Deterministic grid-wide reductions over the caller's work buffer.
//
// A reduction kernel runs at most SA_RED_GRID workgroups; each workgroup reduces its share to one
// value per column and stores it in the partials area, and a one-workgroup finisher launched next
// sums the partials in block order. No floating-point atomics are used, so results are bitwise
// reproducible for a fixed (n, grid).
#pragma once

#include "sa_map.hpp"

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

// Grid-wide reductions run in two launches. Phase 1 (any kernel): each workgroup reduces its C
// values and thread 0 stores them to part[b * C + c] with plain stores. Phase 2 (a one-workgroup
// finisher launched next on the same stream) reads the partials, which the kernel boundary has made
// visible, and sums them in workgroup order. No workgroup ever waits on another, and no device-scope
// fence is issued per workgroup (on gfx950 such a fence writes back the XCD's L2).
template <int C, int OP>
__device__ void block_partials(double (&v)[C], void* work, int slot) {
  __shared__ double red[SA_WAVES_PER_BLOCK];
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
    if (threadIdx.x == 0) part[(int64_t)slot * C + c] = x;
  }
}

template <int C, int OP>
__device__ void block_partials(double (&v)[C], void* work) {
  block_partials<C, OP>(v, work, (int)blockIdx.x);
}

__device__ inline void block_partials_u64(uint64_t v, void* work, int slot) {
  __shared__ uint64_t red[SA_WAVES_PER_BLOCK];
  uint64_t* part = reinterpret_cast<uint64_t*>(partials_of(work));
  const uint64_t x = block_sum(v, red);
  if (threadIdx.x == 0) part[slot] = x;
}

__device__ inline void block_partials_u64(uint64_t v, void* work) { block_partials_u64(v, work, (int)blockIdx.x); }

// This thread's share of a phase-2 sum: part[b * stride + off] for b = threadIdx.x, + blockDim.x, ...
// below G, added in that order. Eight loads are in flight at a time (the partials come from every
// XCD, so each load is a miss); the order of the additions is the same as one at a time.
template <typename T>
__device__ inline T thread_partial_sum(const T* part, int G, int stride, int off) {
  T x = 0;
  const int B = blockDim.x;
  int b = threadIdx.x;
  for (; b + 7 * B < G; b += 8 * B) {
    T q[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) q[u] = part[(int64_t)(b + u * B) * stride + off];
#pragma unroll
    for (int u = 0; u < 8; ++u) x += q[u];
  }
  for (; b < G; b += B) x += part[(int64_t)b * stride + off];
  return x;
}

// Phase 2 helper for one workgroup: sum of part[b * stride + off] over b < G (each thread in
// workgroup order, then a fixed tree); the result is returned to every thread.
__device__ inline double sum_partials(const double* part, int G, int stride, int off) {
  __shared__ double red[SA_WAVES_PER_BLOCK];
  __shared__ double bc;
  double x = thread_partial_sum(part, G, stride, off);
  x = block_sum(x, red);
  if (threadIdx.x == 0) bc = x;
  __syncthreads();
  x = bc;
  __syncthreads();
  return x;
}

// Generic phase 2: up to four columns, each summed (FP64 or wrapping uint64) or maximised, optionally
// square-rooted, and stored at a byte offset of sa_state. Skipped unless only_if < 0 or the state's
// status equals only_if (the phase-1 kernel is gated the same way).
enum FinishOp { FIN_SUM = 0, FIN_MAX = 1, FIN_SUM_U64 = 2 };
struct FinishSpec {
  int32_t cols;
  int32_t only_if;
  int32_t op[4];
  int32_t take_sqrt[4];
  int32_t dst[4];
};

int launch_finish(void* work, int G, const FinishSpec& spec, hipStream_t stream);
int launch_finish_batch(const sa_batch* Bt, const FinishSpec& spec, hipStream_t stream);

// The batch descriptor's validity and its device-side maps (sa_lanczos.hip).
bool batch_ok(const sa_batch* Bt);
BatchMap bmap(const sa_batch* Bt);
BatchFin bfin(const sa_batch* Bt);

}  // namespace sa
