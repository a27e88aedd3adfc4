// Work distribution of the vector kernels: which rows, work buffer and reduction slot a workgroup
// owns. The same kernel bodies serve one problem (OneMap: the grid splits rows [0, n) of one work
// buffer) and a batch of instances packed block-diagonally (BatchMap: instance b owns rows
// [row_off[b], row_off[b+1]) and workgroups [blk_off[b], blk_off[b+1]), and its own work buffer).
// The one-workgroup finishers take the matching Fin type: OneFin runs one workgroup on one work
// buffer; BatchFin runs workgroup b on instance b's. A batch launch may carry an instance mask: the
// workgroups of masked-out instances return at once, leaving their rows and state untouched.
#pragma once

#include "sa_common.hpp"

namespace sa {

struct Slice {
  void* work;
  int64_t lo, hi;  // rows [lo, hi), global (packed) indices, lo a multiple of 64
  int blk;         // this workgroup's reduction slot among its instance's G workgroups
  int G;
  int inst;        // instance index (0 for one problem)
  int64_t base;    // first row of the instance
  bool live;       // false: the launch leaves this instance alone (a masked batch)
};

struct OneMap {
  void* work;
  int64_t n;
  // In one problem the SpMV applies sigma_j itself; a batch's SpMV cannot (sigma_j differs per
  // instance), so batched projections and updates scale W by their instance's sigma_j.
  static constexpr bool kScaleW = false;
  __device__ inline Slice slice() const {
    Slice s;
    s.work = work;
    block_rows(n, s.lo, s.hi);
    s.blk = blockIdx.x;
    s.G = gridDim.x;
    s.inst = 0;
    s.base = 0;
    s.live = true;
    return s;
  }
};

struct BatchMap {
  char* works;
  int64_t stride;
  const int64_t* row_off;  // B + 1, multiples of 64
  const int64_t* row_end;  // B: end of the instance's real rows (padding rows up to row_off[b+1] are skipped)
  const int32_t* blk_off;  // B + 1
  int32_t B;
  int32_t rows_per_block;  // a multiple of 64
  const uint8_t* mask;     // NULL, or B bytes: only instances with a nonzero byte are live
  static constexpr bool kScaleW = true;
  __device__ inline Slice slice() const {
    int lo = 0, hi = B - 1;  // the instance b with blk_off[b] <= blockIdx.x < blk_off[b + 1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (blk_off[mid] <= (int)blockIdx.x) lo = mid; else hi = mid - 1;
    }
    Slice s;
    s.work = works + (int64_t)lo * stride;
    s.blk = (int)blockIdx.x - blk_off[lo];
    s.G = blk_off[lo + 1] - blk_off[lo];
    s.lo = row_off[lo] + (int64_t)s.blk * rows_per_block;
    const int64_t end = row_end[lo];
    s.hi = s.lo + rows_per_block < end ? s.lo + rows_per_block : end;
    s.inst = lo;
    s.base = row_off[lo];
    s.live = !mask || mask[lo] != 0;
    return s;
  }
};

struct OneFin {
  void* work;
  int G;
  static constexpr bool kScaleW = false;
  __device__ inline void* w() const { return work; }
  __device__ inline int g() const { return G; }
  __device__ inline int idx() const { return 0; }
  __device__ inline bool live() const { return true; }
};

struct BatchFin {
  char* works;
  int64_t stride;
  const int32_t* blk_off;
  const uint8_t* mask;  // as BatchMap::mask
  static constexpr bool kScaleW = true;
  __device__ inline void* w() const { return works + (int64_t)blockIdx.x * stride; }
  __device__ inline int g() const { return blk_off[blockIdx.x + 1] - blk_off[blockIdx.x]; }
  __device__ inline int idx() const { return blockIdx.x; }
  __device__ inline bool live() const { return !mask || mask[blockIdx.x] != 0; }
};

}  // namespace sa
