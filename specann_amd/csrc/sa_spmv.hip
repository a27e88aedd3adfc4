// Fused scaled CSR SpMV for gfx950.
//
// The matrix is cut into tiles of tile_nnz stored entries; workgroup t owns the rows whose first
// entry lies in tile t. Phase 1 streams the tile's column indices and values (non-temporal loads,
// fully coalesced), gathers g at the column and keeps the exact FP64 product in LDS. Phase 2 sums
// each row's products with P lanes per row (P from the tile's row count) and a shuffle reduction,
// then applies the epilogue a x_r + b s_r sum. A row that runs past the end of its tile (only the
// last row of a tile can) is summed by the whole workgroup. All sums run in a fixed order, so the
// result is bitwise reproducible; row sums are FP64 even in the FP32 form.
#include "sa_spmv.hpp"

namespace sa {

template <typename T>
__device__ inline void spmv_store(const SpmvParams<T>& p, int32_t r, double acc, double bb) {
  const double sr = p.s ? (double)p.s[r] : 1.0;
  double v = bb * sr * acc;
  if (p.x) v += p.a * (double)p.x[r];
  p.y[r] = (T)v;
}

template <typename T, int TILE>
__global__ __launch_bounds__(SA_BLOCK) void spmv_tile_kernel(SpmvParams<T> p) {
  __shared__ double prod[TILE];
  __shared__ double red[SA_WAVES_PER_BLOCK];
  double bb = p.b;
  if (p.lz) {
    if (p.lz->status != SA_RUNNING) return;
    bb *= p.lz->sigma[p.lz->j];
  }
  const int64_t t = blockIdx.x;
  const int32_t r0 = p.tile_rows[t], r1 = p.tile_rows[t + 1];
  if (r0 >= r1) return;
  const int32_t e0 = p.indptr[r0];
  const int32_t e_last = p.indptr[r1 - 1];
  const int32_t e1 = p.indptr[r1];
  const bool long_last = (e1 - e0) > TILE;
  const int32_t rs = long_last ? r1 - 1 : r1;
  const int32_t es = long_last ? e_last : e1;

  // Phase 1: products of the short rows into LDS.
  for (int32_t e = e0 + (int32_t)threadIdx.x; e < es; e += SA_BLOCK) {
    const int32_t c = stream_load(&p.indices[e]);
    const T v = stream_load(&p.vals[e]);
    prod[e - e0] = (double)v * (double)p.g[c];
  }
  __syncthreads();

  // Phase 2: segmented row sums, P lanes per row.
  const int32_t nrows = rs - r0;
  if (nrows > 0) {
    int P = p.lanes_fixed;
    if (P <= 0) {
      const int q = SA_BLOCK / nrows;
      P = q >= 64 ? 64 : q >= 32 ? 32 : q >= 16 ? 16 : q >= 8 ? 8 : q >= 4 ? 4 : q >= 2 ? 2 : 1;
    }
    const int groups = SA_BLOCK / P;
    const int gid = threadIdx.x / P, lig = threadIdx.x % P;
    for (int32_t rb = r0; rb < rs; rb += groups) {
      const int32_t r = rb + gid;
      double acc = 0.0;
      if (r < rs) {
        const int32_t b0 = p.indptr[r] - e0, b1 = p.indptr[r + 1] - e0;
        for (int32_t e = b0 + lig; e < b1; e += P) acc += prod[e];
      }
      for (int o = P >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if (lig == 0 && r < rs) spmv_store(p, r, acc, bb);
    }
  }

  // A row longer than what is left of its tile: the whole workgroup strides over it.
  if (long_last) {
    const int32_t r = r1 - 1;
    double acc = 0.0;
    for (int32_t e = e_last + (int32_t)threadIdx.x; e < e1; e += SA_BLOCK) {
      const int32_t c = stream_load(&p.indices[e]);
      acc += (double)stream_load(&p.vals[e]) * (double)p.g[c];
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) spmv_store(p, r, acc, bb);
  }
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void prescale_kernel(int64_t n, const T* __restrict__ s,
                                                            const T* __restrict__ x, T* __restrict__ z) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    z[r] = s ? s[r] * x[r] : x[r];
}

template <typename T>
int launch_spmv(const SpmvParams<T>& p, int32_t tile_nnz, hipStream_t stream) {
  if (p.ntiles < 1 || p.ntiles >= (1ll << 31)) {
    set_error("spmv: ntiles out of range");
    return SA_E_ARG;
  }
  const dim3 grid((unsigned)p.ntiles), block(SA_BLOCK);
  switch (tile_nnz) {
    case 512: hipLaunchKernelGGL((spmv_tile_kernel<T, 512>), grid, block, 0, stream, p); break;
    case 1024: hipLaunchKernelGGL((spmv_tile_kernel<T, 1024>), grid, block, 0, stream, p); break;
    case 2048: hipLaunchKernelGGL((spmv_tile_kernel<T, 2048>), grid, block, 0, stream, p); break;
    case 4096: hipLaunchKernelGGL((spmv_tile_kernel<T, 4096>), grid, block, 0, stream, p); break;
    default: set_error("spmv: tile_nnz must be 512, 1024, 2048 or 4096"); return SA_E_ARG;
  }
  return check_launch("spmv");
}

template <typename T>
int launch_prescale(int64_t n, const T* s, const T* x, T* z, hipStream_t stream) {
  hipLaunchKernelGGL(prescale_kernel<T>, dim3(grid_for(n, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK), 0, stream, n, s, x,
                     z);
  return check_launch("prescale");
}

template int launch_spmv<float>(const SpmvParams<float>&, int32_t, hipStream_t);
template int launch_spmv<double>(const SpmvParams<double>&, int32_t, hipStream_t);
template int launch_prescale<float>(int64_t, const float*, const float*, float*, hipStream_t);
template int launch_prescale<double>(int64_t, const double*, const double*, double*, hipStream_t);

template <typename T>
static int spmv_scaled(const sa_csr* A, const T* s, const T* x, double a, double b, T* z, T* y, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && A->indices && A->vals && A->tile_rows && x && z && y,
             "bad arguments");
  SA_REQUIRE(A->nnz >= 0 && A->nnz < (1ll << 31), "nnz must fit int32 indices");
  int rc = launch_prescale<T>(A->n, s, x, z, stream);
  if (rc) return rc;
  SpmvParams<T> p;
  p.n = A->n;
  p.indptr = A->indptr;
  p.indices = A->indices;
  p.vals = (const T*)A->vals;
  p.tile_rows = A->tile_rows;
  p.ntiles = A->ntiles;
  p.lanes_fixed = A->lanes_per_row;
  p.g = z;
  p.s = s;
  p.x = (a != 0.0) ? x : nullptr;
  p.a = a;
  p.b = b;
  p.lz = nullptr;
  p.y = y;
  return launch_spmv<T>(p, A->tile_nnz, stream);
}

template <typename T>
static int spmv_gather(const sa_csr* A, const T* s, const T* g, double a, const T* x, double b, T* y,
                       hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && A->indices && A->vals && A->tile_rows && g && y, "bad arguments");
  SA_REQUIRE(a == 0.0 || x, "x is required when a != 0");
  SpmvParams<T> p;
  p.n = A->n;
  p.indptr = A->indptr;
  p.indices = A->indices;
  p.vals = (const T*)A->vals;
  p.tile_rows = A->tile_rows;
  p.ntiles = A->ntiles;
  p.lanes_fixed = A->lanes_per_row;
  p.g = g;
  p.s = s;
  p.x = (a != 0.0) ? x : nullptr;
  p.a = a;
  p.b = b;
  p.lz = nullptr;
  p.y = y;
  return launch_spmv<T>(p, A->tile_nnz, stream);
}

}  // namespace sa

extern "C" {

int sa_spmv_gather_f32(const sa_csr* A, const float* s, const float* g, double a, const float* x, double b, float* y,
                       hipStream_t stream) {
  return sa::spmv_gather<float>(A, s, g, a, x, b, y, stream);
}

int sa_spmv_gather_f64(const sa_csr* A, const double* s, const double* g, double a, const double* x, double b,
                       double* y, hipStream_t stream) {
  return sa::spmv_gather<double>(A, s, g, a, x, b, y, stream);
}

int sa_spmv_scaled_f32(const sa_csr* A, const float* s, const float* x, double a, double b, float* z, float* y,
                       hipStream_t stream) {
  return sa::spmv_scaled<float>(A, s, x, a, b, z, y, stream);
}

int sa_spmv_scaled_f64(const sa_csr* A, const double* s, const double* x, double a, double b, double* z, double* y,
                       hipStream_t stream) {
  return sa::spmv_scaled<double>(A, s, x, a, b, z, y, stream);
}

}  // extern "C"
