// Set-up and vector kernels of libspecann: value preparation, degrees, normalisation, SpMV tiling,
// per-alpha scaling, norms, the counter-based normal generator, rounding with signature, and the
// error/version entry points.
#include <cstddef>
#include <cstring>
#include <mutex>

#include "sa_reduce.hpp"

namespace sa {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? SA_E_OOM : SA_E_LAUNCH;
  }
  return SA_OK;
}

// ---------------------------------------------------------------------------------------------
// Phase 2 of the two-phase grid reductions (see sa_reduce.hpp).
template <class F>
__global__ __launch_bounds__(SA_BLOCK) void finish_kernel(F fin, FinishSpec spec) {
  __shared__ double red[SA_WAVES_PER_BLOCK];
  __shared__ uint64_t redu[SA_WAVES_PER_BLOCK];
  if (!fin.live()) return;
  void* work = fin.w();
  const int G = fin.g();
  sa_state* st = state_of(work);
  if (spec.only_if >= 0 && st->status != spec.only_if) return;
  const double* part = partials_of(work);
  const uint64_t* partu = reinterpret_cast<const uint64_t*>(part);
  const int C = spec.cols;
  for (int c = 0; c < C; ++c) {
    char* dst = reinterpret_cast<char*>(st) + spec.dst[c];
    if (spec.op[c] == FIN_SUM_U64) {
      uint64_t x = thread_partial_sum(partu, G, C, c);
      x = block_sum(x, redu);
      if (threadIdx.x == 0) *reinterpret_cast<uint64_t*>(dst) = x;
    } else if (spec.op[c] == FIN_MAX) {
      double x = -INFINITY;
      for (int b = threadIdx.x; b < G; b += blockDim.x) x = fmax(x, part[(int64_t)b * C + c]);
      x = wave_max(x);
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      if (lane == 0) red[wave] = x;
      __syncthreads();
      if (threadIdx.x == 0) {
        x = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        *reinterpret_cast<double*>(dst) = spec.take_sqrt[c] ? sqrt(x) : x;
      }
      __syncthreads();
    } else {
      double x = thread_partial_sum(part, G, C, c);
      x = block_sum(x, red);
      if (threadIdx.x == 0) *reinterpret_cast<double*>(dst) = spec.take_sqrt[c] ? sqrt(x) : x;
    }
  }
}

int launch_finish(void* work, int G, const FinishSpec& spec, hipStream_t stream) {
  hipLaunchKernelGGL(finish_kernel<OneFin>, dim3(1), dim3(SA_BLOCK), 0, stream, OneFin{work, G}, spec);
  return check_launch("finish");
}

int launch_finish_batch(const sa_batch* Bt, const FinishSpec& spec, hipStream_t stream) {
  hipLaunchKernelGGL(finish_kernel<BatchFin>, dim3(Bt->B), dim3(SA_BLOCK), 0, stream, bfin(Bt), spec);
  return check_launch("finish_batch");
}

static FinishSpec finish_spec(int cols, int only_if = -1) {
  FinishSpec f{};
  f.cols = cols;
  f.only_if = only_if;
  return f;
}

#define SA_STATE_OFF(field) ((int32_t)offsetof(sa_state, field))

// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void prepare_values_kernel(int64_t nnz, const double* __restrict__ raw,
                                                                  T* __restrict__ vals, int32_t* __restrict__ jint,
                                                                  int32_t* __restrict__ nonint) {
  int bad = 0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const double v = raw[e];
    vals[e] = (T)v;
    const bool integral = (v == rint(v)) && fabs(v) <= 1073741824.0;
    if (jint) jint[e] = integral ? (int32_t)v : 0;
    bad |= !integral;
  }
  if (__any(bad) && (threadIdx.x & 63) == 0 && nonint) atomicOr(nonint, 1);
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void abs_degree_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                              const T* __restrict__ vals, T* __restrict__ deg,
                                                              void* work) {
  double mx[1] = {0.0};
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int32_t e = indptr[r]; e < indptr[r + 1]; ++e) acc += fabs((double)vals[e]);
    const T dv = (T)acc;
    deg[r] = dv;
    mx[0] = fmax(mx[0], (double)dv);
  }
  block_partials<1, RED_MAX>(mx, work);
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void normalize_kernel(int64_t nnz, T* __restrict__ vals, int64_t n,
                                                             const T* __restrict__ deg, T* __restrict__ d, T c) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += stride) vals[e] = vals[e] / c;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += stride) d[r] = deg[r] / c;
}

// Per-row normalisation of a block-diagonal batch: row r and its entries are divided by c_row[r]
// (its instance's factor), the same operations as normalize_kernel per instance.
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void normalize_rows_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                                  T* __restrict__ vals, const T* __restrict__ deg,
                                                                  T* __restrict__ d, const T* __restrict__ c_row) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const T c = c_row[r];
    for (int32_t e = indptr[r]; e < indptr[r + 1]; ++e) vals[e] = vals[e] / c;
    d[r] = deg[r] / c;
  }
}

__global__ __launch_bounds__(SA_BLOCK) void tiles_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                         int32_t tile_nnz, int32_t* __restrict__ tile_rows,
                                                         int64_t ntiles) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t <= ntiles; t += (int64_t)gridDim.x * blockDim.x) {
    if (t == ntiles) {
      tile_rows[t] = (int32_t)n;
      continue;
    }
    const int64_t target = t * (int64_t)tile_nnz;
    int64_t lo = 0, hi = n;  // first r in [0, n] with indptr[r] >= target
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)indptr[mid] >= target) hi = mid; else lo = mid + 1;
    }
    tile_rows[t] = (int32_t)lo;
  }
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void alpha_scaling_kernel(int64_t n, const T* __restrict__ d, T alpha,
                                                                 T one_minus_alpha, T clamp, T* __restrict__ s,
                                                                 void* work) {
  double acc[2] = {0.0, -INFINITY};
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const T dr = d[r];
    const T dv = pow(dr, alpha);
    const T cl = dv > clamp ? dv : clamp;
    s[r] = (T)1 / sqrt(cl);
    acc[0] += (double)dv;
    acc[1] = fmax(acc[1], (double)pow(dr, one_minus_alpha));
  }
  // Column 0 is summed, column 1 is maximised (phase 2 on the host side of the launch).
  double sm[1] = {acc[0]}, mx[1] = {acc[1]};
  __shared__ double red[SA_WAVES_PER_BLOCK];
  double* part = partials_of(work);
  const double bs = block_sum(sm[0], red);
  double bm = wave_max(mx[0]);
  {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = bm;
    __syncthreads();
    if (threadIdx.x == 0) bm = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  }
  if (threadIdx.x == 0) {
    part[2 * (int64_t)blockIdx.x] = bs;
    part[2 * (int64_t)blockIdx.x + 1] = bm;
  }
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void norm_dot_kernel(int64_t n, const T* __restrict__ x,
                                                            const T* __restrict__ y, double a, double b,
                                                            void* work) {
  double acc[2] = {0.0, 0.0};
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const double xv = (double)x[r];
    const double yv = y ? (double)y[r] : 0.0;
    const double u = a * xv + b * yv;
    acc[0] += u * u;
    acc[1] += xv * yv;
  }
  block_partials<2, RED_SUM>(acc, work);
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void scale_copy_kernel(int64_t n, const T* __restrict__ x, double scale,
                                                              T* __restrict__ out) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    out[r] = (T)(scale * (double)x[r]);
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void combine3_kernel(int64_t n, double a, const T* __restrict__ x, double b,
                                                            const T* __restrict__ y, double c,
                                                            const T* __restrict__ z, T* __restrict__ out) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double v = a * (double)x[r];
    if (y) v += b * (double)y[r];
    if (z) v += c * (double)z[r];
    out[r] = (T)v;
  }
}

// out = max(d, 1e-12)^pexp (.) y in T (the diagonal predictor's scaling, utils/predictors.py:19-20).
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void diag_scale_kernel(int64_t n, const T* __restrict__ y,
                                                              const T* __restrict__ d, T pexp,
                                                              T* __restrict__ out) {
  const T clampv = (T)1e-12;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const T dr = d[r] > clampv ? d[r] : clampv;
    out[r] = pow(dr, pexp) * y[r];
  }
}

// Candidate-major int8 +-1 spins [k][n] -> bits [k][nwords] (bit l of word w: spin 64 w + l > 0), one
// ballot per wave of 64 nodes.
__global__ __launch_bounds__(SA_BLOCK) void pack_spins_kernel(int64_t n, int64_t nwords, const int8_t* __restrict__ spins,
                                                              uint64_t* __restrict__ bits) {
  const int c = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < nwords; w += waves) {
    const int64_t i = w * 64 + lane;
    const bool up = i < n && spins[(int64_t)c * n + i] > 0;
    const uint64_t m = __ballot(up);
    if (lane == 0) bits[(int64_t)c * nwords + w] = m;
  }
}

// Box-Muller on two 53-bit uniforms of the stream (seed, stream_id, i); u1 is in (0, 1].
__host__ __device__ inline double normal_at(uint64_t seed, uint64_t stream_id, uint64_t i) {
  const uint64_t h1 = mix3(seed, 0x5A17C0DE00000000ull ^ stream_id, i);
  const uint64_t h2 = splitmix64(h1);
  const double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void fill_normal_kernel(int64_t n, uint64_t seed, uint64_t stream_id,
                                                               T* __restrict__ out) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    out[r] = (T)normal_at(seed, stream_id, (uint64_t)r);
}

// One wavefront per 64-spin word: the ballot of the lanes is the word.
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void round_kernel(int64_t n, const T* __restrict__ y, double scale,
                                                         T* __restrict__ y_last, uint64_t* __restrict__ bits,
                                                         void* work) {
  const int lane = threadIdx.x & 63;
  const int64_t nwords = (n + 63) / 64;
  const int64_t wave0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwave = ((int64_t)gridDim.x * blockDim.x) >> 6;
  uint64_t hsum = 0;
  for (int64_t w = wave0; w < nwords; w += nwave) {
    const int64_t r = w * 64 + lane;
    bool up = false;
    if (r < n) {
      const T v = (T)(scale * (double)y[r]);
      if (y_last) y_last[r] = v;
      up = v >= (T)0;
    }
    const uint64_t word = __ballot(up);
    if (lane == 0) {
      bits[w] = word;
      hsum += splitmix64(word ^ splitmix64((uint64_t)w));
    }
  }
  block_partials_u64(hsum, work);
}

__global__ __launch_bounds__(SA_BLOCK) void bits_compare_kernel(int64_t nwords, const uint64_t* __restrict__ a,
                                                                const uint64_t* __restrict__ b, void* work) {
  uint64_t diff = 0;
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < nwords; w += (int64_t)gridDim.x * blockDim.x)
    diff += (a[w] != b[w]);
  block_partials_u64(diff, work);
}

static int finish_alpha_scaling(void* work, int G, hipStream_t stream) {
  FinishSpec f = finish_spec(2);
  f.op[0] = FIN_SUM;
  f.op[1] = FIN_MAX;
  f.dst[0] = SA_STATE_OFF(r[1]);
  f.dst[1] = SA_STATE_OFF(r[2]);
  return launch_finish(work, G, f, stream);
}

static int finish_norm_dot(void* work, int G, int slot, hipStream_t stream) {
  FinishSpec f = finish_spec(2);
  f.take_sqrt[0] = 1;
  f.dst[0] = SA_STATE_OFF(r[0]) + 8 * slot;
  f.dst[1] = SA_STATE_OFF(r[0]) + 8 * (slot + 1);
  return launch_finish(work, G, f, stream);
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_version(void) { return 100; }

int sa_last_error(char* buf, size_t len) {
  const std::string& m = g_last_error;
  if (buf && len) {
    size_t k = m.size() < len - 1 ? m.size() : len - 1;
    std::memcpy(buf, m.data(), k);
    buf[k] = 0;
  }
  return (int)m.size();
}

size_t sa_state_bytes(void) { return sizeof(sa_state); }

size_t sa_csr_bytes(void) { return sizeof(sa_csr); }

size_t sa_work_bytes(int64_t n) {
  (void)n;
  return state_bytes_padded() + (size_t)(kPartials + kRestart) * sizeof(double);
}

#define SA_GRID(nwork) grid_for((nwork), SA_BLOCK * 4, 4096)
#define SA_RGRID(nwork) grid_for((nwork), SA_BLOCK * 4, SA_RED_GRID)

int sa_prepare_values_f32(int64_t nnz, const double* raw, float* vals, int32_t* jint, int32_t* nonint,
                          hipStream_t stream) {
  SA_REQUIRE(nnz >= 0 && (nnz == 0 || (raw && vals)), "bad arguments");
  if (nnz == 0) return SA_OK;
  hipLaunchKernelGGL(prepare_values_kernel<float>, dim3(SA_GRID(nnz)), dim3(SA_BLOCK), 0, stream, nnz, raw, vals,
                     jint, nonint);
  return check_launch("prepare_values_f32");
}

int sa_prepare_values_f64(int64_t nnz, const double* raw, double* vals, int32_t* jint, int32_t* nonint,
                          hipStream_t stream) {
  SA_REQUIRE(nnz >= 0 && (nnz == 0 || (raw && vals)), "bad arguments");
  if (nnz == 0) return SA_OK;
  hipLaunchKernelGGL(prepare_values_kernel<double>, dim3(SA_GRID(nnz)), dim3(SA_BLOCK), 0, stream, nnz, raw, vals,
                     jint, nonint);
  return check_launch("prepare_values_f64");
}

int sa_abs_degree_f32(const sa_csr* A, float* deg, void* work, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && A->vals && deg && work, "bad arguments");
  const int G = SA_RGRID(A->n);
  hipLaunchKernelGGL(abs_degree_kernel<float>, dim3(G), dim3(SA_BLOCK), 0, stream, A->n, A->indptr,
                     (const float*)A->vals, deg, work);
  int rc = check_launch("abs_degree_f32");
  if (rc) return rc;
  FinishSpec f = finish_spec(1);
  f.op[0] = FIN_MAX;
  f.dst[0] = SA_STATE_OFF(r[0]);
  return launch_finish(work, G, f, stream);
}

int sa_abs_degree_f64(const sa_csr* A, double* deg, void* work, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && A->vals && deg && work, "bad arguments");
  const int G = SA_RGRID(A->n);
  hipLaunchKernelGGL(abs_degree_kernel<double>, dim3(G), dim3(SA_BLOCK), 0, stream, A->n, A->indptr,
                     (const double*)A->vals, deg, work);
  int rc = check_launch("abs_degree_f64");
  if (rc) return rc;
  FinishSpec f = finish_spec(1);
  f.op[0] = FIN_MAX;
  f.dst[0] = SA_STATE_OFF(r[0]);
  return launch_finish(work, G, f, stream);
}

int sa_normalize_f32(int64_t nnz, float* vals, int64_t n, const float* deg, float* d, double c, hipStream_t stream) {
  SA_REQUIRE(nnz >= 0 && n > 0 && vals && deg && d && c > 0, "bad arguments");
  hipLaunchKernelGGL(normalize_kernel<float>, dim3(SA_GRID(nnz > n ? nnz : n)), dim3(SA_BLOCK), 0, stream, nnz, vals,
                     n, deg, d, (float)c);
  return check_launch("normalize_f32");
}

int sa_normalize_f64(int64_t nnz, double* vals, int64_t n, const double* deg, double* d, double c,
                     hipStream_t stream) {
  SA_REQUIRE(nnz >= 0 && n > 0 && vals && deg && d && c > 0, "bad arguments");
  hipLaunchKernelGGL(normalize_kernel<double>, dim3(SA_GRID(nnz > n ? nnz : n)), dim3(SA_BLOCK), 0, stream, nnz,
                     vals, n, deg, d, c);
  return check_launch("normalize_f64");
}

int sa_normalize_rows_f32(const sa_csr* A, float* vals, const float* deg, float* d, const float* c_row,
                          hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && vals && deg && d && c_row, "bad arguments");
  hipLaunchKernelGGL(normalize_rows_kernel<float>, dim3(SA_GRID(A->n)), dim3(SA_BLOCK), 0, stream, A->n, A->indptr,
                     vals, deg, d, c_row);
  return check_launch("normalize_rows_f32");
}

int sa_normalize_rows_f64(const sa_csr* A, double* vals, const double* deg, double* d, const double* c_row,
                          hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && vals && deg && d && c_row, "bad arguments");
  hipLaunchKernelGGL(normalize_rows_kernel<double>, dim3(SA_GRID(A->n)), dim3(SA_BLOCK), 0, stream, A->n, A->indptr,
                     vals, deg, d, c_row);
  return check_launch("normalize_rows_f64");
}

int sa_csr_tiles(int64_t n, const int32_t* indptr, int32_t tile_nnz, int32_t* tile_rows, int64_t ntiles,
                 hipStream_t stream) {
  SA_REQUIRE(n > 0 && indptr && tile_rows && ntiles >= 1, "bad arguments");
  SA_REQUIRE(tile_nnz == 512 || tile_nnz == 1024 || tile_nnz == 2048 || tile_nnz == 4096,
             "tile_nnz must be 512, 1024, 2048 or 4096");
  hipLaunchKernelGGL(tiles_kernel, dim3(grid_for(ntiles + 1, SA_BLOCK, 4096)), dim3(SA_BLOCK), 0, stream, n, indptr,
                     tile_nnz, tile_rows, ntiles);
  return check_launch("csr_tiles");
}

int sa_alpha_scaling_f32(int64_t n, const float* d, double alpha, double clamp, float* s, void* work,
                         hipStream_t stream) {
  SA_REQUIRE(n > 0 && d && s && work, "bad arguments");
  const int G = SA_RGRID(n);
  hipLaunchKernelGGL(alpha_scaling_kernel<float>, dim3(G), dim3(SA_BLOCK), 0, stream, n, d, (float)alpha,
                     (float)(1.0 - alpha), (float)clamp, s, work);
  int rc = check_launch("alpha_scaling_f32");
  if (rc) return rc;
  return finish_alpha_scaling(work, G, stream);
}

int sa_alpha_scaling_f64(int64_t n, const double* d, double alpha, double clamp, double* s, void* work,
                         hipStream_t stream) {
  SA_REQUIRE(n > 0 && d && s && work, "bad arguments");
  const int G = SA_RGRID(n);
  hipLaunchKernelGGL(alpha_scaling_kernel<double>, dim3(G), dim3(SA_BLOCK), 0, stream, n, d, alpha, 1.0 - alpha,
                     clamp, s, work);
  int rc = check_launch("alpha_scaling_f64");
  if (rc) return rc;
  return finish_alpha_scaling(work, G, stream);
}

int sa_norm_dot_f32(void* work, int64_t n, const float* x, const float* y, double a, double b, int32_t slot,
                    hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && x && slot >= 0 && slot <= 6, "bad arguments");
  const int G = SA_RGRID(n);
  hipLaunchKernelGGL(norm_dot_kernel<float>, dim3(G), dim3(SA_BLOCK), 0, stream, n, x, y, a, b, work);
  int rc = check_launch("norm_dot_f32");
  if (rc) return rc;
  return finish_norm_dot(work, G, slot, stream);
}

int sa_norm_dot_f64(void* work, int64_t n, const double* x, const double* y, double a, double b, int32_t slot,
                    hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && x && slot >= 0 && slot <= 6, "bad arguments");
  const int G = SA_RGRID(n);
  hipLaunchKernelGGL(norm_dot_kernel<double>, dim3(G), dim3(SA_BLOCK), 0, stream, n, x, y, a, b, work);
  int rc = check_launch("norm_dot_f64");
  if (rc) return rc;
  return finish_norm_dot(work, G, slot, stream);
}

int sa_scale_copy_f32(int64_t n, const float* x, double scale, float* out, hipStream_t stream) {
  SA_REQUIRE(n > 0 && x && out, "bad arguments");
  hipLaunchKernelGGL(scale_copy_kernel<float>, dim3(SA_GRID(n)), dim3(SA_BLOCK), 0, stream, n, x, scale, out);
  return check_launch("scale_copy_f32");
}

int sa_scale_copy_f64(int64_t n, const double* x, double scale, double* out, hipStream_t stream) {
  SA_REQUIRE(n > 0 && x && out, "bad arguments");
  hipLaunchKernelGGL(scale_copy_kernel<double>, dim3(SA_GRID(n)), dim3(SA_BLOCK), 0, stream, n, x, scale, out);
  return check_launch("scale_copy_f64");
}

int sa_diag_scale_f32(int64_t n, const float* y, const float* d, double pexp, float* out, hipStream_t stream) {
  SA_REQUIRE(n > 0 && y && d && out, "bad arguments");
  hipLaunchKernelGGL(diag_scale_kernel<float>, dim3(SA_GRID(n)), dim3(SA_BLOCK), 0, stream, n, y, d, (float)pexp, out);
  return check_launch("diag_scale_f32");
}

int sa_diag_scale_f64(int64_t n, const double* y, const double* d, double pexp, double* out, hipStream_t stream) {
  SA_REQUIRE(n > 0 && y && d && out, "bad arguments");
  hipLaunchKernelGGL(diag_scale_kernel<double>, dim3(SA_GRID(n)), dim3(SA_BLOCK), 0, stream, n, y, d, pexp, out);
  return check_launch("diag_scale_f64");
}

int sa_pack_spins(int64_t n, int32_t k, const int8_t* spins, int64_t nwords, uint64_t* bits, hipStream_t stream) {
  SA_REQUIRE(n > 0 && k >= 1 && k <= 65535 && spins && bits && nwords * 64 >= n, "bad arguments");
  hipLaunchKernelGGL(pack_spins_kernel, dim3(grid_for(nwords, SA_WAVES_PER_BLOCK, 4096), k), dim3(SA_BLOCK), 0, stream,
                     n, nwords, spins, bits);
  return check_launch("pack_spins");
}

int sa_combine3_f32(int64_t n, double a, const float* x, double b, const float* y, double c, const float* z,
                    float* out, hipStream_t stream) {
  SA_REQUIRE(n > 0 && x && out && (y || b == 0.0) && (z || c == 0.0), "bad arguments");
  hipLaunchKernelGGL(combine3_kernel<float>, dim3(SA_GRID(n)), dim3(SA_BLOCK), 0, stream, n, a, x, y ? b : 0.0, y,
                     z ? c : 0.0, z, out);
  return check_launch("combine3_f32");
}

int sa_combine3_f64(int64_t n, double a, const double* x, double b, const double* y, double c, const double* z,
                    double* out, hipStream_t stream) {
  SA_REQUIRE(n > 0 && x && out && (y || b == 0.0) && (z || c == 0.0), "bad arguments");
  hipLaunchKernelGGL(combine3_kernel<double>, dim3(SA_GRID(n)), dim3(SA_BLOCK), 0, stream, n, a, x, y ? b : 0.0, y,
                     z ? c : 0.0, z, out);
  return check_launch("combine3_f64");
}

int sa_fill_normal_f32(int64_t n, uint64_t seed, uint64_t stream_id, float* out, hipStream_t stream) {
  SA_REQUIRE(n > 0 && out, "bad arguments");
  hipLaunchKernelGGL(fill_normal_kernel<float>, dim3(SA_GRID(n)), dim3(SA_BLOCK), 0, stream, n, seed, stream_id, out);
  return check_launch("fill_normal_f32");
}

int sa_fill_normal_f64(int64_t n, uint64_t seed, uint64_t stream_id, double* out, hipStream_t stream) {
  SA_REQUIRE(n > 0 && out, "bad arguments");
  hipLaunchKernelGGL(fill_normal_kernel<double>, dim3(SA_GRID(n)), dim3(SA_BLOCK), 0, stream, n, seed, stream_id,
                     out);
  return check_launch("fill_normal_f64");
}

int sa_round_f32(void* work, int64_t n, const float* y, double scale, float* y_last, uint64_t* bits,
                 hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && y && bits, "bad arguments");
  const int64_t nwords = (n + 63) / 64;
  const int G = grid_for(nwords, SA_WAVES_PER_BLOCK * 4, SA_RED_GRID);
  hipLaunchKernelGGL(round_kernel<float>, dim3(G), dim3(SA_BLOCK), 0, stream, n, y, scale, y_last, bits, work);
  int rc = check_launch("round_f32");
  if (rc) return rc;
  FinishSpec f = finish_spec(1);
  f.op[0] = FIN_SUM_U64;
  f.dst[0] = SA_STATE_OFF(hash);
  return launch_finish(work, G, f, stream);
}

int sa_round_f64(void* work, int64_t n, const double* y, double scale, double* y_last, uint64_t* bits,
                 hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && y && bits, "bad arguments");
  const int64_t nwords = (n + 63) / 64;
  const int G = grid_for(nwords, SA_WAVES_PER_BLOCK * 4, SA_RED_GRID);
  hipLaunchKernelGGL(round_kernel<double>, dim3(G), dim3(SA_BLOCK), 0, stream, n, y, scale, y_last, bits, work);
  int rc = check_launch("round_f64");
  if (rc) return rc;
  FinishSpec f = finish_spec(1);
  f.op[0] = FIN_SUM_U64;
  f.dst[0] = SA_STATE_OFF(hash);
  return launch_finish(work, G, f, stream);
}

int sa_bits_compare(void* work, int64_t nwords, const uint64_t* a, const uint64_t* b, hipStream_t stream) {
  SA_REQUIRE(work && nwords > 0 && a && b, "bad arguments");
  const int G = SA_RGRID(nwords);
  hipLaunchKernelGGL(bits_compare_kernel, dim3(G), dim3(SA_BLOCK), 0, stream, nwords, a, b, work);
  int rc = check_launch("bits_compare");
  if (rc) return rc;
  FinishSpec f = finish_spec(1);
  f.op[0] = FIN_SUM_U64;
  f.dst[0] = SA_STATE_OFF(i64[0]);
  return launch_finish(work, G, f, stream);
}

}  // extern "C"
