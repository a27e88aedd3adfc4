// Warm-started Lanczos for the leading eigenpair of N_alpha = s (.) (J/c) (.) s on gfx950.
//
// The recurrence runs on N itself, not on the reference's shifted operator S_rho = rho I - N: the
// Krylov spaces of the two are identical (T_S = rho I - T_N), and without the shift the FP32 basis
// keeps its orthogonality (with rho ~ 1 dominating the product, S_rho v loses the N v part to
// cancellation). The caller reports mu = rho - theta, the eigenvalue the reference's eigsh returns.
//
// One step = three launches:
//   1. the fused scaled SpMV  W = sigma_j s (.) (J Z),  Z = s (.) V[j] written by the previous step;
//   2. projections h_m = v_m . W for every basis vector of the cycle (FP64 partial sums, 8 vectors
//      per chunk, last-workgroup reduction);
//   3. classical Gram-Schmidt W - sum h_m v_m into V[j+1] (raw; sigma_{j+1} = 1/beta_j), Z for the
//      next step, beta_j = ||.||, then, in the last workgroup, the FP64 Ritz solve of the
//      tridiagonal T_{j+1}: largest eigenvalue by 257-section Sturm bisection over the whole
//      workgroup, its eigenvector by inverse iteration with a pivoted tridiagonal LU, and the
//      residual estimate beta_j |s_j|, the stopping test the reference's eigsh applies per restart
//      (here it is applied every step).
// Every launch reads the status in the work buffer first and does nothing unless the solve is
// running, so the host may queue several steps (and the restart pair) without a round trip.
#include "sa_reduce.hpp"
#include "sa_spmv.hpp"

namespace sa {

__global__ void lz_begin_kernel(void* work, int32_t ncv, int32_t maxiter, double tol, double bound,
                                double pred_tol) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  sa_state* st = state_of(work);
  st->status = SA_RUNNING;
  st->j = 0;
  st->steps = 0;
  st->cycles = 0;
  st->ncv = ncv;
  st->maxiter = maxiter;
  st->gate = 0;
  st->flags = 0;
  st->theta = 0.0;
  st->res = INFINITY;
  st->first_beta = NAN;
  st->tol = tol;
  st->bound = bound;
  st->pred_tol = pred_tol;
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void lz_start_kernel(void* work, int64_t n, const T* __restrict__ s,
                                                            const T* __restrict__ x, const T* __restrict__ d,
                                                            T pexp, double cx, const T* __restrict__ y, double cy,
                                                            T* __restrict__ V, T* __restrict__ Z, int32_t gate,
                                                            int32_t only_if) {
  sa_state* st = state_of(work);
  if (only_if >= 0 && st->status != only_if) return;
  double acc[1] = {0.0};
  const T clampv = (T)1e-12;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    T v;
    if (y) {
      v = (T)(cx * (double)x[r] + cy * (double)y[r]);
    } else {
      v = x[r];
      if (cx != 1.0) v = (T)(cx * (double)v);
    }
    if (d && pexp != (T)0) {
      const T dr = d[r] > clampv ? d[r] : clampv;
      v = pow(dr, pexp) * v;
    }
    V[r] = v;
    Z[r] = s ? s[r] * v : v;
    acc[0] += (double)v * (double)v;
  }
  double out[1];
  if (grid_reduce<1, RED_SUM>(acc, work, out) && threadIdx.x == 0) {
    const double nrm = sqrt(out[0]);
    st->r[3] = nrm;
    const bool ok = nrm > 0.0 && isfinite(nrm);
    st->sigma[0] = ok ? 1.0 / nrm : 0.0;
    st->j = 0;
    st->gate = gate;
    st->res = INFINITY;
    if (only_if == SA_RESTART) {
      st->cycles += 1;  // a restart keeps the predictor residual of the solve's first step
    } else {
      st->first_beta = NAN;
    }
    st->status = ok ? SA_RUNNING : SA_BROKEN;
  }
}

// W = sigma_j s (.) sum_p partial[p] (XCD-panel layout; panel order fixed).
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void lz_combine_kernel(void* work, int64_t n, int32_t npanels,
                                                              const T* __restrict__ partial, const T* __restrict__ s,
                                                              T* __restrict__ W) {
  const sa_state* st = state_of(work);
  if (st->status != SA_RUNNING) return;
  const double sigma = st->sigma[st->j];
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int p = 0; p < npanels; ++p) acc += (double)partial[(int64_t)p * n + r];
    W[r] = (T)(sigma * (s ? (double)s[r] : 1.0) * acc);
  }
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void lz_dots_kernel(void* work, int64_t n, const T* __restrict__ V,
                                                           const T* __restrict__ W) {
  __shared__ double red[SA_WAVES_PER_BLOCK];
  __shared__ bool last;
  sa_state* st = state_of(work);
  if (st->status != SA_RUNNING) return;
  const int j = st->j;
  const int nch = j / SA_RED_COLS + 1;
  const int ch = blockIdx.y;
  if (ch >= nch) return;
  const int m0 = ch * SA_RED_COLS;
  const int mc = min(SA_RED_COLS, j + 1 - m0);
  double acc[SA_RED_COLS];
#pragma unroll
  for (int c = 0; c < SA_RED_COLS; ++c) acc[c] = 0.0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const double w = (double)W[r];
#pragma unroll
    for (int c = 0; c < SA_RED_COLS; ++c)
      if (c < mc) acc[c] += (double)V[(int64_t)(m0 + c) * n + r] * w;
  }
  double* part = partials_of(work);
  const int64_t G = gridDim.x;
#pragma unroll
  for (int c = 0; c < SA_RED_COLS; ++c) {
    const double bs = block_sum(acc[c], red);
    if (threadIdx.x == 0) part[((int64_t)ch * G + blockIdx.x) * SA_RED_COLS + c] = bs;
  }
  if (!last_block_arrive(&st->counters[0], (unsigned)(G * nch), &last)) return;
  for (int m = 0; m <= j; ++m) {
    const int chm = m / SA_RED_COLS, cm = m % SA_RED_COLS;
    double x = 0.0;
    for (int64_t b = threadIdx.x; b < G; b += blockDim.x)
      x += load_coherent(&part[((int64_t)chm * G + b) * SA_RED_COLS + cm]);
    x = block_sum(x, red);
    if (threadIdx.x == 0) st->h[m] = st->sigma[m] * x;
  }
  if (threadIdx.x == 0) st->a[j] = st->h[j];
}

// Sturm count: number of eigenvalues of the symmetric tridiagonal (a, b) below x.
__device__ inline int sturm_count(int m, const double* a, const double* b2, double x, double pivmin) {
  double q = a[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int cnt = q < 0.0;
  for (int i = 1; i < m; ++i) {
    q = (a[i] - x) - b2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// Largest eigenpair of the m x m tridiagonal (a, b) held in LDS. All threads of the workgroup take
// part; on return theta is valid in every thread and the unit eigenvector is in vec (LDS).
__device__ double ritz_largest(int m, const double* a, const double* b, double* b2, double* work5, int* piv,
                               double* vec) {
  __shared__ double sh[4];
  __shared__ unsigned long long masks[SA_WAVES_PER_BLOCK];
  const int t = threadIdx.x;
  // Gershgorin interval and the pivot floor.
  double glo = INFINITY, ghi = -INFINITY, mb2 = 0.0;
  for (int i = t; i < m; i += blockDim.x) {
    const double bl = i > 0 ? fabs(b[i - 1]) : 0.0;
    const double br = i < m - 1 ? fabs(b[i]) : 0.0;
    glo = fmin(glo, a[i] - bl - br);
    ghi = fmax(ghi, a[i] + bl + br);
    if (i < m - 1) {
      b2[i] = b[i] * b[i];
      mb2 = fmax(mb2, b2[i]);
    }
  }
  glo = -wave_max(-glo);
  ghi = wave_max(ghi);
  mb2 = wave_max(mb2);
  if ((t & 63) == 0) masks[t >> 6] = 0;
  __shared__ double wlo[SA_WAVES_PER_BLOCK], whi[SA_WAVES_PER_BLOCK], wb2[SA_WAVES_PER_BLOCK];
  if ((t & 63) == 0) {
    wlo[t >> 6] = glo;
    whi[t >> 6] = ghi;
    wb2[t >> 6] = mb2;
  }
  __syncthreads();
  if (t == 0) {
    double L = wlo[0], H = whi[0], B = wb2[0];
    for (int w = 1; w < SA_WAVES_PER_BLOCK; ++w) {
      L = fmin(L, wlo[w]);
      H = fmax(H, whi[w]);
      B = fmax(B, wb2[w]);
    }
    const double tn = fmax(fabs(L), fabs(H));
    const double pad = 4.0 * 2.220446049250313e-16 * tn + 1e-300;
    sh[0] = L - pad;
    sh[1] = H + pad;
    sh[2] = 1e-290 * fmax(1.0, B);
    sh[3] = tn;
  }
  __syncthreads();
  double lo = sh[0], hi = sh[1];
  const double pivmin = sh[2], tnorm = sh[3];
  const int nsect = blockDim.x + 1;
  for (int round = 0; round < 24; ++round) {
    const double width = hi - lo;
    if (width <= 4.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi)) + 2.0 * pivmin) break;
    const double x = lo + width * (double)(t + 1) / (double)nsect;
    const int c = sturm_count(m, a, b2, x, pivmin);
    const unsigned long long mk = __ballot(c >= m);
    if ((t & 63) == 0) masks[t >> 6] = mk;
    __syncthreads();
    int first = -1;
    for (int w = 0; w < SA_WAVES_PER_BLOCK; ++w) {
      if (masks[w]) {
        first = w * 64 + __builtin_ctzll(masks[w]);
        break;
      }
    }
    __syncthreads();
    if (first < 0) {
      lo = lo + width * (double)(nsect - 1) / (double)nsect;
    } else {
      const double nhi = lo + width * (double)(first + 1) / (double)nsect;
      const double nlo = first == 0 ? lo : lo + width * (double)first / (double)nsect;
      lo = nlo;
      hi = nhi;
    }
  }
  const double theta = 0.5 * (lo + hi);

  // Inverse iteration for the eigenvector (thread 0; m <= SA_MAX_NCV).
  if (t == 0) {
    double* d = work5;
    double* dl = work5 + SA_MAX_NCV;
    double* du = work5 + 2 * SA_MAX_NCV;
    double* du2 = work5 + 3 * SA_MAX_NCV;
    const double tiny = 2.220446049250313e-16 * fmax(tnorm, 1e-300);
    if (m == 1) {
      vec[0] = 1.0;
    } else {
      for (int i = 0; i < m; ++i) {
        d[i] = a[i] - theta;
        if (i < m - 1) {
          dl[i] = b[i];
          du[i] = b[i];
          du2[i] = 0.0;
        }
        piv[i] = i;
      }
      // LU with partial pivoting of the tridiagonal (row interchanges create a second superdiagonal).
      for (int i = 0; i < m - 1; ++i) {
        if (fabs(d[i]) >= fabs(dl[i])) {
          if (d[i] == 0.0) d[i] = tiny;
          const double f = dl[i] / d[i];
          dl[i] = f;
          d[i + 1] -= f * du[i];
          du2[i] = 0.0;
          piv[i] = i;
        } else {
          const double f = d[i] / dl[i];
          d[i] = dl[i];
          dl[i] = f;
          const double tmp = du[i];
          du[i] = d[i + 1];
          d[i + 1] = tmp - f * d[i + 1];
          if (i < m - 2) {
            du2[i] = du[i + 1];
            du[i + 1] = -f * du[i + 1];
          }
          piv[i] = i + 1;
        }
      }
      if (d[m - 1] == 0.0) d[m - 1] = tiny;
      for (int i = 0; i < m; ++i) vec[i] = 1.0 + 0.01 * (double)(i % 7);
      for (int it = 0; it < 3; ++it) {
        // forward: apply P and L
        for (int i = 0; i < m - 1; ++i) {
          if (piv[i] == i) {
            vec[i + 1] -= dl[i] * vec[i];
          } else {
            const double tmp = vec[i];
            vec[i] = vec[i + 1];
            vec[i + 1] = tmp - dl[i] * vec[i];
          }
        }
        // backward: U with two superdiagonals
        vec[m - 1] /= d[m - 1];
        vec[m - 2] = (vec[m - 2] - du[m - 2] * vec[m - 1]) / d[m - 2];
        for (int i = m - 3; i >= 0; --i) vec[i] = (vec[i] - du[i] * vec[i + 1] - du2[i] * vec[i + 2]) / d[i];
        double mx = 0.0;
        for (int i = 0; i < m; ++i) mx = fmax(mx, fabs(vec[i]));
        if (!(mx > 0.0) || !isfinite(mx)) break;
        double nn = 0.0;
        for (int i = 0; i < m; ++i) {
          vec[i] /= mx;
          nn += vec[i] * vec[i];
        }
        nn = sqrt(nn);
        for (int i = 0; i < m; ++i) vec[i] /= nn;
      }
    }
  }
  __syncthreads();
  return theta;
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void lz_orth_kernel(void* work, int64_t n, const T* __restrict__ s,
                                                           T* __restrict__ V, T* __restrict__ Z,
                                                           const T* __restrict__ W) {
  __shared__ double coef[SA_MAX_NCV + 1];
  __shared__ double red[SA_WAVES_PER_BLOCK];
  __shared__ bool last;
  sa_state* st = state_of(work);
  if (st->status != SA_RUNNING) return;
  const int j = st->j;
  for (int m = threadIdx.x; m <= j; m += blockDim.x) coef[m] = st->h[m] * st->sigma[m];
  __syncthreads();
  T* __restrict__ vn = V + (int64_t)(j + 1) * n;
  double acc = 0.0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double w = (double)W[r];
    for (int m = 0; m <= j; ++m) w -= coef[m] * (double)V[(int64_t)m * n + r];
    const T wt = (T)w;
    vn[r] = wt;
    Z[r] = s ? s[r] * wt : wt;
    acc += (double)wt * (double)wt;
  }
  double* part = partials_of(work);
  const double bs = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
  if (!last_block_arrive(&st->counters[0], gridDim.x, &last)) return;
  double tot = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) tot += load_coherent(&part[b]);
  tot = block_sum(tot, red);
  __shared__ double beta_sh;
  if (threadIdx.x == 0) beta_sh = sqrt(tot);
  __syncthreads();
  const double beta = beta_sh;

  __shared__ double la[SA_MAX_NCV], lb[SA_MAX_NCV], lb2[SA_MAX_NCV], lvec[SA_MAX_NCV];
  __shared__ double lwork[4 * SA_MAX_NCV];
  __shared__ int lpiv[SA_MAX_NCV];
  const int m = j + 1;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    la[i] = st->a[i];
    lb[i] = i < j ? st->b[i] : 0.0;
  }
  __syncthreads();
  const double aj = la[j];
  const bool finite = isfinite(aj) && isfinite(beta);
  const bool big = fabs(aj) > 10.0 * st->bound || beta > 10.0 * st->bound;
  double theta = 0.0;
  bool solved = false;
  bool all_finite = finite && !big;
  if (all_finite) {
    for (int i = 0; i < j; ++i) all_finite = all_finite && isfinite(lb[i]);
  }
  if (all_finite) {
    theta = ritz_largest(m, la, lb, lb2, lwork, lpiv, lvec);
    solved = true;
  }
  if (threadIdx.x == 0) {
    st->b[j] = beta;
    st->steps += 1;
    if (!solved) {
      st->status = SA_BROKEN;
    } else {
      for (int i = 0; i < m; ++i) st->ritz[i] = lvec[i];
      const double res = beta * fabs(lvec[m - 1]);
      st->theta = theta;
      st->res = res;
      int status = SA_RUNNING;
      if (!isfinite(theta) || !isfinite(res)) status = SA_BROKEN;
      if (status == SA_RUNNING && st->gate && j == 0) {
        st->first_beta = beta;
        st->gate = 0;
        if (beta > st->pred_tol) status = SA_REJECTED;
      }
      if (status == SA_RUNNING) {
        if (res <= st->tol) status = SA_CONVERGED;
        else if (st->steps >= st->maxiter) status = SA_MAXITER;
        else if (m >= st->ncv) status = SA_RESTART;
      }
      st->sigma[j + 1] = beta > 0.0 ? 1.0 / beta : 0.0;
      st->j = m;
      st->status = status;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void lz_ritz_kernel(void* work, int64_t n, const T* __restrict__ V,
                                                           const T* __restrict__ y_last, T* __restrict__ y,
                                                           int32_t only_if) {
  __shared__ double coef[SA_MAX_NCV];
  sa_state* st = state_of(work);
  if (only_if >= 0 && st->status != only_if) return;
  const int m = st->j;
  for (int i = threadIdx.x; i < m; i += blockDim.x) coef[i] = st->ritz[i] * st->sigma[i];
  __syncthreads();
  double acc[2] = {0.0, 0.0};
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int i = 0; i < m; ++i) v += coef[i] * (double)V[(int64_t)i * n + r];
    const T vt = (T)v;
    y[r] = vt;
    acc[0] += (double)vt * (double)vt;
    if (y_last) acc[1] += (double)vt * (double)y_last[r];
  }
  double out[2];
  if (grid_reduce<2, RED_SUM>(acc, work, out) && threadIdx.x == 0) {
    st->r[4] = sqrt(out[0]);
    st->r[5] = out[1];
  }
}

static inline int red_grid(int64_t n) { return grid_for(n, SA_BLOCK * 4, SA_RED_GRID); }

template <typename T>
static int lanczos_start(void* work, int64_t n, const T* s, const T* x, const T* d, double pexp, double cx,
                         const T* y, double cy, T* V, T* Z, int32_t gate, int32_t only_if, hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && x && V && Z, "bad arguments");
  hipLaunchKernelGGL(lz_start_kernel<T>, dim3(red_grid(n)), dim3(SA_BLOCK), 0, stream, work, n, s, x, d, (T)pexp, cx,
                     y, cy, V, Z, gate, only_if);
  return check_launch("lanczos_start");
}

template <typename T>
static int lanczos_step(const sa_csr* A, void* work, const T* s, T* V, T* Z, T* W, hipStream_t stream) {
  SA_REQUIRE(A && work && V && Z && W && A->tile_rows, "bad arguments");
  const int64_t n = A->n;
  SpmvParams<T> p = params_from<T>(A);
  p.g = Z;
  p.s = s;
  p.lz = reinterpret_cast<const sa_state*>(work);
  p.y = W;
  int rc = launch_spmv<T>(p, A->tile_nnz, stream);
  if (rc) return rc;
  if (A->npanels > 0) {
    hipLaunchKernelGGL(lz_combine_kernel<T>, dim3(grid_for(n, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK), 0, stream, work,
                       n, A->npanels, (const T*)A->partial, s, W);
    rc = check_launch("lanczos_combine");
    if (rc) return rc;
  }
  const int G = red_grid(n);
  hipLaunchKernelGGL(lz_dots_kernel<T>, dim3(G, (SA_MAX_NCV + SA_RED_COLS - 1) / SA_RED_COLS), dim3(SA_BLOCK), 0,
                     stream, work, n, (const T*)V, (const T*)W);
  rc = check_launch("lanczos_dots");
  if (rc) return rc;
  hipLaunchKernelGGL(lz_orth_kernel<T>, dim3(G), dim3(SA_BLOCK), 0, stream, work, n, s, V, Z, (const T*)W);
  return check_launch("lanczos_orth");
}

template <typename T>
static int lanczos_ritz(void* work, int64_t n, const T* V, const T* y_last, T* y, int32_t only_if,
                        hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && V && y, "bad arguments");
  hipLaunchKernelGGL(lz_ritz_kernel<T>, dim3(red_grid(n)), dim3(SA_BLOCK), 0, stream, work, n, V, y_last, y,
                     only_if);
  return check_launch("lanczos_ritz");
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_lanczos_begin(void* work, int32_t ncv, int32_t maxiter, double tol, double bound, double pred_tol,
                     hipStream_t stream) {
  SA_REQUIRE(work && ncv >= 2 && ncv <= SA_MAX_NCV && maxiter >= 1, "bad arguments (2 <= ncv <= 128)");
  hipLaunchKernelGGL(lz_begin_kernel, dim3(1), dim3(64), 0, stream, work, ncv, maxiter, tol, bound, pred_tol);
  return check_launch("lanczos_begin");
}

int sa_lanczos_start_f32(void* work, int64_t n, const float* s, const float* x, const float* d, double pexp,
                         double cx, const float* y, double cy, float* V, float* Z, int32_t gate, int32_t only_if,
                         hipStream_t stream) {
  return lanczos_start<float>(work, n, s, x, d, pexp, cx, y, cy, V, Z, gate, only_if, stream);
}

int sa_lanczos_start_f64(void* work, int64_t n, const double* s, const double* x, const double* d, double pexp,
                         double cx, const double* y, double cy, double* V, double* Z, int32_t gate, int32_t only_if,
                         hipStream_t stream) {
  return lanczos_start<double>(work, n, s, x, d, pexp, cx, y, cy, V, Z, gate, only_if, stream);
}

int sa_lanczos_step_f32(const sa_csr* A, void* work, const float* s, float* V, float* Z, float* W,
                        hipStream_t stream) {
  return lanczos_step<float>(A, work, s, V, Z, W, stream);
}

int sa_lanczos_step_f64(const sa_csr* A, void* work, const double* s, double* V, double* Z, double* W,
                        hipStream_t stream) {
  return lanczos_step<double>(A, work, s, V, Z, W, stream);
}

int sa_lanczos_ritz_f32(void* work, int64_t n, const float* V, const float* y_last, float* y, int32_t only_if,
                        hipStream_t stream) {
  return lanczos_ritz<float>(work, n, V, y_last, y, only_if, stream);
}

int sa_lanczos_ritz_f64(void* work, int64_t n, const double* V, const double* y_last, double* y, int32_t only_if,
                        hipStream_t stream) {
  return lanczos_ritz<double>(work, n, V, y_last, y, only_if, stream);
}

}  // extern "C"
