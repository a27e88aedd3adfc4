// Deflated MINRES on gfx950: the eigenvector tangent dy/dalpha and its adjoint.
//
// Solves P (N - lambda I) P x = P r, P = I - y y^T, by the MINRES recurrence of Paige & Saunders
// (1975): a Lanczos process on the deflated operator plus Givens rotations of its tridiagonal, the
// iterate x updated along the search directions w. Vectors are FP32 or FP64 (the operator's
// precision); every scalar is FP64 and stays on the device, so iterations are queued without host
// round trips. One iteration is five launches:
//   1. the fused scaled SpMV  q = N r2 (gathering z = s (.) r2, written by the previous update);
//   2. one pass over q, r2, r1, y for five dots (two-phase reduction);
//   3. a finisher: g = y . (q - lambda r2) (the deflated component), alfa;
//   4. one pass: the pending update of x by the previous iteration's direction, then the next raw
//      Lanczos vector r_new = (q - lambda r2 - g y) / beta - cr1 r1 - (alfa / beta) r2 (written over
//      r1), z = s (.) r_new, and ||r_new||^2 partials;
//   5. a finisher: beta_new, the Givens rotation, the stopping test phibar <= tol beta1.
// The two raw Lanczos vectors alternate between two buffers by the parity of the iteration count,
// and the three directions rotate modulo 3, both read from the device state, so a launch skipped
// after the solve has stopped leaves the roles consistent.
#include "sa_reduce.hpp"
#include "sa_spmv.hpp"

namespace sa {

__device__ inline sa_minres_state* mr_of(void* work) { return reinterpret_cast<sa_minres_state*>(work); }

template <typename T>
struct MrBufs {
  T* r[2];
  T* w[3];
  T* x;
  T* z;
  T* q;
};

template <typename T>
static MrBufs<T> bufs_of(const sa_minres_bufs* B) {
  MrBufs<T> b;
  for (int i = 0; i < 2; ++i) b.r[i] = reinterpret_cast<T*>(B->r[i]);
  for (int i = 0; i < 3; ++i) b.w[i] = reinterpret_cast<T*>(B->w[i]);
  b.x = reinterpret_cast<T*>(B->x);
  b.z = reinterpret_cast<T*>(B->z);
  b.q = reinterpret_cast<T*>(B->q);
  return b;
}

#define SA_GRID_STRIDE(r, n) \
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < (n); r += (int64_t)gridDim.x * blockDim.x)

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void tangent_rhs_kernel(int64_t n, const T* __restrict__ y,
                                                               const T* __restrict__ d, double clamp,
                                                               T* __restrict__ u) {
  SA_GRID_STRIDE(r, n) {
    const double dr = fmax((double)d[r], clamp);
    u[r] = (T)(0.5 * log(dr) * (double)y[r]);
  }
}

// ---- begin --------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void mr_begin_dots_kernel(void* work, int64_t n, const T* __restrict__ y,
                                                                 const T* __restrict__ r) {
  double acc[1] = {0.0};
  SA_GRID_STRIDE(i, n) acc[0] += (double)y[i] * (double)r[i];
  block_partials<1, RED_SUM>(acc, work);
}

__global__ __launch_bounds__(SA_BLOCK) void mr_begin_finish1_kernel(void* work, int G, double shift, double tol,
                                                                    int32_t maxit) {
  sa_minres_state* st = mr_of(work);
  const double c0 = sum_partials(partials_of(work), G, 1, 0);
  if (threadIdx.x == 0) {
    st->status = SA_RUNNING;
    st->itn = 0;
    st->maxit = maxit;
    st->pending = 0;
    st->shift = shift;
    st->tol = tol;
    st->c0 = c0;
    st->rnorm = NAN;
  }
}

// b = r - (y . r) y into r[0]; z = s (.) b; x = w = 0; ||b||^2 partials.
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void mr_begin_vectors_kernel(void* work, int64_t n, const T* __restrict__ s,
                                                                    const T* __restrict__ y,
                                                                    const T* __restrict__ r, MrBufs<T> B) {
  const double c0 = mr_of(work)->c0;
  double acc[1] = {0.0};
  SA_GRID_STRIDE(i, n) {
    const T b = (T)((double)r[i] - c0 * (double)y[i]);
    B.r[0][i] = b;
    B.z[i] = s ? s[i] * b : b;
    B.x[i] = (T)0;
    B.w[0][i] = (T)0;
    B.w[1][i] = (T)0;
    B.w[2][i] = (T)0;
    acc[0] += (double)b * (double)b;
  }
  block_partials<1, RED_SUM>(acc, work);
}

__global__ __launch_bounds__(SA_BLOCK) void mr_begin_finish2_kernel(void* work, int G) {
  sa_minres_state* st = mr_of(work);
  const double bb = sum_partials(partials_of(work), G, 1, 0);
  if (threadIdx.x == 0) {
    const double beta1 = sqrt(bb);
    st->beta1 = beta1;
    st->beta = beta1;
    st->oldb = 0.0;
    st->cr1 = 0.0;
    st->alfa = 0.0;
    st->g = 0.0;
    st->dbar = 0.0;
    st->epsln = 0.0;
    st->cs = -1.0;
    st->sn = 0.0;
    st->phibar = beta1;
    st->rnorm = beta1;
    if (!isfinite(beta1)) st->status = SA_BROKEN;
    else if (beta1 == 0.0 || st->maxit <= 0) st->status = SA_CONVERGED;   // x = 0 solves it
  }
}

// ---- one iteration ------------------------------------------------------------------------------
// Dots of the projection pass: [r2.q, r2.r2, y.q, y.r2, r2.r1].
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void mr_dots_kernel(void* work, int64_t n, const T* __restrict__ y,
                                                           MrBufs<T> B) {
  const sa_minres_state* st = mr_of(work);
  if (st->status != SA_RUNNING) return;
  const int par = st->itn & 1;
  const T* __restrict__ r2 = B.r[par];
  const T* __restrict__ r1 = B.r[par ^ 1];
  const bool use_r1 = st->cr1 != 0.0;
  const T* __restrict__ q = B.q;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  SA_GRID_STRIDE(i, n) {
    const double a = (double)r2[i], qi = (double)q[i], yi = (double)y[i];
    acc[0] += a * qi;
    acc[1] += a * a;
    acc[2] += yi * qi;
    acc[3] += yi * a;
    if (use_r1) acc[4] += a * (double)r1[i];
  }
  block_partials<5, RED_SUM>(acc, work);
}

__global__ __launch_bounds__(SA_BLOCK) void mr_dots_finish_kernel(void* work, int G) {
  sa_minres_state* st = mr_of(work);
  if (st->status != SA_RUNNING) return;
  const double* part = partials_of(work);
  double d[5];
  for (int c = 0; c < 5; ++c) d[c] = sum_partials(part, G, 5, c);
  if (threadIdx.x == 0) {
    const double lam = st->shift, beta = st->beta;
    // A v with v = r2 / beta: ((q - lam r2) - g y) / beta, g = y . (q - lam r2)
    const double g = d[2] - lam * d[3];
    const double vAv = (d[0] - lam * d[1] - g * d[3]) / (beta * beta);
    st->g = g;
    st->alfa = vAv - st->cr1 * d[4] / beta;
  }
}

// Pending x update of the previous iteration k: w_k = (v_k - oldeps w_{k-2} - delta w_{k-1}) denom,
// x += phi w_k, with v_k = r1 / oldb; then r_new over r1, z = s (.) r_new, ||r_new||^2 partials.
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void mr_update_kernel(void* work, int64_t n, const T* __restrict__ s,
                                                             const T* __restrict__ y, MrBufs<T> B) {
  const sa_minres_state* st = mr_of(work);
  if (st->status != SA_RUNNING) return;
  const int itn = st->itn;
  const int par = itn & 1;
  const T* __restrict__ r2 = B.r[par];
  T* __restrict__ r1 = B.r[par ^ 1];
  const bool pend = st->pending != 0;
  const int k = itn - 1;
  T* __restrict__ wk = B.w[((k % 3) + 3) % 3];
  const T* __restrict__ wk2 = B.w[((k + 1) % 3 + 3) % 3];
  const T* __restrict__ wk1 = B.w[((k + 2) % 3 + 3) % 3];
  const double ioldb = st->oldb > 0.0 ? 1.0 / st->oldb : 0.0;
  const double oe = st->pe_oldeps, de = st->pe_delta, dn = st->pe_denom, ph = st->pe_phi;
  const double lam = st->shift, beta = st->beta, ib = 1.0 / beta, g = st->g, cr1 = st->cr1, alfa = st->alfa;
  const bool use_r1 = cr1 != 0.0 || pend;
  double acc[1] = {0.0};
  SA_GRID_STRIDE(i, n) {
    const double a1 = use_r1 ? (double)r1[i] : 0.0;
    if (pend) {
      const T w = (T)((a1 * ioldb - oe * (double)wk2[i] - de * (double)wk1[i]) * dn);
      wk[i] = w;
      B.x[i] = (T)((double)B.x[i] + ph * (double)w);
    }
    const double a2 = (double)r2[i];
    const T rn = (T)(((double)B.q[i] - lam * a2 - g * (double)y[i]) * ib - cr1 * a1 - alfa * ib * a2);
    r1[i] = rn;
    B.z[i] = s ? s[i] * rn : rn;
    acc[0] += (double)rn * (double)rn;
  }
  block_partials<1, RED_SUM>(acc, work);
}

__global__ __launch_bounds__(SA_BLOCK) void mr_update_finish_kernel(void* work, int G) {
  sa_minres_state* st = mr_of(work);
  if (st->status != SA_RUNNING) return;
  const double bb = sum_partials(partials_of(work), G, 1, 0);
  if (threadIdx.x != 0) return;
  const double beta_new = sqrt(bb);
  const double oldeps = st->epsln;
  const double delta = st->cs * st->dbar + st->sn * st->alfa;
  const double gbar = st->sn * st->dbar - st->cs * st->alfa;
  st->epsln = st->sn * beta_new;
  st->dbar = -st->cs * beta_new;
  double gamma = sqrt(gbar * gbar + beta_new * beta_new);
  gamma = fmax(gamma, 2.220446049250313e-16);
  const double cs = gbar / gamma, sn = beta_new / gamma;
  const double phi = cs * st->phibar;
  st->cs = cs;
  st->sn = sn;
  st->phibar = sn * st->phibar;
  st->pe_oldeps = oldeps;
  st->pe_delta = delta;
  st->pe_denom = 1.0 / gamma;
  st->pe_phi = phi;
  st->pending = 1;
  st->oldb = st->beta;
  st->beta = beta_new;
  st->cr1 = beta_new / st->oldb;
  st->itn += 1;
  st->rnorm = st->phibar;
  int status = SA_RUNNING;
  if (!isfinite(beta_new) || !isfinite(phi) || !isfinite(st->phibar)) status = SA_BROKEN;
  else if (st->phibar <= st->tol * st->beta1 || beta_new == 0.0) status = SA_CONVERGED;
  else if (st->itn >= st->maxit) status = SA_MAXITER;
  st->status = status;
}

// ---- end ----------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void mr_end_kernel(void* work, int64_t n, const T* __restrict__ y,
                                                          MrBufs<T> B) {
  const sa_minres_state* st = mr_of(work);
  const bool pend = st->pending != 0 && st->status != SA_BROKEN;
  const int itn = st->itn;
  const int par = itn & 1;
  const T* __restrict__ r1 = B.r[par ^ 1];
  const int k = itn - 1;
  T* __restrict__ wk = B.w[((k % 3) + 3) % 3];
  const T* __restrict__ wk2 = B.w[((k + 1) % 3 + 3) % 3];
  const T* __restrict__ wk1 = B.w[((k + 2) % 3 + 3) % 3];
  const double ioldb = st->oldb > 0.0 ? 1.0 / st->oldb : 0.0;
  const double oe = st->pe_oldeps, de = st->pe_delta, dn = st->pe_denom, ph = st->pe_phi;
  double acc[1] = {0.0};
  SA_GRID_STRIDE(i, n) {
    double x = (double)B.x[i];
    if (pend) {
      const T w = (T)(((double)r1[i] * ioldb - oe * (double)wk2[i] - de * (double)wk1[i]) * dn);
      wk[i] = w;
      x += ph * (double)w;
      B.x[i] = (T)x;
      x = (double)(T)x;
    }
    acc[0] += (double)y[i] * x;
  }
  block_partials<1, RED_SUM>(acc, work);
}

__global__ __launch_bounds__(SA_BLOCK) void mr_end_finish_kernel(void* work, int G) {
  sa_minres_state* st = mr_of(work);
  const double c = sum_partials(partials_of(work), G, 1, 0);
  if (threadIdx.x == 0) {
    st->c0 = c;
    st->pending = 0;
  }
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void mr_deflate_kernel(void* work, int64_t n, const T* __restrict__ y,
                                                              T* __restrict__ x) {
  const double c = mr_of(work)->c0;
  SA_GRID_STRIDE(i, n) x[i] = (T)((double)x[i] - c * (double)y[i]);
}

// ---- launchers ----------------------------------------------------------------------------------
static inline int mr_grid(int64_t n) { return grid_for(n, SA_BLOCK * 4, SA_RED_GRID); }

static bool bufs_ok(const sa_minres_bufs* B) {
  return B && B->r[0] && B->r[1] && B->w[0] && B->w[1] && B->w[2] && B->x && B->z && B->q;
}

template <typename T>
static int tangent_rhs(int64_t n, const T* y, const T* d, double clamp, T* u, hipStream_t stream) {
  SA_REQUIRE(n > 0 && y && d && u && clamp > 0.0, "bad arguments");
  hipLaunchKernelGGL(tangent_rhs_kernel<T>, dim3(grid_for(n, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK), 0, stream, n, y,
                     d, clamp, u);
  return check_launch("tangent_rhs");
}

template <typename T>
static int minres_begin(void* work, int64_t n, const T* s, const T* y, const T* r, double shift, double tol,
                        int32_t maxit, const sa_minres_bufs* Bp, hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && y && r && bufs_ok(Bp) && tol >= 0.0, "bad arguments");
  const MrBufs<T> B = bufs_of<T>(Bp);
  const int G = mr_grid(n);
  hipLaunchKernelGGL(mr_begin_dots_kernel<T>, dim3(G), dim3(SA_BLOCK), 0, stream, work, n, y, r);
  int rc = check_launch("minres_begin_dots");
  if (rc) return rc;
  hipLaunchKernelGGL(mr_begin_finish1_kernel, dim3(1), dim3(SA_BLOCK), 0, stream, work, G, shift, tol, maxit);
  rc = check_launch("minres_begin_finish1");
  if (rc) return rc;
  hipLaunchKernelGGL(mr_begin_vectors_kernel<T>, dim3(G), dim3(SA_BLOCK), 0, stream, work, n, s, y, r, B);
  rc = check_launch("minres_begin_vectors");
  if (rc) return rc;
  hipLaunchKernelGGL(mr_begin_finish2_kernel, dim3(1), dim3(SA_BLOCK), 0, stream, work, G);
  return check_launch("minres_begin_finish2");
}

template <typename T>
static int minres_iterate(const sa_csr* A, void* work, const T* s, const T* y, int32_t count,
                          const sa_minres_bufs* Bp, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->tile_rows && work && y && bufs_ok(Bp) && count >= 0, "bad arguments");
  const int64_t n = A->n;
  const MrBufs<T> B = bufs_of<T>(Bp);
  const int G = mr_grid(n);
  const int32_t* gate = &reinterpret_cast<const sa_minres_state*>(work)->status;
  for (int32_t it = 0; it < count; ++it) {
    int rc = spmv_apply<T>(A, s, B.z, 0.0, nullptr, 1.0, B.q, gate, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(mr_dots_kernel<T>, dim3(G), dim3(SA_BLOCK), 0, stream, work, n, y, B);
    rc = check_launch("minres_dots");
    if (rc) return rc;
    hipLaunchKernelGGL(mr_dots_finish_kernel, dim3(1), dim3(SA_BLOCK), 0, stream, work, G);
    rc = check_launch("minres_dots_finish");
    if (rc) return rc;
    hipLaunchKernelGGL(mr_update_kernel<T>, dim3(G), dim3(SA_BLOCK), 0, stream, work, n, s, y, B);
    rc = check_launch("minres_update");
    if (rc) return rc;
    hipLaunchKernelGGL(mr_update_finish_kernel, dim3(1), dim3(SA_BLOCK), 0, stream, work, G);
    rc = check_launch("minres_update_finish");
    if (rc) return rc;
  }
  return SA_OK;
}

template <typename T>
static int minres_end(void* work, int64_t n, const T* y, const sa_minres_bufs* Bp, hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && y && bufs_ok(Bp), "bad arguments");
  const MrBufs<T> B = bufs_of<T>(Bp);
  const int G = mr_grid(n);
  hipLaunchKernelGGL(mr_end_kernel<T>, dim3(G), dim3(SA_BLOCK), 0, stream, work, n, y, B);
  int rc = check_launch("minres_end");
  if (rc) return rc;
  hipLaunchKernelGGL(mr_end_finish_kernel, dim3(1), dim3(SA_BLOCK), 0, stream, work, G);
  rc = check_launch("minres_end_finish");
  if (rc) return rc;
  hipLaunchKernelGGL(mr_deflate_kernel<T>, dim3(grid_for(n, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK), 0, stream, work, n,
                     y, B.x);
  return check_launch("minres_deflate");
}

}  // namespace sa

using namespace sa;

extern "C" {

size_t sa_minres_state_bytes(void) { return sizeof(sa_minres_state); }

int sa_tangent_rhs_f32(int64_t n, const float* y, const float* d, double clamp, float* u, hipStream_t stream) {
  return tangent_rhs<float>(n, y, d, clamp, u, stream);
}

int sa_tangent_rhs_f64(int64_t n, const double* y, const double* d, double clamp, double* u, hipStream_t stream) {
  return tangent_rhs<double>(n, y, d, clamp, u, stream);
}

int sa_minres_begin_f32(void* work, int64_t n, const float* s, const float* y, const float* r, double shift,
                        double tol, int32_t maxit, const sa_minres_bufs* B, hipStream_t stream) {
  return minres_begin<float>(work, n, s, y, r, shift, tol, maxit, B, stream);
}

int sa_minres_begin_f64(void* work, int64_t n, const double* s, const double* y, const double* r, double shift,
                        double tol, int32_t maxit, const sa_minres_bufs* B, hipStream_t stream) {
  return minres_begin<double>(work, n, s, y, r, shift, tol, maxit, B, stream);
}

int sa_minres_iterate_f32(const sa_csr* A, void* work, const float* s, const float* y, int32_t count,
                          const sa_minres_bufs* B, hipStream_t stream) {
  return minres_iterate<float>(A, work, s, y, count, B, stream);
}

int sa_minres_iterate_f64(const sa_csr* A, void* work, const double* s, const double* y, int32_t count,
                          const sa_minres_bufs* B, hipStream_t stream) {
  return minres_iterate<double>(A, work, s, y, count, B, stream);
}

int sa_minres_end_f32(void* work, int64_t n, const float* y, const sa_minres_bufs* B, hipStream_t stream) {
  return minres_end<float>(work, n, y, B, stream);
}

int sa_minres_end_f64(void* work, int64_t n, const double* y, const sa_minres_bufs* B, hipStream_t stream) {
  return minres_end<double>(work, n, y, B, stream);
}

}  // extern "C"
