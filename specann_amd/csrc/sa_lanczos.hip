// Warm-started Lanczos for the leading eigenpair of N_alpha = s (.) (J/c) (.) s on gfx950.
//
// The recurrence runs on N itself, not on the reference's shifted operator S_rho = rho I - N: the
// Krylov spaces of the two are identical (T_S = rho I - T_N), and without the shift the FP32 basis
// keeps its orthogonality (with rho ~ 1 dominating the product, S_rho v loses the N v part to
// cancellation). The caller reports mu = rho - theta, the eigenvalue the reference's eigsh returns.
//
// One step = three launches:
//   1. the fused scaled SpMV  W = sigma_j s (.) (J Z),  Z = s (.) V[j] written by the previous step;
//   2. projections h_m = v_m . W for every basis vector of the cycle (FP64 partial sums, up to 12
//      shadow rows and the two newest full-precision rows per pass over W, then a one-workgroup
//      finisher that sums the partials in workgroup order);
//   3. classical Gram-Schmidt W - sum h_m v_m into V[j+1] (raw; sigma_{j+1} = 1/beta_j), Z for the
//      next step, beta_j = ||.||, then, in the last workgroup, the FP64 Ritz solve of the projected
//      matrix: largest eigenvalue by 257-section Sturm bisection over the whole workgroup, the last
//      component s_j of its eigenvector from the derivative of the last LDL^T pivot, and the
//      residual estimate beta_j |s_j|, the stopping test the reference's eigsh applies per restart
//      (here it is applied every step). A small or non-finite estimate, and every stop, is decided
//      instead by the full eigenvector (inverse iteration on the dense projected matrix, pivoted LU
//      shared by the workgroup), which also gives the Ritz vector's coefficients.
// Every launch reads the status in the work buffer first and does nothing unless the solve is
// running, so the host may queue several steps (and the restart pair) without a round trip.
#include <cstddef>

#include "sa_reduce.hpp"
#include "sa_ritz.hpp"
#include "sa_spmv.hpp"

namespace sa {

// The shadow basis: a 16-bit (FP32 form) or 32-bit (FP64 form) copy of every basis row, read only to
// measure the projections on rows older than j - 1 (which decide and correct components of order
// sqrt(eps); the copy's rounding, ~1e-7 of ||W|| for FP16, is far below that). Row m holds the unit
// vector sigma_m v_m times shscale = sqrt(n), so that its entries sit well inside the FP16 range; it is
// written by the projections of step m (which read v_m in full precision, with sigma_m known) and by
// the thick restart for the rows it rotates.
template <typename T>
struct Shadow;
template <>
struct Shadow<float> {
  typedef _Float16 type;
  typedef _Float16 vec __attribute__((ext_vector_type(4)));    // the L = 4 entries of a 16-byte load of V
  typedef _Float16 vec16 __attribute__((ext_vector_type(8)));  // one 16-byte load of the shadow
};
template <>
struct Shadow<double> {
  typedef float type;
  typedef float vec __attribute__((ext_vector_type(2)));
  typedef float vec16 __attribute__((ext_vector_type(4)));
};
template <typename T>
using SH = typename Shadow<T>::type;

template <class F>
__global__ void lz_begin_kernel(F fin, int32_t ncv, int32_t maxiter, double tol, double bound, const double* bounds,
                                double pred_tol, int32_t keep, double orth_tau) {
  if (threadIdx.x != 0 || !fin.live()) return;
  sa_state* st = state_of(fin.w());
  if (bounds) bound = bounds[fin.idx()];
  st->status = SA_RUNNING;
  st->j = 0;
  st->steps = 0;
  st->cycles = 0;
  st->ncv = ncv;
  st->maxiter = maxiter;
  st->gate = 0;
  st->kept = 0;
  st->keep = keep;
  st->theta = 0.0;
  st->res = INFINITY;
  st->first_beta = NAN;
  st->tol = tol;
  st->bound = bound;
  st->pred_tol = pred_tol;
  st->orth_tau = orth_tau;
  st->orth_lo = 0;
  st->counters[0] = 0;
  st->counters[1] = 0;
}

template <typename T, class M>
__global__ __launch_bounds__(SA_BLOCK) void lz_start_kernel(M map, const T* __restrict__ s, const T* __restrict__ x,
                                                            const T* __restrict__ d, T pexp,
                                                            const double* __restrict__ pexp_each, double cx,
                                                            const T* __restrict__ y, double cy, T* __restrict__ V,
                                                            T* __restrict__ Z, int32_t gate, int32_t only_if) {
  const Slice sl = map.slice();
  if (!sl.live) return;
  const sa_state* st = state_of(sl.work);
  if (only_if >= 0 && st->status != only_if) return;
  if (pexp_each) pexp = (T)pexp_each[sl.inst];  // a batch whose instances sit at different alphas
  double acc[1] = {0.0};
  const T clampv = (T)1e-12;
  for (int64_t r = sl.lo + threadIdx.x; r < sl.hi; r += blockDim.x) {
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
  block_partials<1, RED_SUM>(acc, sl.work, sl.blk);
}

// ||x0||, sigma_0 and the fresh state of a started solve (phase 2 of the norm of lz_start).
template <class F>
__global__ __launch_bounds__(SA_BLOCK) void lz_start_finish_kernel(F fin, int32_t gate, int32_t only_if) {
  if (!fin.live()) return;
  sa_state* st = state_of(fin.w());
  if (only_if >= 0 && st->status != only_if) return;
  const double tot = sum_partials(partials_of(fin.w()), fin.g(), 1, 0);
  if (threadIdx.x == 0) {
    const double nrm = sqrt(tot);
    st->r[3] = nrm;
    const bool ok = nrm > 0.0 && isfinite(nrm);
    st->sigma[0] = ok ? 1.0 / nrm : 0.0;
    st->j = 0;
    st->kept = 0;
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
                                                              const T* __restrict__ partial, int64_t pld,
                                                              const T* __restrict__ s, T* __restrict__ W) {
  const sa_state* st = state_of(work);
  if (st->status != SA_RUNNING) return;
  const double sigma = st->sigma[st->j];
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int p = 0; p < npanels; ++p) acc += (double)partial[(int64_t)p * pld + r];
    W[r] = (T)(sigma * (s ? (double)s[r] : 1.0) * acc);
  }
}

// h_m = sigma_m v_m . W for every basis vector of the cycle. Each thread takes 2 L consecutive entries
// per step (one 16-byte shadow load, two 16-byte loads of W and of a full-precision row). The shadow
// rows below mh go in passes of kDotsShRows rows over the workgroup's rows, the first pass also
// carrying W's norm and the full-precision rows j - 1 and j, so W is read once per pass and a pass has
// all its rows' loads in flight; without the shadow the rows go SA_RED_COLS per pass in full precision.
// Per-workgroup partials are stored column-major (part[m * G + b]); column j + 1 takes ||W||^2. The
// finisher sums them in workgroup order. The pass over row j also writes its shadow row.
// With npanels > 0 the workgroup first forms W on its own rows from the SpMV's panel partials (the
// panel combine: sigma_j s_r times their FP64 sum in panel order, as lz_combine_kernel computes it)
// and writes it for the update kernel; partial rows and W are then 16-byte aligned.
// Shadow rows per pass of the projections: twelve 16-byte shadow loads plus W and two full rows per
// lane keep the kernel at two waves per SIMD without spills (sixteen spill, 32 do not fit).
constexpr int kDotsShRows = 12;
constexpr int kDotsWLds = 1024;  // units of 2 L entries of W kept in LDS between passes (32 KiB)
static_assert(kDotsShRows >= SA_RED_COLS, "a pass without the shadow takes SA_RED_COLS rows");

template <typename T, class M>
__global__ __launch_bounds__(SA_BLOCK, 2) void lz_dots_kernel(M map, const T* __restrict__ V, int64_t ldv,
                                                           SH<T>* __restrict__ Vh, double shscale, T* __restrict__ W,
                                                           const T* __restrict__ partial, int64_t pld, int npanels,
                                                           const T* __restrict__ s) {
  using VT = typename vec16<T>::type;
  constexpr int L = vec16<T>::L;
  __shared__ double red[2][SA_WAVES_PER_BLOCK][kDotsShRows + 3];
  const Slice sl = map.slice();
  const sa_state* st = state_of(sl.work);
  if (st->status != SA_RUNNING) return;
  const int j = st->j;
  // rows below mh come from the shadow (the first step of a cycle measures the kept rows, whose
  // projections are the arrow couplings, in full precision)
  const int mh = (Vh && j > st->kept) ? j - 1 : 0;
  const T put = (T)(st->sigma[j] * shscale);  // v_j -> its shadow row
  double* part = partials_of(sl.work);
  const int G = sl.G;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t lo = sl.lo, hi = sl.hi;
  const int64_t nv = hi > lo ? (hi - lo) / L : 0;
  if (npanels > 0) {
    const double sigma = st->sigma[j];
    for (int64_t g = threadIdx.x; g < nv; g += SA_BLOCK) {
      const int64_t r = lo + g * L;
      double acc[L];
#pragma unroll
      for (int l = 0; l < L; ++l) acc[l] = 0.0;
      for (int p0 = 0; p0 < npanels; p0 += 8) {  // npanels is a multiple of 8
        VT pv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) pv[q] = *reinterpret_cast<const VT*>(&partial[(int64_t)(p0 + q) * pld + r]);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
#pragma unroll
          for (int l = 0; l < L; ++l) acc[l] += (double)pv[q][l];
        }
      }
      VT w;
      if (s) {
        const VT sv = *reinterpret_cast<const VT*>(&s[r]);
#pragma unroll
        for (int l = 0; l < L; ++l) w[l] = (T)(sigma * (double)sv[l] * acc[l]);
      } else {
#pragma unroll
        for (int l = 0; l < L; ++l) w[l] = (T)(sigma * 1.0 * acc[l]);
      }
      *reinterpret_cast<VT*>(&W[r]) = w;
    }
    for (int64_t r = lo + nv * L + threadIdx.x; r < hi; r += SA_BLOCK) {
      double acc = 0.0;
      for (int q = 0; q < npanels; ++q) acc += (double)partial[(int64_t)q * pld + r];
      W[r] = (T)(sigma * (s ? (double)s[r] : 1.0) * acc);
    }
    __syncthreads();  // the chunks below read rows other threads of this workgroup wrote
  }
  // One pass over the workgroup's rows per group of basis rows, W loaded once per pass. With the
  // shadow on (mh > 0) a group is up to kDotsShRows shadow rows below mh, and the first pass also
  // carries the full-precision rows j - 1 and j (mh = j - 1); without it a group is SA_RED_COLS
  // full-precision rows. Every load of a group is unconditional (rows past the group clamped to its
  // last row, their sums dropped), so the whole group's loads are in flight together.
  constexpr int E = 2 * L;
  using S16 = typename Shadow<T>::vec16;
  const int64_t ne = hi > lo ? (hi - lo) / E : 0;
  const int64_t tail = lo + ne * E;
  const bool sh = mh > 0;
  const int gend = sh ? mh : j + 1;  // the grouped rows are [0, gend); with the shadow, rows mh..j ride along
  double ww = 0.0;                   // ||W||^2, accumulated in the first pass
  // The workgroup's slice of W stays in LDS after the first pass when it fits (every thread reads back
  // its own units, so no barrier), and later passes stream only basis rows.
  __shared__ VT wl[kDotsWLds][2];
  const bool wlds = ne <= kDotsWLds;
  int pass = 0;
  for (int p0 = 0; p0 < gend; p0 += sh ? kDotsShRows : SA_RED_COLS, ++pass) {
    const int cnt = min(sh ? kDotsShRows : SA_RED_COLS, gend - p0);
    const bool extra = sh && pass == 0;
    double acc[kDotsShRows];
#pragma unroll
    for (int q = 0; q < kDotsShRows; ++q) acc[q] = 0.0;
    double f0 = 0.0, f1 = 0.0;  // rows j - 1 and j (extra)
    for (int64_t u = threadIdx.x; u < ne; u += SA_BLOCK) {
      const int64_t r = lo + u * E;
      VT w0, w1;
      if (wlds && pass > 0) {
        w0 = wl[u][0];
        w1 = wl[u][1];
      } else {
        w0 = *reinterpret_cast<const VT*>(&W[r]);
        w1 = *reinterpret_cast<const VT*>(&W[r + L]);
        if (wlds) {
          wl[u][0] = w0;
          wl[u][1] = w1;
        }
      }
      double wd[E];
#pragma unroll
      for (int l = 0; l < L; ++l) {
        wd[l] = (double)w0[l];
        wd[L + l] = (double)w1[l];
      }
      if (pass == 0) {
#pragma unroll
        for (int e = 0; e < E; ++e) ww += wd[e] * wd[e];
      }
      if (sh) {
        S16 h[kDotsShRows];
#pragma unroll
        for (int q = 0; q < kDotsShRows; ++q)
          h[q] = *reinterpret_cast<const S16*>(&Vh[(int64_t)(p0 + min(q, cnt - 1)) * ldv + r]);
        VT a0, a1, b0, b1;
        if (extra) {
          a0 = *reinterpret_cast<const VT*>(&V[(int64_t)(j - 1) * ldv + r]);
          a1 = *reinterpret_cast<const VT*>(&V[(int64_t)(j - 1) * ldv + r + L]);
          b0 = *reinterpret_cast<const VT*>(&V[(int64_t)j * ldv + r]);
          b1 = *reinterpret_cast<const VT*>(&V[(int64_t)j * ldv + r + L]);
        }
#pragma unroll
        for (int q = 0; q < kDotsShRows; ++q) {
          if (q < cnt) {
#pragma unroll
            for (int e = 0; e < E; ++e) acc[q] += (double)h[q][e] * wd[e];
          }
        }
        if (extra) {
          S16 hj;
#pragma unroll
          for (int l = 0; l < L; ++l) {
            f0 += (double)a0[l] * wd[l] + (double)a1[l] * wd[L + l];
            f1 += (double)b0[l] * wd[l] + (double)b1[l] * wd[L + l];
            hj[l] = (SH<T>)(b0[l] * put);
            hj[L + l] = (SH<T>)(b1[l] * put);
          }
          *reinterpret_cast<S16*>(&Vh[(int64_t)j * ldv + r]) = hj;  // the shadow row of v_j
        }
      } else {
        VT a[SA_RED_COLS][2];
#pragma unroll
        for (int q = 0; q < SA_RED_COLS; ++q) {
          const int64_t at = (int64_t)(p0 + min(q, cnt - 1)) * ldv + r;
          a[q][0] = *reinterpret_cast<const VT*>(&V[at]);
          a[q][1] = *reinterpret_cast<const VT*>(&V[at + L]);
        }
#pragma unroll
        for (int q = 0; q < SA_RED_COLS; ++q) {
          if (q < cnt) {
#pragma unroll
            for (int l = 0; l < L; ++l) acc[q] += (double)a[q][0][l] * wd[l] + (double)a[q][1][l] * wd[L + l];
            if (Vh && p0 + q == j) {
              S16 hj;
#pragma unroll
              for (int l = 0; l < L; ++l) {
                hj[l] = (SH<T>)(a[q][0][l] * put);
                hj[L + l] = (SH<T>)(a[q][1][l] * put);
              }
              *reinterpret_cast<S16*>(&Vh[(int64_t)j * ldv + r]) = hj;
            }
          }
        }
      }
    }
    for (int64_t r = tail + threadIdx.x; r < hi; r += SA_BLOCK) {  // rows past the last whole unit
      const double w = (double)W[r];
      if (pass == 0) ww += w * w;
#pragma unroll
      for (int q = 0; q < kDotsShRows; ++q) {
        if (q < cnt) {
          const int64_t at = (int64_t)(p0 + q) * ldv + r;
          acc[q] += (sh ? (double)Vh[at] : (double)V[at]) * w;
          if (!sh && Vh && p0 + q == j) Vh[at] = (SH<T>)(V[at] * put);
        }
      }
      if (extra) {
        f0 += (double)V[(int64_t)(j - 1) * ldv + r] * w;
        const T vj = V[(int64_t)j * ldv + r];
        f1 += (double)vj * w;
        Vh[(int64_t)j * ldv + r] = (SH<T>)(vj * put);
      }
    }
    // values: the group's rows, then (extra) rows j - 1 and j, then (first pass) ||W||^2
    double (*rd)[kDotsShRows + 3] = red[pass & 1];
#pragma unroll
    for (int q = 0; q < kDotsShRows; ++q) {
      if (q < cnt) {
        const double x = wave_sum(acc[q]);
        if (lane == 0) rd[wave][q] = x;
      }
    }
    if (extra) {
      const double x0 = wave_sum(f0), x1 = wave_sum(f1);
      if (lane == 0) {
        rd[wave][kDotsShRows] = x0;
        rd[wave][kDotsShRows + 1] = x1;
      }
    }
    if (pass == 0) {
      const double x = wave_sum(ww);
      if (lane == 0) rd[wave][kDotsShRows + 2] = x;
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < kDotsShRows + 3) {
      const bool row = t < cnt;
      const bool ext = extra && (t == kDotsShRows || t == kDotsShRows + 1);
      const bool wsq = pass == 0 && t == kDotsShRows + 2;
      if (row || ext || wsq) {
        double x = (rd[0][t] + rd[1][t]) + (rd[2][t] + rd[3][t]);
        int64_t col;
        if (row) {
          col = p0 + t;
          if (sh) x *= 1.0 / (st->sigma[p0 + t] * shscale);  // shadow row -> raw scale of v_m
        } else if (ext) {
          col = t == kDotsShRows ? j - 1 : j;
        } else {
          col = j + 1;
        }
        part[col * G + sl.blk] = x;
      }
    }
  }
}

// Phase 2 of the projections: column m summed over the G workgroups by a 32-lane group, eight loads
// in flight per lane, in workgroup order per lane and a fixed shuffle tree across lanes.
template <class F>
__global__ __launch_bounds__(SA_BLOCK) void lz_dots_finish_kernel(F fin) {
  sa_state* st = state_of(fin.w());
  if (st->status != SA_RUNNING) return;
  const int j = st->j;
  const int G = fin.g();
  const double ws = F::kScaleW ? st->sigma[j] : 1.0;  // batched: W was not scaled by sigma_j
  const double* part = partials_of(fin.w());
  const int lane32 = threadIdx.x & 31, grp = threadIdx.x >> 5;
  __shared__ double lww;
  __shared__ double lh[SA_MAX_NCV + 1];  // the projections again, for the decision below (LDS, not L2)
  for (int mb = 0; mb <= j + 1; mb += SA_BLOCK / 32) {
    const int m = mb + grp;
    double x = 0.0;
    if (m <= j + 1) {
      const double* pc = part + (int64_t)m * G;
      int b = lane32;
      for (; b + 32 * 7 < G; b += 32 * 8) {
        double q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = pc[b + 32 * u];
#pragma unroll
        for (int u = 0; u < 8; ++u) x += q[u];
      }
      for (; b < G; b += 32) x += pc[b];
    }
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) x += __shfl_xor(x, o, 32);
    if (lane32 == 0 && m <= j) {
      const double h = st->sigma[m] * ws * x;
      st->h[m] = h;
      lh[m] = h;
    }
    if (lane32 == 0 && m == j + 1) lww = ws * ws * x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    st->a[j] = lh[j];
    // Selective reorthogonalisation. The projections h_m of W on rows m < j - 1 vanish in exact
    // arithmetic past the first step of a cycle (the kept rows couple to row k only); left in, they
    // become components h_m / beta_j of the next basis vector. The update subtracts every row when one
    // of them exceeds tau beta_j (beta_j^2 ~ ||W||^2 - sum h_m^2), else only rows j - 1 and j, which
    // keeps the basis orthogonal to tau (semi-orthogonality, tau ~ sqrt(eps) of the vectors' type).
    double s2 = 0.0, hmax = 0.0;
    for (int m = 0; m <= j; ++m) {
      s2 += lh[m] * lh[m];
      if (m + 1 < j) hmax = fmax(hmax, fabs(lh[m]));
    }
    const double beta_est = sqrt(fmax(lww - s2, 1e-30 * lww));
    const double tau = st->orth_tau;
    const bool full = !(hmax <= tau * beta_est);  // NaN: full
    st->orth_lo = full ? 0 : max(0, j - 1);
    st->counters[full ? 0 : 1] += 1;
  }
}

// ---------------------------------------------------------------------------------------------
// The projected matrix of the current cycle. Rows 0..k-1 hold the Ritz values kept by the last thick
// restart (diagonal a[i]) and couple only to row k, through arrow[i]; rows k..m-1 are tridiagonal
// (a, b), b[i] coupling rows i and i+1. With k = 0 it is the plain Lanczos tridiagonal.
// ---------------------------------------------------------------------------------------------
struct Proj {
  int m, k;
  const double* a;
  const double* b;
  const double* ar;
  const double* b2;
  const double* ar2;
};

// Number of eigenvalues below x: the inertia of T - x I from its LDL^T factorisation in the row order
// 0..m-1 (the kept rows first; their Schur complement only touches the pivot of row k).
__device__ inline int sturm_count(const Proj& T, double x, double pivmin) {
  int cnt = 0;
  double acc = 0.0;
  for (int i = 0; i < T.k; ++i) {
    double q = T.a[i] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
    acc += T.ar2[i] / q;
  }
  double q = (T.a[T.k] - x) - acc;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0.0;
  for (int i = T.k + 1; i < T.m; ++i) {
    q = (T.a[i] - x) - T.b2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// Gershgorin radius of row i.
__device__ inline double gersh_radius(const Proj& T, int i) {
  if (i < T.k) return fabs(T.ar[i]);
  double r = 0.0;
  if (i == T.k) {
    for (int u = 0; u < T.k; ++u) r += fabs(T.ar[u]);
  } else {
    r += fabs(T.b[i - 1]);
  }
  if (i < T.m - 1) r += fabs(T.b[i]);
  return r;
}

// s_{m-1}^2 of the unit eigenvector of the eigenvalue x, from the derivative of the last pivot of the
// LDL^T factorisation in sturm_count's order: 1 / D_last(y) = sum_l s_l^2 / (theta_l - y) near an
// eigenvalue theta gives D_last'(theta) = -1 / s^2. D' follows D through the same recurrence:
// D_i' = -1 + b_{i-1}^2 D_{i-1}' / D_{i-1}^2 (kept rows: D_k' = -1 - sum ar_i^2 / (a_i - x)^2), so
// D' <= -1 and the result lies in (0, 1]. At the largest eigenvalue every earlier pivot is negative
// (the leading blocks lie below it), so the recurrence meets no small pivots unless T splits.
__device__ inline double last_component_sq(const Proj& T, double x, double pivmin) {
  double acc = 0.0, dacc = 0.0;
  for (int i = 0; i < T.k; ++i) {
    double q = T.a[i] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    const double iq = 1.0 / q;
    const double t = T.ar2[i] * iq;
    acc += t;
    dacc += t * iq;
  }
  double q = (T.a[T.k] - x) - acc;
  double dq = -1.0 - dacc;
  for (int i = T.k + 1; i < T.m; ++i) {
    if (fabs(q) < pivmin) q = -pivmin;
    const double iq = 1.0 / q;
    const double t = T.b2[i - 1] * iq;
    dq = -1.0 + t * iq * dq;
    q = (T.a[i] - x) - t;
  }
  return -1.0 / dq;
}

// Inverse iteration for the eigenvector of theta (one thread). The kept rows are eliminated first
// (s_i = (r_i - ar_i s_k) / (a_i - theta)), which leaves a tridiagonal system on rows k..m-1 with a
// modified first pivot; that system is solved by LU with partial pivoting (row interchanges create
// a second superdiagonal). The recurrences carry their running values in registers and the pivots
// are stored as reciprocals, so the three solves have no divisions and no LDS round trips on their
// dependency chains. Scratch: d, dl, du, du2, den, rhs (m doubles each) and piv (m ints).
__device__ void inverse_iteration(const Proj& T, double theta, double tiny, int seed, double* __restrict__ d,
                                  double* __restrict__ dl, double* __restrict__ du, double* __restrict__ du2,
                                  double* __restrict__ den, double* __restrict__ rhs, int* __restrict__ piv,
                                  double* __restrict__ vec) {
  const int m = T.m, k = T.k, mt = m - k;
  double corr = 0.0;
  for (int i = 0; i < k; ++i) {
    double q = T.a[i] - theta;
    if (fabs(q) < tiny) q = q < 0.0 ? -tiny : tiny;
    const double iq = 1.0 / q;
    den[i] = iq;
    corr += T.ar2[i] * iq;
  }
  // LU of the tridiagonal part; dc/uc are the current row's diagonal and superdiagonal.
  double dc = (T.a[k] - theta) - corr;
  double uc = mt > 1 ? T.b[k] : 0.0;
  for (int t = 0; t < mt - 1; ++t) {
    const double l = T.b[k + t];
    double dn = T.a[k + t + 1] - theta;
    double un = t < mt - 2 ? T.b[k + t + 1] : 0.0;
    if (fabs(dc) >= fabs(l)) {
      if (dc == 0.0) dc = tiny;
      const double f = l / dc;
      d[t] = dc;
      dl[t] = f;
      du[t] = uc;
      du2[t] = 0.0;
      piv[t] = t;
      dn -= f * uc;
    } else {
      const double f = dc / l;
      d[t] = l;
      dl[t] = f;
      du[t] = dn;
      du2[t] = un;
      piv[t] = t + 1;
      dn = uc - f * dn;
      un = -f * un;
    }
    dc = dn;
    uc = un;
  }
  if (dc == 0.0) dc = tiny;
  d[mt - 1] = dc;
  for (int t = 0; t < mt; ++t) d[t] = 1.0 / d[t];
  for (int i = 0; i < m; ++i) vec[i] = 1.0 + 0.01 * (double)((i + 3 * seed) % 7) + 0.001 * (double)seed;
  for (int it = 0; it < 3; ++it) {
    double cur = vec[k];
    for (int i = 0; i < k; ++i) cur -= T.ar[i] * vec[i] * den[i];
    for (int t = 0; t < mt - 1; ++t) {
      const double nxt = vec[k + t + 1];
      if (piv[t] == t) {
        rhs[t] = cur;
        cur = nxt - dl[t] * cur;
      } else {
        rhs[t] = nxt;
        cur = cur - dl[t] * nxt;
      }
    }
    double x1 = cur * d[mt - 1], x2 = 0.0;
    vec[k + mt - 1] = x1;
    for (int t = mt - 2; t >= 0; --t) {
      const double x0 = (rhs[t] - du[t] * x1 - du2[t] * x2) * d[t];
      vec[k + t] = x0;
      x2 = x1;
      x1 = x0;
    }
    for (int i = 0; i < k; ++i) vec[i] = (vec[i] - T.ar[i] * x1) * den[i];
    double mx = 0.0;
    for (int i = 0; i < m; ++i) mx = fmax(mx, fabs(vec[i]));
    if (!(mx > 0.0) || !isfinite(mx)) break;
    const double imx = 1.0 / mx;
    double nn = 0.0;
    for (int i = 0; i < m; ++i) {
      const double v = vec[i] * imx;
      vec[i] = v;
      nn += v * v;
    }
    const double inn = 1.0 / sqrt(nn);
    for (int i = 0; i < m; ++i) vec[i] *= inn;
  }
}

// Gershgorin interval [lo, hi], pivot floor and norm estimate of T over the whole workgroup; b2 and
// ar2 (LDS) receive the squares.
__device__ void proj_bounds(const Proj& T, double* b2, double* ar2, double* out4) {
  __shared__ double wlo[SA_WAVES_PER_BLOCK], whi[SA_WAVES_PER_BLOCK], wb2[SA_WAVES_PER_BLOCK];
  const int t = threadIdx.x;
  double glo = INFINITY, ghi = -INFINITY, mb2 = 0.0;
  for (int i = t; i < T.m; i += blockDim.x) {
    const double r = gersh_radius(T, i);
    glo = fmin(glo, T.a[i] - r);
    ghi = fmax(ghi, T.a[i] + r);
    if (i < T.k) {
      ar2[i] = T.ar[i] * T.ar[i];
      mb2 = fmax(mb2, ar2[i]);
    } else if (i < T.m - 1) {
      b2[i] = T.b[i] * T.b[i];
      mb2 = fmax(mb2, b2[i]);
    }
  }
  glo = -wave_max(-glo);
  ghi = wave_max(ghi);
  mb2 = wave_max(mb2);
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
    out4[0] = L - pad;
    out4[1] = H + pad;
    out4[2] = 1e-290 * fmax(1.0, B);
    out4[3] = tn;
  }
  __syncthreads();
}

// Largest eigenvalue of T by multisection: every thread evaluates one Sturm count per round.
// `hint` (finite) is the largest eigenvalue of a leading principal submatrix of T (the previous
// step's Ritz value: by interlacing the answer is at least that large). The first round then places
// its points geometrically above it, base + (hi - base) 2^((t - 255) / 4) with base a few ulps below
// the hint (its own rounding), so that a Ritz value that
// moved by d since the previous step is bracketed to about d / 5 at once (late in a cycle d is tiny and
// two or three uniform rounds finish it instead of six or seven from the Gershgorin interval). If every
// eigenvalue lies below the lowest point (a hint from rounding above the answer), the search goes on
// in [lo, that point].
__device__ double largest_eigenvalue(const Proj& T, const double* bounds, double hint) {
  __shared__ unsigned long long masks[SA_WAVES_PER_BLOCK];
  const int t = threadIdx.x;
  double lo = bounds[0], hi = bounds[1];
  const double pivmin = bounds[2];
  const int nsect = blockDim.x + 1;
  const bool geo = isfinite(hint) && hint > lo && hint < hi && blockDim.x == 256;
  for (int round = 0; round < 24; ++round) {
    const double width = hi - lo;
    if (width <= 4.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi)) + 2.0 * pivmin) break;
    if (geo && round == 0) {
      const double base = hint - 8.0 * 2.220446049250313e-16 * fabs(hint) - pivmin;
      const double span = hi - base;
      auto point = [&](int u) { return u >= 255 ? hi : base + span * exp2((double)(u - 255) * 0.25); };
      const int c = sturm_count(T, point(t), pivmin);
      const unsigned long long mk = __ballot(c >= T.m);
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
      // point(255) = hi lies above every eigenvalue, so first >= 0
      if (first == 0) {
        hi = point(0);
      } else if (first > 0) {
        lo = fmax(lo, point(first - 1));
        hi = point(first);
      }
      continue;
    }
    const double x = lo + width * (double)(t + 1) / (double)nsect;
    const int c = sturm_count(T, x, pivmin);
    const unsigned long long mk = __ballot(c >= T.m);
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
  return 0.5 * (lo + hi);
}

// Phase 1 of the Gram-Schmidt step: v_{j+1} = W - sum_{m >= orth_lo} h_m sigma_m v_m, raw (W = N sigma_j v_j;
// a batch scales W by sigma_j here), in FP64 per entry with the rows in eight partial sums; Z = s v_{j+1}
// and each workgroup's partial of ||v_{j+1}||^2. orth_lo is j - 1 on a selective step, 0 on a full one.
template <typename T, class M>
__global__ __launch_bounds__(SA_BLOCK) void lz_orth_kernel(M map, const T* __restrict__ s, T* __restrict__ V,
                                                           int64_t ldv, T* __restrict__ Z, const T* __restrict__ W) {
  __shared__ double coef[SA_MAX_NCV + 1];
  __shared__ double red[SA_WAVES_PER_BLOCK];
  const Slice sl = map.slice();
  const sa_state* st = state_of(sl.work);
  if (st->status != SA_RUNNING) return;
  const int j = st->j;
  const double ws = M::kScaleW ? st->sigma[j] : 1.0;  // batched: W was not scaled by sigma_j
  const int mlo = st->orth_lo;  // rows mlo..j are subtracted (set by the projections' finisher)
  for (int m = threadIdx.x; m <= j; m += blockDim.x) coef[m] = st->h[m] * st->sigma[m];
  __syncthreads();
  T* __restrict__ vn = V + (int64_t)(j + 1) * ldv;
  using VT = typename vec16<T>::type;
  constexpr int L = vec16<T>::L;
  double acc = 0.0;
  const int64_t lo = sl.lo, hi = sl.hi;
  const int64_t nv = hi > lo ? (hi - lo) / L : 0;
  // 16-byte loads: L entries per lane, eight rows per group; the rows left after the last full group
  // are loaded as one clamped group, all in flight together. A selective step (rows j - 1 and j) loads
  // just those two rows and forms the same sums: the clamped group's other six partial sums stay 0.0.
  const bool two = mlo == j - 1 && j >= 1;
  if (two) {
    const double c0 = coef[j - 1], c1 = coef[j];
    const T* __restrict__ v0 = V + (int64_t)(j - 1) * ldv;
    const T* __restrict__ v1 = V + (int64_t)j * ldv;
    for (int64_t g = threadIdx.x; g < nv; g += blockDim.x) {
      const int64_t r = lo + g * L;
      const VT w0 = *reinterpret_cast<const VT*>(&W[r]);
      const VT a0 = *reinterpret_cast<const VT*>(&v0[r]);
      const VT a1 = *reinterpret_cast<const VT*>(&v1[r]);
      VT sv;
      if (s) sv = *reinterpret_cast<const VT*>(&s[r]);
      VT o, z;
#pragma unroll
      for (int l = 0; l < L; ++l) {
        double x0 = ws * (double)w0[l];
        x0 -= c0 * (double)a0[l];
        double x1 = 0.0;
        x1 -= c1 * (double)a1[l];
        const double w = ((x0 + x1) + 0.0) + 0.0;
        o[l] = (T)w;
        acc += (double)o[l] * (double)o[l];
        z[l] = s ? sv[l] * o[l] : o[l];
      }
      *reinterpret_cast<VT*>(&vn[r]) = o;
      *reinterpret_cast<VT*>(&Z[r]) = z;
    }
  }
  for (int64_t g = two ? nv : threadIdx.x; g < nv; g += blockDim.x) {
    const int64_t r = lo + g * L;
    const VT w0 = *reinterpret_cast<const VT*>(&W[r]);
    double wv[8][L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      wv[0][l] = ws * (double)w0[l];
#pragma unroll
      for (int u = 1; u < 8; ++u) wv[u][l] = 0.0;
    }
    int m = mlo;
    for (; m + 7 <= j; m += 8) {
      VT a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = *reinterpret_cast<const VT*>(&V[(int64_t)(m + u) * ldv + r]);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
#pragma unroll
        for (int l = 0; l < L; ++l) wv[u][l] -= coef[m + u] * (double)a[u][l];
      }
    }
    if (m <= j) {
      VT a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = *reinterpret_cast<const VT*>(&V[(int64_t)min(m + u, j) * ldv + r]);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (m + u <= j) {
#pragma unroll
          for (int l = 0; l < L; ++l) wv[u][l] -= coef[m + u] * (double)a[u][l];
        }
      }
    }
    VT o, z;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const double w = ((wv[0][l] + wv[1][l]) + (wv[2][l] + wv[3][l])) + ((wv[4][l] + wv[5][l]) + (wv[6][l] + wv[7][l]));
      o[l] = (T)w;
      acc += (double)o[l] * (double)o[l];
    }
    if (s) {
      const VT sv = *reinterpret_cast<const VT*>(&s[r]);
#pragma unroll
      for (int l = 0; l < L; ++l) z[l] = sv[l] * o[l];
    } else {
      z = o;
    }
    *reinterpret_cast<VT*>(&vn[r]) = o;
    *reinterpret_cast<VT*>(&Z[r]) = z;
  }
  for (int64_t r = lo + nv * L + threadIdx.x; r < hi; r += blockDim.x) {
    double wv = ws * (double)W[r];
    for (int m = mlo; m <= j; ++m) wv -= coef[m] * (double)V[(int64_t)m * ldv + r];
    const T wt = (T)wv;
    vn[r] = wt;
    Z[r] = s ? s[r] * wt : wt;
    acc += (double)wt * (double)wt;
  }
  double* part = partials_of(sl.work);
  const double bs = block_sum(acc, red);
  if (threadIdx.x == 0) part[sl.blk] = bs;
}

// The first wave of the workgroup as the team that shares dense_ritz_vector's factorisation and
// substitutions: its meeting points are wave barriers with a workgroup-scope fence (the workspace may
// be in LDS or in global memory), far cheaper than a workgroup barrier at the ~3m of them per solve.
struct WaveTeam {
  __device__ int rank() const { return threadIdx.x & 63; }
  __device__ int size() const { return 64; }
  __device__ void sync() const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
};

// Phase 2 of the Gram-Schmidt step: beta_j from the G partial squared norms, then the FP64 Ritz solve
// of the projected matrix and the stopping test, in one workgroup.
template <class F>
__global__ __launch_bounds__(SA_BLOCK) void lz_orth_finish_kernel(F fin) {
  sa_state* st = state_of(fin.w());
  if (st->status != SA_RUNNING) return;
  const int j = st->j;
  const double beta = sqrt(sum_partials(partials_of(fin.w()), fin.g(), 1, 0));

  __shared__ double la[SA_MAX_NCV], lb[SA_MAX_NCV], lb2[SA_MAX_NCV], lar[SA_MAX_NCV], lar2[SA_MAX_NCV];
  __shared__ double lvec[SA_MAX_NCV], lbounds[4];
  __shared__ int lpiv[SA_MAX_NCV];
  __shared__ double ldense[kDenseRitz * kDenseRitz];
  __shared__ int lstatus, ldecide;
  __shared__ double lres;
  const int m = j + 1;
  const int k = st->kept;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    la[i] = st->a[i];
    lb[i] = i < j ? st->b[i] : 0.0;
    lar[i] = i < k ? st->arrow[i] : 0.0;
  }
  __syncthreads();
  const double aj = la[j];
  const bool finite = isfinite(aj) && isfinite(beta);
  const bool big = fabs(aj) > 10.0 * st->bound || beta > 10.0 * st->bound;
  double theta = 0.0;
  bool all_finite = finite && !big;
  if (all_finite) {
    for (int i = 0; i < j; ++i) all_finite = all_finite && isfinite(lb[i]) && isfinite(lar[i]);
  }
  const Proj P{m, k, la, lb, lar, lb2, lar2};
  if (all_finite) {
    proj_bounds(P, lb2, lar2, lbounds);
    theta = largest_eigenvalue(P, lbounds, j >= 1 ? st->theta : NAN);
  }
  if (threadIdx.x == 0) {
    st->b[j] = beta;
    st->steps += 1;
    int status = SA_BROKEN;
    double res = 0.0;
    if (all_finite && isfinite(theta)) {
      status = SA_RUNNING;
      st->theta = theta;
      res = beta * sqrt(last_component_sq(P, theta, lbounds[2]));
      if (st->gate && j == 0) {
        st->first_beta = beta;
        st->gate = 0;
        if (beta > st->pred_tol) status = SA_REJECTED;
      }
    }
    lstatus = status;
    lres = res;
    // The derivative estimate cancels (to 0, or to inf * 0) once theta equals a pivot to rounding, as
    // it does when a kept or start vector is already the eigenvector: every small or non-finite
    // estimate, and every stop, is decided by the pivoted dense eigenvector instead.
    ldecide = status == SA_RUNNING && (!(res > st->tol) || !isfinite(res) || st->steps >= st->maxiter);
  }
  __syncthreads();
  const bool decide = ldecide != 0;
  if (decide) {
    // LDS workspace while it fits; the restart scratch of the work buffer (idle between restarts) beyond
    double* ws = m <= kDenseRitz ? ldense : restart_of(fin.w());
    const double tiny = 2.220446049250313e-16 * fmax(lbounds[3], 1e-300);
    if (threadIdx.x < 64) dense_ritz_vector(WaveTeam{}, m, k, la, lb, lar, theta, tiny, ws, lpiv, lvec);
    __syncthreads();  // lvec for thread 0 below
  }
  if (threadIdx.x == 0) {
    int status = lstatus;
    const double res = decide ? beta * fabs(lvec[m - 1]) : lres;
    if (all_finite) st->res = res;
    if (status == SA_RUNNING) {
      if (!isfinite(res)) status = SA_BROKEN;
      else if (res <= st->tol) status = SA_CONVERGED;
      else if (st->steps >= st->maxiter) status = SA_MAXITER;
      else if (m >= st->ncv) status = SA_RESTART;
    }
    if (status == SA_CONVERGED || status == SA_MAXITER) {
      // the solve stops here (decide was set): its Ritz vector's coefficients, for the final combination
      for (int i = 0; i < m; ++i) st->ritz[i] = lvec[i];
    }
    if (all_finite) {
      st->sigma[j + 1] = beta > 0.0 ? 1.0 / beta : 0.0;
      st->j = m;
    }
    st->status = status;
  }
}

// Thick restart, part 1 (one workgroup): the `keep` largest eigenpairs of the projected matrix.
// Eigenvalue m-1-u is refined by the 16 lanes of group u with a 17-section of its bracket per round
// (Sturm counts); thread u then finds its vector by inverse iteration, and thread 0 orthonormalises
// the vectors in order (modified Gram-Schmidt) and writes the new projected matrix: diag(theta)
// bordered by arrow_u = beta_{m-1} s_u[m-1]. Scratch lives in LDS while m <= kPrepLds.
constexpr int kPrepLds = 48;
// Per-thread strides (doubles) of the inverse-iteration scratch and of the vectors in LDS: odd, so the
// threads that walk their own arrays in step fall on different LDS banks.
constexpr int kPrepScr = 8 * kPrepLds + 1;
constexpr int kPrepVec = kPrepLds + 1;

template <class F>
__global__ __launch_bounds__(SA_BLOCK) void lz_thick_prepare_kernel(F fin) {
  void* work = fin.w();
  __shared__ double la[SA_MAX_NCV], lb[SA_MAX_NCV], lb2[SA_MAX_NCV], lar[SA_MAX_NCV], lar2[SA_MAX_NCV];
  __shared__ double lbounds[4], ltheta[SA_MAX_KEEP];
  __shared__ double lscr[SA_MAX_KEEP * kPrepScr];
  __shared__ double lvec[SA_MAX_KEEP * kPrepVec];
  sa_state* st = state_of(work);
  if (st->status != SA_RESTART) return;
  const int m = st->j;
  const int k = st->kept;
  int keep = st->keep;
  if (keep > SA_MAX_KEEP) keep = SA_MAX_KEEP;
  if (keep > m - 1) keep = m - 1;
  if (keep < 1) keep = 1;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    la[i] = st->a[i];
    lb[i] = i < m - 1 ? st->b[i] : 0.0;
    lar[i] = i < k ? st->arrow[i] : 0.0;
  }
  __syncthreads();
  const Proj P{m, k, la, lb, lar, lb2, lar2};
  proj_bounds(P, lb2, lar2, lbounds);
  const int t = threadIdx.x;
  const int grp = t >> 4, lig = t & 15;
  {
    const int idx = m - 1 - (grp < keep ? grp : 0);
    double lo = lbounds[0], hi = lbounds[1];
    const double pivmin = lbounds[2];
    for (int round = 0; round < 15; ++round) {
      const double width = hi - lo;
      const double x = lo + width * (double)(lig + 1) / 17.0;
      const bool above = sturm_count(P, x, pivmin) > idx;   // x lies above eigenvalue idx
      const unsigned long long mk = __ballot(above);
      const unsigned seg = (unsigned)((mk >> (16 * (grp & 3))) & 0xFFFFull);
      const int first = seg ? __builtin_ctz(seg) : 16;
      const double nhi = first < 16 ? lo + width * (double)(first + 1) / 17.0 : hi;
      const double nlo = first == 0 ? lo : lo + width * (double)first / 17.0;
      lo = nlo;
      hi = nhi;
    }
    if (lig == 0 && grp < keep) ltheta[grp] = 0.5 * (lo + hi);
  }
  __syncthreads();
  double* S = restart_of(work);                              // S[u * SA_MAX_NCV + p], read by part 2
  double* gscr = S + (int64_t)SA_MAX_KEEP * SA_MAX_NCV;
  const bool in_lds = m <= kPrepLds;
  const int vstride = in_lds ? kPrepVec : SA_MAX_NCV;
  double* vecs = in_lds ? lvec : S;
  if (t < keep) {
    double* base = in_lds ? lscr + t * kPrepScr : gscr + (int64_t)t * 8 * SA_MAX_NCV;
    const int sc = in_lds ? kPrepLds : SA_MAX_NCV;
    const double tiny = 2.220446049250313e-16 * fmax(lbounds[3], 1e-300);
    inverse_iteration(P, ltheta[t], tiny, t, base, base + sc, base + 2 * sc, base + 3 * sc, base + 4 * sc,
                      base + 5 * sc, reinterpret_cast<int*>(base + 6 * sc), vecs + (int64_t)t * vstride);
  }
  __threadfence_block();
  __syncthreads();
  // Modified Gram-Schmidt of the vectors in order, by wave 0 with two components per lane (m <= 128);
  // the butterfly sums leave the same value in every lane, so kk stays uniform.
  __shared__ int lkk;
  if (t < 64) {
    const int p0 = t, p1 = t + 64;
    int kk = 0;
    for (int u = 0; u < keep; ++u) {
      const double* v = vecs + (int64_t)u * vstride;
      double v0 = p0 < m ? v[p0] : 0.0, v1 = p1 < m ? v[p1] : 0.0;
      for (int w = 0; w < kk; ++w) {
        const double* q = vecs + (int64_t)w * vstride;
        const double q0 = p0 < m ? q[p0] : 0.0, q1 = p1 < m ? q[p1] : 0.0;
        const double dot = wave_sum(q0 * v0 + q1 * v1);
        v0 -= dot * q0;
        v1 -= dot * q1;
      }
      const double nn = sqrt(wave_sum(v0 * v0 + v1 * v1));
      if (!(nn > 1e-8) || !isfinite(nn)) continue;     // a vector lost in the span of the earlier ones
      const double inv = 1.0 / nn;
      double* dst = vecs + (int64_t)kk * vstride;
      if (p0 < m) dst[p0] = v0 * inv;
      if (p1 < m) dst[p1] = v1 * inv;
      if (t == 0) st->a[kk] = ltheta[u];
      ++kk;
    }
    if (t == 0) lkk = kk;
  }
  __threadfence_block();
  __syncthreads();
  {
    const int kk = lkk;
    const double beta_last = st->b[m - 1];   // beta of the last step: couples v_{m-1} to v_m
    for (int u = t; u < kk; u += blockDim.x) st->arrow[u] = beta_last * vecs[(int64_t)u * vstride + (m - 1)];
    if (t == 0) st->kept = kk;
  }
  __syncthreads();
  if (in_lds) {
    for (int q = t; q < keep * m; q += blockDim.x) {
      const int u = q / m, p = q % m;
      S[(int64_t)u * SA_MAX_NCV + p] = lvec[u * kPrepVec + p];
    }
  }
}

// Thick restart, part 2: rotate the basis in place. Each thread reads all of its rows' basis entries
// before writing any, so rows 0..kept-1 can take the Ritz vectors and row kept the last Lanczos vector.
// A thread owns 16 bytes of rows (4 floats or 2 doubles) and keeps four basis loads in flight; the
// coefficients are stored basis-row-major in LDS so one broadcast read serves a whole row group.
template <typename T, class M>
__global__ __launch_bounds__(SA_BLOCK) void lz_thick_vectors_kernel(M map, const T* __restrict__ s,
                                                                    T* __restrict__ V, int64_t ldv,
                                                                    SH<T>* __restrict__ Vh, double shscale,
                                                                    T* __restrict__ Z) {
  const Slice sl = map.slice();
  void* work = sl.work;
  using VT = typename vec16<T>::type;
  constexpr int L = vec16<T>::L;
  __shared__ double coef[SA_MAX_NCV * SA_MAX_KEEP];   // coef[p * SA_MAX_KEEP + u]
  const sa_state* st = state_of(work);
  if (st->status != SA_RESTART) return;
  const int m = st->j;
  const int kk = st->kept;
  const double* S = restart_of(work);
  // shadow rows of the rotated basis: the kept Ritz vectors are unit, row kept keeps sigma_m
  const T put_kept = (T)shscale, put_last = (T)(st->sigma[m] * shscale);
  for (int q = threadIdx.x; q < kk * m; q += blockDim.x) {
    const int u = q / m, p = q % m;
    coef[p * SA_MAX_KEEP + u] = S[(int64_t)u * SA_MAX_NCV + p] * st->sigma[p];
  }
  __syncthreads();
  const int64_t lo = sl.lo, hi = sl.hi;
  const int64_t nv = hi > lo ? (hi - lo) / L : 0;
  for (int64_t g = threadIdx.x; g < nv; g += blockDim.x) {
    const int64_t r = lo + g * L;
    double acc[SA_MAX_KEEP][L];
#pragma unroll
    for (int u = 0; u < SA_MAX_KEEP; ++u) {
#pragma unroll
      for (int l = 0; l < L; ++l) acc[u][l] = 0.0;
    }
    int p = 0;
    for (; p + 7 < m; p += 8) {  // eight rows' loads in flight
      VT v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = *reinterpret_cast<const VT*>(&V[(int64_t)(p + q) * ldv + r]);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
#pragma unroll
        for (int u = 0; u < SA_MAX_KEEP; ++u) {
          if (u < kk) {
            const double c = coef[(p + q) * SA_MAX_KEEP + u];
#pragma unroll
            for (int l = 0; l < L; ++l) acc[u][l] += c * (double)v[q][l];
          }
        }
      }
    }
    for (; p + 3 < m; p += 4) {
      VT v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = *reinterpret_cast<const VT*>(&V[(int64_t)(p + q) * ldv + r]);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int u = 0; u < SA_MAX_KEEP; ++u) {
          if (u < kk) {
            const double c = coef[(p + q) * SA_MAX_KEEP + u];
#pragma unroll
            for (int l = 0; l < L; ++l) acc[u][l] += c * (double)v[q][l];
          }
        }
      }
    }
    for (; p < m; ++p) {
      const VT v = *reinterpret_cast<const VT*>(&V[(int64_t)p * ldv + r]);
#pragma unroll
      for (int u = 0; u < SA_MAX_KEEP; ++u) {
        if (u < kk) {
          const double c = coef[p * SA_MAX_KEEP + u];
#pragma unroll
          for (int l = 0; l < L; ++l) acc[u][l] += c * (double)v[l];
        }
      }
    }
    const VT lastv = *reinterpret_cast<const VT*>(&V[(int64_t)m * ldv + r]);
#pragma unroll
    for (int u = 0; u < SA_MAX_KEEP; ++u) {
      if (u < kk) {
        VT o;
#pragma unroll
        for (int l = 0; l < L; ++l) o[l] = (T)acc[u][l];
        *reinterpret_cast<VT*>(&V[(int64_t)u * ldv + r]) = o;
        if (Vh) {
          typename Shadow<T>::vec hv;  // the L shadow entries in one store
#pragma unroll
          for (int l = 0; l < L; ++l) hv[l] = (SH<T>)(o[l] * put_kept);
          *reinterpret_cast<typename Shadow<T>::vec*>(&Vh[(int64_t)u * ldv + r]) = hv;
        }
      }
    }
    *reinterpret_cast<VT*>(&V[(int64_t)kk * ldv + r]) = lastv;
    if (Vh) {
      typename Shadow<T>::vec hv;
#pragma unroll
      for (int l = 0; l < L; ++l) hv[l] = (SH<T>)(lastv[l] * put_last);
      *reinterpret_cast<typename Shadow<T>::vec*>(&Vh[(int64_t)kk * ldv + r]) = hv;
    }
    VT z = lastv;
    if (s) z = *reinterpret_cast<const VT*>(&s[r]) * lastv;
    *reinterpret_cast<VT*>(&Z[r]) = z;
  }
  for (int64_t r = lo + nv * L + threadIdx.x; r < hi; r += blockDim.x) {
    double acc[SA_MAX_KEEP];
#pragma unroll
    for (int u = 0; u < SA_MAX_KEEP; ++u) acc[u] = 0.0;
    for (int p = 0; p < m; ++p) {
      const double vp = (double)V[(int64_t)p * ldv + r];
#pragma unroll
      for (int u = 0; u < SA_MAX_KEEP; ++u)
        if (u < kk) acc[u] += coef[p * SA_MAX_KEEP + u] * vp;
    }
    const T lastv = V[(int64_t)m * ldv + r];
#pragma unroll
    for (int u = 0; u < SA_MAX_KEEP; ++u)
      if (u < kk) {
        V[(int64_t)u * ldv + r] = (T)acc[u];
        if (Vh) Vh[(int64_t)u * ldv + r] = (SH<T>)((T)acc[u] * put_kept);
      }
    V[(int64_t)kk * ldv + r] = lastv;
    if (Vh) Vh[(int64_t)kk * ldv + r] = (SH<T>)(lastv * put_last);
    Z[r] = s ? s[r] * lastv : lastv;
  }
}

// The state after the rotation: kept vectors are unit (sigma 1), the last Lanczos vector moves to
// row kept, and the cycle continues from j = kept.
template <class F>
__global__ void lz_thick_finish_kernel(F fin) {
  sa_state* st = state_of(fin.w());
  if (st->status != SA_RESTART || threadIdx.x != 0) return;
  const int m = st->j;
  const int kk = st->kept;
  const double sig_last = st->sigma[m];
  for (int u = 0; u < kk; ++u) st->sigma[u] = 1.0;
  st->sigma[kk] = sig_last;
  st->j = kk;
  st->cycles += 1;
  st->status = SA_RUNNING;
}

template <typename T, class M>
__global__ __launch_bounds__(SA_BLOCK) void lz_ritz_kernel(M map, const T* __restrict__ V, int64_t ldv,
                                                           const T* __restrict__ y_last, T* __restrict__ y,
                                                           int32_t only_if) {
  __shared__ double coef[SA_MAX_NCV];
  const Slice sl = map.slice();
  if (!sl.live) return;
  const sa_state* st = state_of(sl.work);
  if (only_if >= 0 && st->status != only_if) return;
  const int m = st->j;
  for (int i = threadIdx.x; i < m; i += blockDim.x) coef[i] = st->ritz[i] * st->sigma[i];
  __syncthreads();
  double acc[2] = {0.0, 0.0};
  for (int64_t r = sl.lo + threadIdx.x; r < sl.hi; r += blockDim.x) {
    double v4[4] = {0.0, 0.0, 0.0, 0.0};
    int i = 0;
    for (; i + 3 < m; i += 4) {
#pragma unroll
      for (int q = 0; q < 4; ++q) v4[q] += coef[i + q] * (double)V[(int64_t)(i + q) * ldv + r];
    }
    for (; i < m; ++i) v4[0] += coef[i] * (double)V[(int64_t)i * ldv + r];
    const double v = (v4[0] + v4[1]) + (v4[2] + v4[3]);
    const T vt = (T)v;
    y[r] = vt;
    acc[0] += (double)vt * (double)vt;
    if (y_last) acc[1] += (double)vt * (double)y_last[r];
  }
  block_partials<2, RED_SUM>(acc, sl.work, sl.blk);
}

static inline int red_grid(int64_t n) { return grid_for(n, SA_BLOCK * 4, SA_RED_GRID); }

// The Gram-Schmidt update reduces a single column, so it can afford more workgroups; the projections
// keep (j+1) partials per workgroup, so their grid is capped by the partials area (2048 x 136).
constexpr int SA_ORTH_GRID = 2048;
constexpr int SA_DOTS_GRID = 512;

template <typename T>
static int lanczos_start(void* work, int64_t n, const T* s, const T* x, const T* d, double pexp, double cx,
                         const T* y, double cy, T* V, int64_t ldv, void* Vh, T* Z, int32_t gate, int32_t only_if,
                         hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && x && V && Z && ldv >= n, "bad arguments");
  const int G = red_grid(n);
  hipLaunchKernelGGL((lz_start_kernel<T, OneMap>), dim3(G), dim3(SA_BLOCK), 0, stream, OneMap{work, n}, s, x, d,
                     (T)pexp, (const double*)nullptr, cx, y, cy, V, Z, gate, only_if);
  int rc = check_launch("lanczos_start");
  if (rc) return rc;
  hipLaunchKernelGGL(lz_start_finish_kernel<OneFin>, dim3(1), dim3(SA_BLOCK), 0, stream, OneFin{work, G}, gate,
                     only_if);
  return check_launch("lanczos_start_finish");
}

// One Lanczos step in three phases, in this order on one stream:
//   phase 0: the SpMV (plain CSR layout: also the sigma_j scaling)
//   phase 1: panel combine, projections and their finisher, the Gram-Schmidt update
//   phase 2: beta_j, the Ritz solve and the stopping test (one workgroup)
template <typename T>
static int lanczos_step_phase(const sa_csr* A, void* work, const T* s, T* V, int64_t ldv, void* Vh, T* Z, T* W,
                              int phase, hipStream_t stream) {
  SA_REQUIRE(A && work && V && Z && W && A->tile_rows && ldv >= A->n, "bad arguments");
  SA_REQUIRE(aligned16(V) && aligned16(W) && aligned16(Z) && (!s || aligned16(s)) && ldv % vec16<T>::L == 0,
             "V, W, Z and s must be 16-byte aligned and ldv a multiple of 16 bytes");
  SA_REQUIRE(!Vh || (aligned16(Vh) && ldv % (2 * vec16<T>::L) == 0),
             "the shadow basis must be 16-byte aligned and ldv a multiple of 32 bytes");
  SA_REQUIRE(phase >= 0 && phase <= 2, "phase must be 0, 1 or 2");
  const int64_t n = A->n;
  const int Go = grid_for(n, SA_BLOCK * 2, SA_ORTH_GRID);
  if (phase == 0) {
    SpmvParams<T> p = params_from<T>(A);
    p.g = Z;
    p.s = s;
    p.lz = reinterpret_cast<const sa_state*>(work);
    p.gate = &p.lz->status;
    p.y = W;
    return launch_spmv<T>(p, A->tile_nnz, stream);
  }
  if (phase == 2) {
    hipLaunchKernelGGL(lz_orth_finish_kernel<OneFin>, dim3(1), dim3(SA_BLOCK), 0, stream, OneFin{work, Go});
    return check_launch("lanczos_orth_finish");
  }
  int rc = 0;
  // the panel combine runs inside the projections kernel when the partial rows are 16-byte aligned
  const bool fuse = A->npanels > 0 && A->npanels % 8 == 0 && aligned16(A->partial) &&
                    A->pld % vec16<T>::L == 0 && (!s || aligned16(s));
  if (A->npanels > 0 && !fuse) {
    hipLaunchKernelGGL(lz_combine_kernel<T>, dim3(grid_for(n, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK), 0, stream, work,
                       n, A->npanels, (const T*)A->partial, A->pld, s, W);
    rc = check_launch("lanczos_combine");
    if (rc) return rc;
  }
  const int Gd = grid_for(n, SA_BLOCK * 4, SA_DOTS_GRID);
  hipLaunchKernelGGL((lz_dots_kernel<T, OneMap>), dim3(Gd), dim3(SA_BLOCK), 0, stream, OneMap{work, n},
                     (const T*)V, ldv, (SH<T>*)Vh, sqrt((double)n), W, fuse ? (const T*)A->partial : nullptr,
                     A->pld, fuse ? (int)A->npanels : 0, s);
  rc = check_launch("lanczos_dots");
  if (rc) return rc;
  hipLaunchKernelGGL(lz_dots_finish_kernel<OneFin>, dim3(1), dim3(SA_BLOCK), 0, stream, OneFin{work, Gd});
  rc = check_launch("lanczos_dots_finish");
  if (rc) return rc;
  hipLaunchKernelGGL((lz_orth_kernel<T, OneMap>), dim3(Go), dim3(SA_BLOCK), 0, stream, OneMap{work, n}, s, V, ldv,
                     Z, (const T*)W);
  return check_launch("lanczos_orth");
}

template <typename T>
static int lanczos_step(const sa_csr* A, void* work, const T* s, T* V, int64_t ldv, void* Vh, T* Z, T* W,
                        hipStream_t stream) {
  for (int phase = 0; phase < 3; ++phase) {
    const int rc = lanczos_step_phase<T>(A, work, s, V, ldv, Vh, Z, W, phase, stream);
    if (rc) return rc;
  }
  return 0;
}

template <typename T>
static int lanczos_thick_restart(void* work, int64_t n, const T* s, T* V, int64_t ldv, void* Vh, T* Z,
                                 hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && V && Z && ldv >= n, "bad arguments");
  SA_REQUIRE(aligned16(V) && aligned16(Z) && aligned16(s) && ldv % vec16<T>::L == 0,
             "V, Z and s must be 16-byte aligned and ldv a multiple of 16 bytes");
  hipLaunchKernelGGL(lz_thick_prepare_kernel<OneFin>, dim3(1), dim3(SA_BLOCK), 0, stream, OneFin{work, 1});
  int rc = check_launch("lanczos_thick_prepare");
  if (rc) return rc;
  hipLaunchKernelGGL((lz_thick_vectors_kernel<T, OneMap>), dim3(grid_for(n, SA_BLOCK * 2, SA_ORTH_GRID)),
                     dim3(SA_BLOCK), 0, stream, OneMap{work, n}, s, V, ldv, (SH<T>*)Vh, sqrt((double)n), Z);
  rc = check_launch("lanczos_thick_vectors");
  if (rc) return rc;
  hipLaunchKernelGGL(lz_thick_finish_kernel<OneFin>, dim3(1), dim3(64), 0, stream, OneFin{work, 1});
  return check_launch("lanczos_thick_finish");
}

template <typename T>
static int lanczos_ritz(void* work, int64_t n, const T* V, int64_t ldv, const T* y_last, T* y, int32_t only_if,
                        hipStream_t stream) {
  SA_REQUIRE(work && n > 0 && V && y && ldv >= n, "bad arguments");
  const int G = red_grid(n);
  hipLaunchKernelGGL((lz_ritz_kernel<T, OneMap>), dim3(G), dim3(SA_BLOCK), 0, stream, OneMap{work, n}, V, ldv,
                     y_last, y, only_if);
  int rc = check_launch("lanczos_ritz");
  if (rc) return rc;
  FinishSpec f{};
  f.cols = 2;
  f.only_if = only_if;
  f.take_sqrt[0] = 1;
  f.dst[0] = (int32_t)offsetof(sa_state, r[4]);
  f.dst[1] = (int32_t)offsetof(sa_state, r[5]);
  return launch_finish(work, G, f, stream);
}

// ---- batched forms ------------------------------------------------------------------------------
bool batch_ok(const sa_batch* Bt) {
  return Bt && Bt->B >= 1 && Bt->row_off && Bt->row_end && Bt->blk_off && Bt->works && Bt->nblocks >= Bt->B &&
         Bt->nblocks < (1ll << 31) && Bt->rows_per_block >= 64 && Bt->rows_per_block % 64 == 0 &&
         Bt->stride >= (int64_t)(state_bytes_padded() + (kPartials + kRestart) * sizeof(double));
}

BatchMap bmap(const sa_batch* Bt) {
  return BatchMap{reinterpret_cast<char*>(Bt->works), Bt->stride, Bt->row_off, Bt->row_end, Bt->blk_off, Bt->B,
                  Bt->rows_per_block, Bt->inst_mask};
}

BatchFin bfin(const sa_batch* Bt) {
  return BatchFin{reinterpret_cast<char*>(Bt->works), Bt->stride, Bt->blk_off, Bt->inst_mask};
}

static int batch_begin(const sa_batch* Bt, int32_t ncv, int32_t maxiter, double tol, const double* bounds,
                       double pred_tol, int32_t keep, double orth_tau, hipStream_t stream) {
  SA_REQUIRE(batch_ok(Bt) && bounds, "bad arguments");
  SA_REQUIRE(ncv >= 2 && ncv <= SA_MAX_NCV && maxiter >= 1, "bad arguments (2 <= ncv <= 128)");
  SA_REQUIRE(keep >= 1 && keep <= SA_MAX_KEEP && keep < ncv, "keep must be in [1, min(16, ncv - 1)]");
  SA_REQUIRE(orth_tau >= 0.0 && orth_tau < 1.0, "orth_tau must be in [0, 1)");
  hipLaunchKernelGGL(lz_begin_kernel<BatchFin>, dim3(Bt->B), dim3(64), 0, stream, bfin(Bt), ncv, maxiter, tol, 0.0,
                     bounds, pred_tol, keep, orth_tau);
  return check_launch("batch_lanczos_begin");
}

template <typename T>
static int batch_start(const sa_batch* Bt, const T* s, const T* x, const T* d, double pexp, const double* pexp_each,
                       double cx, const T* y, double cy, T* V, int64_t ldv, T* Z, int32_t gate, int32_t only_if,
                       hipStream_t stream) {
  SA_REQUIRE(batch_ok(Bt) && x && V && Z, "bad arguments");
  hipLaunchKernelGGL((lz_start_kernel<T, BatchMap>), dim3((unsigned)Bt->nblocks), dim3(SA_BLOCK), 0, stream, bmap(Bt),
                     s, x, d, (T)pexp, pexp_each, cx, y, cy, V, Z, gate, only_if);
  int rc = check_launch("batch_lanczos_start");
  if (rc) return rc;
  hipLaunchKernelGGL(lz_start_finish_kernel<BatchFin>, dim3(Bt->B), dim3(SA_BLOCK), 0, stream, bfin(Bt), gate,
                     only_if);
  return check_launch("batch_lanczos_start_finish");
}

// One step for every running instance: the plain fused SpMV over the packed CSR (sigma_j is applied
// by the projections and the update, per instance), projections, update, Ritz solves.
template <typename T>
static int batch_step(const sa_csr* A, const sa_batch* Bt, const T* s, T* V, int64_t ldv, T* Z, T* W,
                      hipStream_t stream) {
  SA_REQUIRE(A && batch_ok(Bt) && V && Z && W && A->tile_rows && ldv >= A->n, "bad arguments");
  SA_REQUIRE(!Bt->inst_mask, "the batched step selects instances by status, not by a mask");
  SA_REQUIRE(A->npanels == 0, "the batched step uses the plain CSR layout");
  SA_REQUIRE(aligned16(V) && aligned16(W) && aligned16(Z) && (!s || aligned16(s)) && ldv % vec16<T>::L == 0,
             "V, W, Z and s must be 16-byte aligned and ldv a multiple of 16 bytes");
  SpmvParams<T> p = params_from<T>(A);
  p.g = Z;
  p.s = s;
  p.y = W;
  // the SpMV skips the tiles (or instance-staged workgroups) of instances that are not running
  p.tile_inst = Bt->tile_inst;
  p.gate_base = reinterpret_cast<const char*>(Bt->works) + offsetof(sa_state, status);
  p.gate_stride = Bt->stride;
  p.tile_list = Bt->tile_list;
  p.ntile_list = Bt->ntile_list;
  int rc = launch_spmv<T>(p, A->tile_nnz, stream);
  if (rc) return rc;
  const dim3 grid((unsigned)Bt->nblocks), fgrid(Bt->B);
  hipLaunchKernelGGL((lz_dots_kernel<T, BatchMap>), grid, dim3(SA_BLOCK), 0, stream, bmap(Bt), (const T*)V, ldv,
                     (SH<T>*)nullptr, 1.0, W, (const T*)nullptr, (int64_t)0, 0, (const T*)nullptr);
  rc = check_launch("batch_lanczos_dots");
  if (rc) return rc;
  hipLaunchKernelGGL(lz_dots_finish_kernel<BatchFin>, fgrid, dim3(SA_BLOCK), 0, stream, bfin(Bt));
  rc = check_launch("batch_lanczos_dots_finish");
  if (rc) return rc;
  hipLaunchKernelGGL((lz_orth_kernel<T, BatchMap>), grid, dim3(SA_BLOCK), 0, stream, bmap(Bt), s, V, ldv, Z,
                     (const T*)W);
  rc = check_launch("batch_lanczos_orth");
  if (rc) return rc;
  hipLaunchKernelGGL(lz_orth_finish_kernel<BatchFin>, fgrid, dim3(SA_BLOCK), 0, stream, bfin(Bt));
  return check_launch("batch_lanczos_orth_finish");
}

template <typename T>
static int batch_thick_restart(const sa_batch* Bt, const T* s, T* V, int64_t ldv, T* Z, hipStream_t stream) {
  SA_REQUIRE(batch_ok(Bt) && V && Z, "bad arguments");
  SA_REQUIRE(!Bt->inst_mask, "the batched restart selects instances by status, not by a mask");
  SA_REQUIRE(aligned16(V) && aligned16(Z) && aligned16(s) && ldv % vec16<T>::L == 0,
             "V, Z and s must be 16-byte aligned and ldv a multiple of 16 bytes");
  hipLaunchKernelGGL(lz_thick_prepare_kernel<BatchFin>, dim3(Bt->B), dim3(SA_BLOCK), 0, stream, bfin(Bt));
  int rc = check_launch("batch_thick_prepare");
  if (rc) return rc;
  hipLaunchKernelGGL((lz_thick_vectors_kernel<T, BatchMap>), dim3((unsigned)Bt->nblocks), dim3(SA_BLOCK), 0, stream,
                     bmap(Bt), s, V, ldv, (SH<T>*)nullptr, 1.0, Z);
  rc = check_launch("batch_thick_vectors");
  if (rc) return rc;
  hipLaunchKernelGGL(lz_thick_finish_kernel<BatchFin>, dim3(Bt->B), dim3(64), 0, stream, bfin(Bt));
  return check_launch("batch_thick_finish");
}

template <typename T>
static int batch_ritz(const sa_batch* Bt, const T* V, int64_t ldv, const T* y_last, T* y, int32_t only_if,
                      hipStream_t stream) {
  SA_REQUIRE(batch_ok(Bt) && V && y, "bad arguments");
  hipLaunchKernelGGL((lz_ritz_kernel<T, BatchMap>), dim3((unsigned)Bt->nblocks), dim3(SA_BLOCK), 0, stream, bmap(Bt),
                     V, ldv, y_last, y, only_if);
  int rc = check_launch("batch_lanczos_ritz");
  if (rc) return rc;
  FinishSpec f{};
  f.cols = 2;
  f.only_if = only_if;
  f.take_sqrt[0] = 1;
  f.dst[0] = (int32_t)offsetof(sa_state, r[4]);
  f.dst[1] = (int32_t)offsetof(sa_state, r[5]);
  return launch_finish_batch(Bt, f, stream);
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_lanczos_begin(void* work, int32_t ncv, int32_t maxiter, double tol, double bound, double pred_tol,
                     int32_t keep, double orth_tau, hipStream_t stream) {
  SA_REQUIRE(work && ncv >= 2 && ncv <= SA_MAX_NCV && maxiter >= 1, "bad arguments (2 <= ncv <= 128)");
  SA_REQUIRE(keep >= 1 && keep <= SA_MAX_KEEP && keep < ncv, "keep must be in [1, min(16, ncv - 1)]");
  SA_REQUIRE(orth_tau >= 0.0 && orth_tau < 1.0, "orth_tau must be in [0, 1)");
  hipLaunchKernelGGL(lz_begin_kernel<OneFin>, dim3(1), dim3(64), 0, stream, OneFin{work, 1}, ncv, maxiter, tol, bound,
                     (const double*)nullptr, pred_tol, keep, orth_tau);
  return check_launch("lanczos_begin");
}

int sa_lanczos_start_f32(void* work, int64_t n, const float* s, const float* x, const float* d, double pexp,
                         double cx, const float* y, double cy, float* V, int64_t ldv, void* Vh, float* Z, int32_t gate,
                         int32_t only_if, hipStream_t stream) {
  return lanczos_start<float>(work, n, s, x, d, pexp, cx, y, cy, V, ldv, Vh, Z, gate, only_if, stream);
}

int sa_lanczos_start_f64(void* work, int64_t n, const double* s, const double* x, const double* d, double pexp,
                         double cx, const double* y, double cy, double* V, int64_t ldv, void* Vh, double* Z, int32_t gate,
                         int32_t only_if, hipStream_t stream) {
  return lanczos_start<double>(work, n, s, x, d, pexp, cx, y, cy, V, ldv, Vh, Z, gate, only_if, stream);
}

int sa_lanczos_step_f32(const sa_csr* A, void* work, const float* s, float* V, int64_t ldv, void* Vh, float* Z,
                        float* W, hipStream_t stream) {
  return lanczos_step<float>(A, work, s, V, ldv, Vh, Z, W, stream);
}

int sa_lanczos_step_f64(const sa_csr* A, void* work, const double* s, double* V, int64_t ldv, void* Vh, double* Z,
                        double* W, hipStream_t stream) {
  return lanczos_step<double>(A, work, s, V, ldv, Vh, Z, W, stream);
}

int sa_lanczos_thick_restart_f32(void* work, int64_t n, const float* s, float* V, int64_t ldv, void* Vh,
                                 float* Z, hipStream_t stream) {
  return lanczos_thick_restart<float>(work, n, s, V, ldv, Vh, Z, stream);
}

int sa_lanczos_thick_restart_f64(void* work, int64_t n, const double* s, double* V, int64_t ldv, void* Vh,
                                 double* Z, hipStream_t stream) {
  return lanczos_thick_restart<double>(work, n, s, V, ldv, Vh, Z, stream);
}

int sa_lanczos_ritz_f32(void* work, int64_t n, const float* V, int64_t ldv, const float* y_last, float* y,
                        int32_t only_if, hipStream_t stream) {
  return lanczos_ritz<float>(work, n, V, ldv, y_last, y, only_if, stream);
}

int sa_lanczos_ritz_f64(void* work, int64_t n, const double* V, int64_t ldv, const double* y_last, double* y,
                        int32_t only_if, hipStream_t stream) {
  return lanczos_ritz<double>(work, n, V, ldv, y_last, y, only_if, stream);
}

int sa_batch_lanczos_begin(const sa_batch* Bt, int32_t ncv, int32_t maxiter, double tol, const double* bounds,
                           double pred_tol, int32_t keep, double orth_tau, hipStream_t stream) {
  return batch_begin(Bt, ncv, maxiter, tol, bounds, pred_tol, keep, orth_tau, stream);
}

int sa_batch_lanczos_start_f32(const sa_batch* Bt, const float* s, const float* x, const float* d, double pexp,
                               double cx, const float* y, double cy, float* V, int64_t ldv, float* Z, int32_t gate,
                               int32_t only_if, hipStream_t stream) {
  return batch_start<float>(Bt, s, x, d, pexp, nullptr, cx, y, cy, V, ldv, Z, gate, only_if, stream);
}

int sa_batch_lanczos_start_each_f32(const sa_batch* Bt, const float* s, const float* x, const float* d,
                                    const double* pexp, float* V, int64_t ldv, float* Z, int32_t gate,
                                    hipStream_t stream) {
  SA_REQUIRE(pexp && d, "bad arguments");
  return batch_start<float>(Bt, s, x, d, 0.0, pexp, 1.0, nullptr, 0.0, V, ldv, Z, gate, -1, stream);
}

int sa_batch_lanczos_start_f64(const sa_batch* Bt, const double* s, const double* x, const double* d, double pexp,
                               double cx, const double* y, double cy, double* V, int64_t ldv, double* Z,
                               int32_t gate, int32_t only_if, hipStream_t stream) {
  return batch_start<double>(Bt, s, x, d, pexp, nullptr, cx, y, cy, V, ldv, Z, gate, only_if, stream);
}

int sa_batch_lanczos_start_each_f64(const sa_batch* Bt, const double* s, const double* x, const double* d,
                                    const double* pexp, double* V, int64_t ldv, double* Z, int32_t gate,
                                    hipStream_t stream) {
  SA_REQUIRE(pexp && d, "bad arguments");
  return batch_start<double>(Bt, s, x, d, 0.0, pexp, 1.0, nullptr, 0.0, V, ldv, Z, gate, -1, stream);
}

int sa_batch_lanczos_step_f32(const sa_csr* A, const sa_batch* Bt, const float* s, float* V, int64_t ldv, float* Z,
                              float* W, hipStream_t stream) {
  return batch_step<float>(A, Bt, s, V, ldv, Z, W, stream);
}

int sa_batch_lanczos_step_f64(const sa_csr* A, const sa_batch* Bt, const double* s, double* V, int64_t ldv,
                              double* Z, double* W, hipStream_t stream) {
  return batch_step<double>(A, Bt, s, V, ldv, Z, W, stream);
}

int sa_batch_lanczos_thick_restart_f32(const sa_batch* Bt, const float* s, float* V, int64_t ldv, float* Z,
                                       hipStream_t stream) {
  return batch_thick_restart<float>(Bt, s, V, ldv, Z, stream);
}

int sa_batch_lanczos_thick_restart_f64(const sa_batch* Bt, const double* s, double* V, int64_t ldv, double* Z,
                                       hipStream_t stream) {
  return batch_thick_restart<double>(Bt, s, V, ldv, Z, stream);
}

int sa_batch_lanczos_ritz_f32(const sa_batch* Bt, const float* V, int64_t ldv, const float* y_last, float* y,
                              int32_t only_if, hipStream_t stream) {
  return batch_ritz<float>(Bt, V, ldv, y_last, y, only_if, stream);
}

int sa_batch_lanczos_ritz_f64(const sa_batch* Bt, const double* V, int64_t ldv, const double* y_last, double* y,
                              int32_t only_if, hipStream_t stream) {
  return batch_ritz<double>(Bt, V, ldv, y_last, y, only_if, stream);
}

}  // extern "C"
