// Eigenvector of the projected matrix of the Lanczos solve for one eigenvalue, by inverse iteration on
// the dense matrix with partial pivoting. Host and device code: the factorisation is shared by a team
// of threads (one wavefront on the device; one thread, or host threads with a barrier, in the CPU tests
// that check it against numpy).
//
// The projected matrix of order m: rows 0..k-1 are kept Ritz values (diagonal a[i]) coupled to row k
// by the arrow ar[i]; rows k..m-1 are tridiagonal (diagonal a[i], off-diagonal b[i] between i, i + 1).
// The derivative formula of the stopping estimate and the kept-rows-first elimination both divide by
// a_i - theta, which cancels to rounding once a start or kept vector is (nearly) the eigenvector;
// pivoting over the whole matrix does not, so this vector decides convergence and forms the final
// Ritz vector.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace sa {

constexpr int kDenseRitz = 48;  // largest order whose workspace (kDenseRitz^2 doubles) fits in LDS

// A team of one: the whole routine runs on the calling thread.
struct SerialTeam {
  __host__ __device__ int rank() const { return 0; }
  __host__ __device__ int size() const { return 1; }
  __host__ __device__ void sync() const {}
};

// x (m entries, unit norm) from two inverse-iteration solves of (P - theta I) x = b starting from the
// all-ones vector; pivots below tiny are set to +-tiny. M: m * m doubles of workspace, perm: m ints,
// both visible to the whole team, as is x (m entries), which every rank sees complete on return.
// Every member of the team calls this with the same arguments. The arithmetic of each entry does not
// depend on the team's size, so every team computes the same bits.
template <class Team>
__host__ __device__ inline void dense_ritz_vector(const Team& team, int m, int k, const double* a, const double* b,
                                                  const double* ar, double theta, double tiny, double* M, int* perm,
                                                  double* x) {
  const int t = team.rank(), nt = team.size();
  for (int i = t; i < m * m; i += nt) M[i] = 0.0;
  team.sync();
  for (int i = t; i < m; i += nt) {
    M[i * m + i] = a[i] - theta;
    if (i < k) {
      M[i * m + k] = ar[i];
      M[k * m + i] = ar[i];
    } else if (i + 1 < m) {
      M[i * m + i + 1] = b[i];
      M[(i + 1) * m + i] = b[i];
    }
  }
  // P M = L U with full-row interchanges: the multipliers end in their final rows
  for (int c = 0; c < m; ++c) {
    team.sync();  // the previous column's multipliers and this column's values are in place
    if (t == 0) {  // the first row of largest magnitude (the running maximum kept in a register)
      int p = c;
      double best = fabs(M[c * m + c]);
      for (int r = c + 1; r < m; ++r) {
        const double v = fabs(M[r * m + c]);
        if (v > best) {
          best = v;
          p = r;
        }
      }
      perm[c] = p;
    }
    team.sync();
    const int p = perm[c];
    if (p != c)
      for (int cc = t; cc < m; cc += nt) {
        const double v = M[c * m + cc];
        M[c * m + cc] = M[p * m + cc];
        M[p * m + cc] = v;
      }
    team.sync();
    double pv = M[c * m + c];
    if (!(fabs(pv) >= tiny)) pv = pv < 0.0 ? -tiny : tiny;
    const double inv = 1.0 / pv;
    // each rank eliminates its own rows of the trailing block (a rank of a slower team reading the
    // pivot after rank 0 stored its clamped value clamps it to the same value)
    if (t == 0) M[c * m + c] = pv;
    for (int r = c + 1 + t; r < m; r += nt) {
      double* __restrict__ row = M + r * m;
      const double* __restrict__ piv = M + c * m;
      const double l = row[c] * inv;
      for (int cc = c + 1; cc < m; ++cc) row[cc] -= l * piv[cc];
      row[c] = l;
    }
  }
  // Two inverse-iteration solves. Rank 0 applies the row interchanges and normalises; the
  // substitutions go column by column, each rank updating its own rows (r = t mod nt), so every x[r]
  // takes its updates in the same order (columns ascending for L, descending for U) whatever the team.
  for (int i = t; i < m; i += nt) x[i] = 1.0;
  for (int round = 0; round < 2; ++round) {
    team.sync();
    if (t == 0 && round > 0) {  // (the first solve's right-hand side, all ones, is invariant under P)
      for (int c = 0; c < m; ++c) {  // x <- P x: the interchanges in order
        const int p = perm[c];
        if (p != c) {
          const double v = x[c];
          x[c] = x[p];
          x[p] = v;
        }
      }
    }
    for (int c = 0; c < m; ++c) {  // L y = x (unit lower)
      team.sync();
      const double xc = x[c];
      for (int r = c + 1 + t; r < m; r += nt) x[r] -= M[r * m + c] * xc;
    }
    for (int c = m - 1; c >= 0; --c) {  // U x = y
      team.sync();
      if (t == 0) x[c] = x[c] / M[c * m + c];
      team.sync();
      const double xc = x[c];
      for (int r = t; r < c; r += nt) x[r] -= M[r * m + c] * xc;
    }
    team.sync();
    if (t == 0) {
      double mx = 0.0;  // normalise (by the largest entry first, so that the squares cannot overflow)
      for (int i = 0; i < m; ++i) mx = fmax(mx, fabs(x[i]));
      const double s = mx > 0.0 ? 1.0 / mx : 0.0;
      double nn = 0.0;
      for (int i = 0; i < m; ++i) {
        x[i] *= s;
        nn += x[i] * x[i];
      }
      const double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
      for (int i = 0; i < m; ++i) x[i] *= inv;
    }
  }
  team.sync();
}

}  // namespace sa
