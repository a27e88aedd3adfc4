// Eigenvector of the projected matrix of the Lanczos solve for one eigenvalue, by inverse iteration on
// the dense matrix with partial pivoting. Host and device code: the factorisation is shared by a team
// of threads (a workgroup on the device; one thread, or host threads with a barrier, in the CPU tests
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
// both visible to the whole team; x is written by rank 0 (and visible to it only when this returns).
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
    if (t == 0) {
      int p = c;
      for (int r = c + 1; r < m; ++r)
        if (fabs(M[r * m + c]) > fabs(M[p * m + c])) p = r;
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
    const int w = m - c - 1;  // order of the trailing block
    for (int e = t; e < w * w; e += nt) {
      const int r = c + 1 + e / w, cc = c + 1 + e % w;
      M[r * m + cc] -= M[r * m + c] * inv * M[c * m + cc];
    }
    team.sync();  // every update has read column c
    if (t == 0) M[c * m + c] = pv;
    for (int r = c + 1 + t; r < m; r += nt) M[r * m + c] *= inv;
  }
  team.sync();
  if (t != 0) return;
  for (int i = 0; i < m; ++i) x[i] = 1.0;
  for (int round = 0; round < 2; ++round) {
    for (int c = 0; c < m; ++c) {  // x <- P x: the interchanges in order
      const int p = perm[c];
      if (p != c) {
        const double v = x[c];
        x[c] = x[p];
        x[p] = v;
      }
    }
    for (int c = 0; c < m; ++c)  // L y = x (unit lower)
      for (int r = c + 1; r < m; ++r) x[r] -= M[r * m + c] * x[c];
    for (int r = m - 1; r >= 0; --r) {  // U x = y
      double v = x[r];
      for (int cc = r + 1; cc < m; ++cc) v -= M[r * m + cc] * x[cc];
      x[r] = v / M[r * m + r];
    }
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

}  // namespace sa
