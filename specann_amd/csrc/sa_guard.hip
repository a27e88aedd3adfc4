// Guarded local-search policies on gfx950 (the ls_policy values fixed_guard, adaptive_guard,
// independent_set and sequential_oracle of models/local_search_guards/policies.py:13-18).
//
// The candidates use the local search's node-major layout: spins int8 [n][kp], fields h = J s kept
// exact (int32 for integer couplings, FP64 for FP32 couplings J / c) [n][kp], kp = 64 * groups, and a
// wavefront handles one (node, group), 64 candidates in its lanes.
//
// Fixed / Adaptive Guard (models/spectral_annealing_safe.py:223-797): a round proposes, for every
// active candidate, the Jacobi flips of a random share of the nodes (each selected node whose spin
// disagrees with its field's sign, judged on the current spins), and keeps the proposal only if the
// exact energy strictly drops. Here the proposal is a bit mask per (node, group), its exact energy
// change comes from the maintained fields in one pass over the flipped rows,
//   dE = sum_{i in F} 2 s_i h_i - 2 sum_{i in F} s_i sum_{j in F, j ~ i} J_ij s_j,
// (the second term: adjacent flipped pairs, which the Jacobi step does not see), and only accepted
// candidates' flips are committed, with the field updates of the default search.
//
// Independent-Set (conflict_free.py:27-80): one colour class per step; a node flips when its single-
// flip change 2 s h is below minus the candidate's tolerance, which for a class (no two nodes adjacent)
// sums to the exact change of the step.
//
// Sequential-Oracle (sequential_oracle.py:22-121): every candidate flips its best single spin
// (the smallest 2 s_i h_i, first node on ties) while that is below minus its tolerance; one
// 1024-thread workgroup runs all steps of one candidate group.
#include <cstddef>

#include "sa_reduce.hpp"

namespace sa {

namespace {

__device__ inline int32_t guniform(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename Src>
__device__ inline Src guniform_val(Src v) {
  static_assert(sizeof(Src) == 4, "4-byte couplings");
  return __builtin_bit_cast(Src, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int32_t, v)));
}

__device__ inline uint64_t guniform64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

template <typename F>
__device__ inline bool disagrees(int32_t s, F h) {
  return s > 0 ? h < (F)0 : h > (F)0;
}

__device__ inline void add_to(int64_t* p, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}
__device__ inline void add_to(double* p, double v) { atomicAdd(p, v); }

}  // namespace

// ---- Fixed / Adaptive Guard -----------------------------------------------------------------
// The proposal of round `round`: node i is selected for candidate c iff (mix3(seed, round, i) >> 11) <
// thresh[c] (thresh = 0: inactive, 2^53: every node); mask[i * groups + g] bit l: candidate 64 g + l
// flips i. counts[c] += proposed flips.
template <typename F>
__global__ __launch_bounds__(SA_BLOCK) void guard_mask_kernel(int64_t n, int32_t groups,
                                                              const int8_t* __restrict__ spins,
                                                              const F* __restrict__ fields, uint64_t seed,
                                                              uint64_t round, const uint64_t* __restrict__ thresh,
                                                              uint64_t* __restrict__ mask, int32_t* __restrict__ counts) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.y;
  const int c = g * 64 + lane;
  const int64_t kp = (int64_t)groups * 64;
  const uint64_t t = thresh[c];
  const int64_t nw = (int64_t)gridDim.x * SA_WAVES_PER_BLOCK;
  int32_t cnt = 0;
  for (int64_t i = (int64_t)blockIdx.x * SA_WAVES_PER_BLOCK + wave; i < n; i += nw) {
    const uint64_t h = mix3(seed, round, (uint64_t)i) >> 11;
    const bool flip = h < t && disagrees<F>(spins[i * kp + c], fields[i * kp + c]);
    const uint64_t m = __ballot(flip);
    if (lane == 0) mask[i * groups + g] = m;
    cnt += flip ? 1 : 0;
  }
  if (cnt) atomicAdd(&counts[c], cnt);
}

// delta[c] += the exact energy change of candidate c's proposal (see the header comment), from the
// rows of its flipped nodes.
template <typename Src, typename F, typename D>
__global__ __launch_bounds__(SA_BLOCK) void guard_delta_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                               const int32_t* __restrict__ indices,
                                                               const Src* __restrict__ jval, int32_t groups,
                                                               const int8_t* __restrict__ spins,
                                                               const F* __restrict__ fields,
                                                               const uint64_t* __restrict__ mask,
                                                               D* __restrict__ delta) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.y;
  const int c = g * 64 + lane;
  const int64_t kp = (int64_t)groups * 64;
  const int64_t nw = (int64_t)gridDim.x * SA_WAVES_PER_BLOCK;
  D acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * SA_WAVES_PER_BLOCK + wave; i < n; i += nw) {
    const uint64_t mi = guniform64(mask[i * groups + g]);
    if (mi == 0) continue;
    const bool mine = (mi >> lane) & 1ull;
    const int32_t si = spins[i * kp + c];
    const F hi = fields[i * kp + c];
    D inner = 0;  // sum over flipped neighbours j of J_ij s_j
    const int32_t e0 = guniform(indptr[i]), e1 = guniform(indptr[i + 1]);
    for (int32_t e = e0; e < e1; ++e) {
      const int32_t j = guniform(indices[e]);
      const uint64_t mj = guniform64(mask[(int64_t)j * groups + g]);
      if (!(mj & mi)) continue;
      const Src jv = guniform_val(jval[e]);
      const int32_t sj = spins[(int64_t)j * kp + c];
      if ((mj >> lane) & 1ull) inner += (D)jv * (D)sj;
    }
    if (mine) acc += (D)2 * (D)si * ((D)hi - inner);
  }
  if (acc != (D)0) add_to(&delta[c], acc);
}

// The proposal's flips of the accepted candidates (accepted[g] bit l: candidate 64 g + l): spins
// flipped and the neighbours' fields updated, as in the default search.
template <typename Src, typename F>
__global__ __launch_bounds__(SA_BLOCK) void guard_commit_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                                const int32_t* __restrict__ indices,
                                                                const Src* __restrict__ jval, int32_t groups,
                                                                int8_t* __restrict__ spins, F* __restrict__ fields,
                                                                const uint64_t* __restrict__ mask,
                                                                const uint64_t* __restrict__ accepted) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.y;
  const int c = g * 64 + lane;
  const int64_t kp = (int64_t)groups * 64;
  const uint64_t acc = accepted[g];
  const int64_t nw = (int64_t)gridDim.x * SA_WAVES_PER_BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * SA_WAVES_PER_BLOCK + wave; i < n; i += nw) {
    const uint64_t bits = guniform64(mask[i * groups + g]) & acc;
    if (bits == 0) continue;
    const bool flip = (bits >> lane) & 1ull;
    const int32_t snew = -(int32_t)spins[i * kp + c];
    if (flip) spins[i * kp + c] = (int8_t)snew;
    const int32_t e0 = guniform(indptr[i]), e1 = guniform(indptr[i + 1]);
    for (int32_t e = e0; e < e1; ++e) {
      const int32_t j = guniform(indices[e]);
      const Src jv = guniform_val(jval[e]);
      if (flip) atomicAdd(&fields[(int64_t)j * kp + c], (F)2 * (F)jv * (F)snew);
    }
  }
}

// ---- Independent-Set ------------------------------------------------------------------------
// One colour class: node i flips candidate c iff c is active (active[g] bit) and s h < -thr[c]; the
// flips, and their energy change sum 2 s h (field units), are added to flips[c] and delta[c].
template <typename Src, typename F, typename D>
__global__ __launch_bounds__(SA_BLOCK) void guard_class_kernel(const int32_t* __restrict__ indptr,
                                                               const int32_t* __restrict__ indices,
                                                               const Src* __restrict__ jval, int32_t groups,
                                                               const int32_t* __restrict__ nodes, int64_t count,
                                                               const uint64_t* __restrict__ active,
                                                               const F* __restrict__ thr, int8_t* __restrict__ spins,
                                                               F* __restrict__ fields, int32_t* __restrict__ flips,
                                                               D* __restrict__ delta) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.y;
  const int c = g * 64 + lane;
  const int64_t kp = (int64_t)groups * 64;
  const bool on = (active[g] >> lane) & 1ull;
  const F t = thr[c];
  const int64_t nw = (int64_t)gridDim.x * SA_WAVES_PER_BLOCK;
  int32_t cnt = 0;
  D acc = 0;
  for (int64_t u = (int64_t)blockIdx.x * SA_WAVES_PER_BLOCK + wave; u < count; u += nw) {
    const int32_t i = guniform(nodes[u]);
    const int32_t s = spins[(int64_t)i * kp + c];
    const F h = fields[(int64_t)i * kp + c];
    const bool flip = on && (F)s * h < -t;
    if (!__any(flip)) continue;
    const int32_t snew = -s;
    if (flip) {
      spins[(int64_t)i * kp + c] = (int8_t)snew;
      ++cnt;
      acc += (D)2 * (D)s * (D)h;
    }
    const int32_t e0 = guniform(indptr[i]), e1 = guniform(indptr[i + 1]);
    for (int32_t e = e0; e < e1; ++e) {
      const int32_t j = guniform(indices[e]);
      const Src jv = guniform_val(jval[e]);
      if (flip) atomicAdd(&fields[(int64_t)j * kp + c], (F)2 * (F)jv * (F)snew);
    }
  }
  if (cnt) atomicAdd(&flips[c], cnt);
  if (acc != (D)0) add_to(&delta[c], acc);
}

// Tolerance of an energy (models/safe_update_policy.py:4-27, common.py:59-73): abs + rel * max(1, |E|)
// on the normalised energy E / scale; a flip changes the energy by 2 s h, so it passes iff
// s h < -tol * scale / 2. For exact integer fields the bound becomes floor(tol * scale / 2) (s h is an
// integer).
template <typename F, typename D>
__device__ inline F flip_threshold(D energy, double scale, double abs_tol, double rel_tol) {
  const double en = (double)energy / scale;
  const double x = (abs_tol + rel_tol * fmax(1.0, fabs(en))) * scale * 0.5;
  if constexpr (sizeof(F) == 4) return (F)fmin(floor(x), 2147483647.0);
  else return (F)x;
}

// Before a step: E[c] += delta[c], sweep[c] += flips[c], both cleared, thr[c] from the new E[c].
template <typename F, typename D>
__global__ __launch_bounds__(SA_BLOCK) void guard_class_prep_kernel(int64_t kp, double scale, double abs_tol,
                                                                    double rel_tol, D* __restrict__ energy,
                                                                    D* __restrict__ delta, int32_t* __restrict__ flips,
                                                                    int32_t* __restrict__ sweep, F* __restrict__ thr) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kp) return;
  const D e = energy[c] + delta[c];
  energy[c] = e;
  delta[c] = 0;
  sweep[c] += flips[c];
  flips[c] = 0;
  thr[c] = flip_threshold<F, D>(e, scale, abs_tol, rel_tol);
}

// ---- Sequential-Oracle ----------------------------------------------------------------------
constexpr int kSeqThreads = 1024;
constexpr int kSeqWaves = kSeqThreads / 64;

// Workgroup g runs candidate group g to a one-flip local optimum: per step, every lane (candidate)
// finds its node of smallest s_i h_i (first node on ties), and where 2 s h < -tol of its energy, flips
// it and updates its neighbours' fields (lane c alone touches column c). energy[c] (in field units,
// E = -sum s h / 2) is updated in place; steps[g] = steps with a flip (at most max_steps), or -1 if
// improving flips remain after max_steps steps.
template <typename Src, typename F, typename D>
__global__ __launch_bounds__(kSeqThreads) void guard_seq_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                                const int32_t* __restrict__ indices,
                                                                const Src* __restrict__ jval, int32_t k,
                                                                int32_t groups, int8_t* __restrict__ spins,
                                                                F* __restrict__ fields, D* __restrict__ energy,
                                                                double scale, double abs_tol, double rel_tol,
                                                                int64_t max_steps, int64_t* __restrict__ steps,
                                                                int32_t* __restrict__ flips) {
  __shared__ F best_v[kSeqWaves][64];
  __shared__ int32_t best_i[kSeqWaves][64];
  __shared__ int32_t any_flip;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.x;
  const int c = g * 64 + lane;
  const int64_t kp = (int64_t)groups * 64;
  int64_t step = 0;
  bool converged = false;
  int32_t nflips = 0;
  for (;;) {
    F bv = 0;
    int32_t bi = -1;
    for (int64_t i = wave; i < n; i += kSeqWaves) {
      const F v = (F)spins[i * kp + c] * fields[i * kp + c];
      if (bi < 0 || v < bv) {
        bv = v;
        bi = (int32_t)i;
      }
    }
    best_v[wave][lane] = bv;
    best_i[wave][lane] = bi;
    if (threadIdx.x == 0) any_flip = 0;
    __syncthreads();
    if (wave == 0) {
      F v = best_v[0][lane];
      int32_t at = best_i[0][lane];
      for (int w = 1; w < kSeqWaves; ++w) {
        const int32_t iw = best_i[w][lane];
        const F vw = best_v[w][lane];
        if (iw >= 0 && (at < 0 || vw < v || (vw == v && iw < at))) {
          v = vw;
          at = iw;
        }
      }
      const F t = flip_threshold<F, D>(energy[c], scale, abs_tol, rel_tol);
      const bool improving = c < k && at >= 0 && v < -t;
      const bool flip = improving && step < max_steps;
      if (flip) {
        const int32_t s = spins[(int64_t)at * kp + c];
        const int32_t snew = -s;
        spins[(int64_t)at * kp + c] = (int8_t)snew;
        energy[c] += (D)2 * (D)v;
        ++nflips;
        for (int32_t e = indptr[at]; e < indptr[at + 1]; ++e) {
          const int32_t j = indices[e];
          fields[(int64_t)j * kp + c] += (F)2 * (F)jval[e] * (F)snew;
        }
      }
      const bool any_improving = __any(improving), any_flipped = __any(flip);  // votes of the whole wave
      if (lane == 0) any_flip = any_improving ? (any_flipped ? 1 : 2) : 0;
    }
    __syncthreads();
    if (any_flip != 1) {  // 0: a one-flip optimum; 2: improving flips remain after max_steps steps
      converged = any_flip == 0;
      break;
    }
    ++step;
    __syncthreads();  // every wave has read any_flip before wave 0 clears it
  }
  if (wave == 0) {
    if (c < k) flips[c] = nflips;
    if (lane == 0) steps[g] = converged ? step : -1;
  }
}

template <typename Src, typename F, typename D>
static int guard_propose(const sa_csr* A, const Src* jval, int32_t k, int32_t groups, const int8_t* spins,
                         const F* fields, uint64_t seed, uint64_t round, const uint64_t* thresh, uint64_t* mask,
                         int32_t* counts, D* delta, hipStream_t stream, const char* what) {
  SA_REQUIRE(A && A->n > 0 && jval && spins && fields && thresh && mask && counts && delta, "bad arguments");
  SA_REQUIRE(k >= 1 && groups >= 1 && k <= 64 * groups && groups <= 1024, "bad candidate count");
  const dim3 grid(grid_for(A->n, SA_WAVES_PER_BLOCK * 4, 8192), groups);
  hipLaunchKernelGGL(guard_mask_kernel<F>, grid, dim3(SA_BLOCK), 0, stream, A->n, groups, spins, fields, seed, round,
                     thresh, mask, counts);
  int rc = check_launch(what);
  if (rc) return rc;
  hipLaunchKernelGGL((guard_delta_kernel<Src, F, D>), grid, dim3(SA_BLOCK), 0, stream, A->n, A->indptr, A->indices,
                     jval, groups, spins, fields, mask, delta);
  return check_launch(what);
}

template <typename Src, typename F>
static int guard_commit(const sa_csr* A, const Src* jval, int32_t groups, int8_t* spins, F* fields,
                        const uint64_t* mask, const uint64_t* accepted, hipStream_t stream, const char* what) {
  SA_REQUIRE(A && A->n > 0 && jval && spins && fields && mask && accepted && groups >= 1 && groups <= 1024,
             "bad arguments");
  hipLaunchKernelGGL((guard_commit_kernel<Src, F>), dim3(grid_for(A->n, SA_WAVES_PER_BLOCK * 4, 8192), groups),
                     dim3(SA_BLOCK), 0, stream, A->n, A->indptr, A->indices, jval, groups, spins, fields, mask,
                     accepted);
  return check_launch(what);
}

template <typename Src, typename F, typename D>
static int guard_class_step(const sa_csr* A, const Src* jval, int32_t groups, const int32_t* nodes, int64_t count,
                            const uint64_t* active, const F* thr, int8_t* spins, F* fields, int32_t* flips, D* delta,
                            hipStream_t stream, const char* what) {
  SA_REQUIRE(A && jval && nodes && active && thr && spins && fields && flips && delta && groups >= 1 &&
                 groups <= 1024,
             "bad arguments");
  if (count <= 0) return SA_OK;
  hipLaunchKernelGGL((guard_class_kernel<Src, F, D>), dim3(grid_for(count, SA_WAVES_PER_BLOCK * 2, 8192), groups),
                     dim3(SA_BLOCK), 0, stream, A->indptr, A->indices, jval, groups, nodes, count, active, thr, spins,
                     fields, flips, delta);
  return check_launch(what);
}

template <typename F, typename D>
static int guard_class_prep(int32_t groups, double scale, double abs_tol, double rel_tol, D* energy, D* delta,
                            int32_t* flips, int32_t* sweep, F* thr, hipStream_t stream, const char* what) {
  SA_REQUIRE(groups >= 1 && energy && delta && flips && sweep && thr && scale > 0.0 && abs_tol >= 0.0 &&
                 rel_tol >= 0.0,
             "bad arguments");
  const int64_t kp = (int64_t)groups * 64;
  hipLaunchKernelGGL((guard_class_prep_kernel<F, D>), dim3((unsigned)((kp + SA_BLOCK - 1) / SA_BLOCK)),
                     dim3(SA_BLOCK), 0, stream, kp, scale, abs_tol, rel_tol, energy, delta, flips, sweep, thr);
  return check_launch(what);
}

template <typename Src, typename F, typename D>
static int guard_sequential(const sa_csr* A, const Src* jval, int32_t k, int32_t groups, int8_t* spins, F* fields,
                            D* energy, double scale, double abs_tol, double rel_tol, int64_t max_steps,
                            int64_t* steps, int32_t* flips, hipStream_t stream, const char* what) {
  SA_REQUIRE(A && A->n > 0 && jval && spins && fields && energy && steps && flips && max_steps >= 0 &&
                 scale > 0.0,
             "bad arguments");
  SA_REQUIRE(k >= 1 && groups >= 1 && k <= 64 * groups && groups <= 1024, "bad candidate count");
  hipLaunchKernelGGL((guard_seq_kernel<Src, F, D>), dim3(groups), dim3(kSeqThreads), 0, stream, A->n, A->indptr,
                     A->indices, jval, k, groups, spins, fields, energy, scale, abs_tol, rel_tol, max_steps, steps,
                     flips);
  return check_launch(what);
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_guard_propose(const sa_csr* A, int32_t k, int32_t groups, const int8_t* spins, const int32_t* fields,
                     uint64_t seed, uint64_t round, const uint64_t* thresh, uint64_t* mask, int32_t* counts,
                     int64_t* delta, hipStream_t stream) {
  SA_REQUIRE(A && A->jint, "integer couplings required");
  return guard_propose(A, A->jint, k, groups, spins, fields, seed, round, thresh, mask, counts, delta, stream,
                       "guard_propose");
}

int sa_guard_propose_f64(const sa_csr* A, const float* fvals, int32_t k, int32_t groups, const int8_t* spins,
                         const double* fields, uint64_t seed, uint64_t round, const uint64_t* thresh, uint64_t* mask,
                         int32_t* counts, double* delta, hipStream_t stream) {
  return guard_propose(A, fvals, k, groups, spins, fields, seed, round, thresh, mask, counts, delta, stream,
                       "guard_propose_f64");
}

int sa_guard_commit(const sa_csr* A, int32_t groups, int8_t* spins, int32_t* fields, const uint64_t* mask,
                    const uint64_t* accepted, hipStream_t stream) {
  SA_REQUIRE(A && A->jint, "integer couplings required");
  return guard_commit(A, A->jint, groups, spins, fields, mask, accepted, stream, "guard_commit");
}

int sa_guard_commit_f64(const sa_csr* A, const float* fvals, int32_t groups, int8_t* spins, double* fields,
                        const uint64_t* mask, const uint64_t* accepted, hipStream_t stream) {
  return guard_commit(A, fvals, groups, spins, fields, mask, accepted, stream, "guard_commit_f64");
}

int sa_guard_class_step(const sa_csr* A, int32_t groups, const int32_t* class_nodes, int64_t count,
                        const uint64_t* active, const int32_t* thr, int8_t* spins, int32_t* fields, int32_t* flips,
                        int64_t* delta, hipStream_t stream) {
  SA_REQUIRE(A && A->jint, "integer couplings required");
  return guard_class_step(A, A->jint, groups, class_nodes, count, active, thr, spins, fields, flips, delta, stream,
                          "guard_class_step");
}

int sa_guard_class_step_f64(const sa_csr* A, const float* fvals, int32_t groups, const int32_t* class_nodes,
                            int64_t count, const uint64_t* active, const double* thr, int8_t* spins, double* fields,
                            int32_t* flips, double* delta, hipStream_t stream) {
  return guard_class_step(A, fvals, groups, class_nodes, count, active, thr, spins, fields, flips, delta, stream,
                          "guard_class_step_f64");
}

int sa_guard_class_prep(int32_t groups, double scale, double abs_tol, double rel_tol, int64_t* energy,
                        int64_t* delta, int32_t* flips, int32_t* sweep, int32_t* thr, hipStream_t stream) {
  return guard_class_prep(groups, scale, abs_tol, rel_tol, energy, delta, flips, sweep, thr, stream,
                          "guard_class_prep");
}

int sa_guard_class_prep_f64(int32_t groups, double abs_tol, double rel_tol, double* energy, double* delta,
                            int32_t* flips, int32_t* sweep, double* thr, hipStream_t stream) {
  return guard_class_prep(groups, 1.0, abs_tol, rel_tol, energy, delta, flips, sweep, thr, stream,
                          "guard_class_prep_f64");
}

int sa_guard_sequential(const sa_csr* A, int32_t k, int32_t groups, int8_t* spins, int32_t* fields, int64_t* energy,
                        double scale, double abs_tol, double rel_tol, int64_t max_steps, int64_t* steps,
                        int32_t* flips, hipStream_t stream) {
  SA_REQUIRE(A && A->jint, "integer couplings required");
  return guard_sequential(A, A->jint, k, groups, spins, fields, energy, scale, abs_tol, rel_tol, max_steps, steps,
                          flips, stream, "guard_sequential");
}

int sa_guard_sequential_f64(const sa_csr* A, const float* fvals, int32_t k, int32_t groups, int8_t* spins,
                            double* fields, double* energy, double abs_tol, double rel_tol, int64_t max_steps,
                            int64_t* steps, int32_t* flips, hipStream_t stream) {
  return guard_sequential(A, fvals, k, groups, spins, fields, energy, 1.0, abs_tol, rel_tol, max_steps, steps, flips,
                          stream, "guard_sequential_f64");
}

}  // extern "C"
