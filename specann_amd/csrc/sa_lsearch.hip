// Exact energies, the batched deterministic local search and the greedy colouring for gfx950.
//
// Local search layout (node-major, all candidates of a node side by side):
//   spins  int8  [n][kp]   kp = 64 * groups, candidate c in lane c % 64 of group c / 64
//   fields int32 [n][kp]   h_ic = sum_j J_ij s_jc, kept exact by integer updates
// A wavefront handles one (node, group): the node's row is read once (wave-uniform loads) for 64
// candidates, and the field updates of one neighbour are one coalesced 256-byte atomic.
//
// The flip order is a sequence of colour classes (independent sets). Within a class no two nodes
// are adjacent, so flipping all improving nodes of the class at once equals flipping them one at a
// time in any order: the parallel pass is a sequential single-flip descent, bit for bit.
#include <cstddef>

#include "sa_reduce.hpp"

namespace sa {

__device__ inline int32_t uniform(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// A wave-uniform 4-byte coupling (int32 or float) read as a scalar.
template <typename Src>
__device__ inline Src uniform_val(Src v) {
  static_assert(sizeof(Src) == 4, "4-byte couplings");
  return __builtin_bit_cast(Src, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int32_t, v)));
}

// ---------------------------------------------------------------------------------------------
// Exact energies of candidate bit vectors: e2[c] = sum_i s_i sum_j J_ij s_j.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(SA_BLOCK) void energy_bits_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                               const int32_t* __restrict__ indices,
                                                               const int32_t* __restrict__ jint,
                                                               const uint64_t* __restrict__ cand, int64_t nwords,
                                                               int64_t* __restrict__ e2) {
  __shared__ int64_t red[SA_WAVES_PER_BLOCK];
  const int c = blockIdx.y;
  const uint64_t* __restrict__ bits = cand + (int64_t)c * nwords;
  int64_t acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t h = 0;
    for (int32_t e = indptr[i]; e < indptr[i + 1]; ++e) {
      const int32_t j = stream_load(&indices[e]);
      const int32_t sj = ((bits[j >> 6] >> (j & 63)) & 1ull) ? 1 : -1;
      h += (int64_t)stream_load(&jint[e]) * sj;
    }
    const int64_t si = ((bits[i >> 6] >> (i & 63)) & 1ull) ? 1 : -1;
    acc += si * h;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0 && acc != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&e2[c]), (unsigned long long)acc);
}

// ---------------------------------------------------------------------------------------------
// Candidate-major bits -> node-major bits (word i*groups+g, bit l = candidate 64 g + l).
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(SA_BLOCK) void bits_transpose_kernel(int64_t n, int32_t k,
                                                                  const uint64_t* __restrict__ cand, int64_t nwords,
                                                                  int32_t groups, uint64_t* __restrict__ sb) {
  const int64_t total = n * groups;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = q / groups;
    const int g = (int)(q % groups);
    uint64_t word = 0;
    for (int l = 0; l < 64; ++l) {
      const int c = g * 64 + l;
      if (c >= k) break;
      word |= ((cand[(int64_t)c * nwords + (i >> 6)] >> (i & 63)) & 1ull) << l;
    }
    sb[q] = word;
  }
}

// ---------------------------------------------------------------------------------------------
// Fields and raw energies of the start candidates. blockIdx.y = group.
// ---------------------------------------------------------------------------------------------
// Src: the couplings the search adds (int32 J, or FP32 J / c); F: the fields (int32, or FP64, exact for
// FP32 couplings whose nonzero magnitudes are at least 2^-30 of the largest absolute degree, which the
// host checks: every partial sum is then a multiple of the smallest coupling's last bit below 2^53 of
// them). e2 may be NULL (the FP64 energies are summed in a fixed order by ls_energy_f64 instead).
template <typename Src, typename F>
__global__ __launch_bounds__(SA_BLOCK) void ls_init_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                           const int32_t* __restrict__ indices,
                                                           const Src* __restrict__ jval, int32_t k,
                                                           int32_t groups, const uint64_t* __restrict__ sb,
                                                           int8_t* __restrict__ spins, F* __restrict__ fields,
                                                           int64_t* __restrict__ e2) {
  __shared__ int64_t red[SA_WAVES_PER_BLOCK][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.y;
  const int c = g * 64 + lane;
  const int64_t kp = (int64_t)groups * 64;
  const int64_t nw = (int64_t)gridDim.x * SA_WAVES_PER_BLOCK;
  int64_t acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * SA_WAVES_PER_BLOCK + wave; i < n; i += nw) {
    const int32_t e0 = uniform(indptr[i]), e1 = uniform(indptr[i + 1]);
    F h = 0;
    int32_t e = e0;
    // eight neighbours' spin words in flight at a time (exact sums: the order does not matter)
    for (; e + 8 <= e1; e += 8) {
      int32_t j[8];
      Src jv[8];
      uint64_t w[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        j[u] = uniform(indices[e + u]);
        jv[u] = uniform_val(jval[e + u]);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) w[u] = sb[(int64_t)j[u] * groups + g];
#pragma unroll
      for (int u = 0; u < 8; ++u) h += ((w[u] >> lane) & 1ull) ? (F)jv[u] : -(F)jv[u];
    }
    for (; e < e1; ++e) {
      const int32_t j = uniform(indices[e]);
      const Src jv = uniform_val(jval[e]);
      const uint64_t w = sb[(int64_t)j * groups + g];
      h += ((w >> lane) & 1ull) ? (F)jv : -(F)jv;
    }
    const uint64_t wi = sb[i * groups + g];
    const int32_t si = ((wi >> lane) & 1ull) ? 1 : -1;
    spins[i * kp + c] = (int8_t)si;
    fields[i * kp + c] = h;
    if constexpr (sizeof(F) == 4) acc += (int64_t)si * (int64_t)h;
  }
  if (!e2) return;
  red[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && c < k) {
    const int64_t v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    if (v != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&e2[c]), (unsigned long long)v);
  }
}

// ---------------------------------------------------------------------------------------------
// One colour class of one pass.
// ---------------------------------------------------------------------------------------------
template <typename Src, typename F>
struct LsProblem {
  const int32_t* __restrict__ indptr;
  const int32_t* __restrict__ indices;
  const Src* __restrict__ jval;
  int32_t k, groups;
  int8_t* __restrict__ spins;
  F* __restrict__ fields;
};

// Spin and field accesses of the search: plain, or (kCoherent, for the persistent kernel, whose
// workgroups exchange them without a kernel boundary) relaxed atomics at device scope, which are
// coherent across the XCDs' L2 caches without the L2 write-back and invalidation a device-scope fence
// costs.
template <bool kCoherent, typename T>
__device__ inline T ls_load(const T* p) {
  if constexpr (kCoherent) return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *p;
}
template <bool kCoherent, typename T>
__device__ inline void ls_store(T* p, T v) {
  if constexpr (kCoherent) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}

// U nodes of one class for candidate group g, nodes[u0 + q * stride] for q < U (those below count),
// their loads issued together: a selected node whose spin disagrees with the sign of its field flips
// in the lanes c < k where it does, and its neighbours' fields take the change. No two nodes of a
// class are adjacent and the field updates are exact, so the order of the nodes does not matter.
// True in the lanes that flipped a node.
template <bool kCoherent, typename Src, typename F, int U>
__device__ inline bool class_step(const LsProblem<Src, F>& P, const int32_t* __restrict__ nodes, int64_t count,
                                  int64_t u0, int64_t stride, int g, uint64_t seed, uint64_t sel_id, uint64_t thresh,
                                  int32_t all, int lane) {
  const int c = g * 64 + lane;
  const int64_t kp = (int64_t)P.groups * 64;
  bool any_flip = false;
  int32_t node[U], s[U];
  F h[U];
  bool sel[U];
#pragma unroll
  for (int q = 0; q < U; ++q) {
    const int64_t u = u0 + q * stride;
    node[q] = uniform(nodes[u < count ? u : count - 1]);
    sel[q] = u < count && (all || (mix3(seed, sel_id, (uint64_t)node[q]) >> 11) < thresh);
  }
#pragma unroll
  for (int q = 0; q < U; ++q) {
    const int64_t off = (int64_t)node[q] * kp + c;
    s[q] = ls_load<kCoherent>(&P.spins[off]);
    h[q] = ls_load<kCoherent>(&P.fields[off]);
  }
#pragma unroll
  for (int q = 0; q < U; ++q) {
    const bool flip = sel[q] && (c < P.k) && ((s[q] > 0) ? (h[q] < (F)0) : (h[q] > (F)0));
    if (!__any(flip)) continue;
    const int32_t i = node[q];
    const int32_t snew = -s[q];
    if (flip) ls_store<kCoherent>(&P.spins[(int64_t)i * kp + c], (int8_t)snew);
    any_flip |= flip;
    const int32_t e0 = uniform(P.indptr[i]), e1 = uniform(P.indptr[i + 1]);
    int32_t e = e0;
    for (; e + 8 <= e1; e += 8) {
      int32_t j[8];
      Src jv[8];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        j[v] = uniform(P.indices[e + v]);
        jv[v] = uniform_val(P.jval[e + v]);
      }
      if (flip) {
#pragma unroll
        for (int v = 0; v < 8; ++v) atomicAdd(&P.fields[(int64_t)j[v] * kp + c], (F)2 * (F)jv[v] * (F)snew);
      }
    }
    for (; e < e1; ++e) {
      const int32_t j = uniform(P.indices[e]);
      const Src jv = uniform_val(P.jval[e]);
      if (flip) atomicAdd(&P.fields[(int64_t)j * kp + c], (F)2 * (F)jv * (F)snew);
    }
  }
  return any_flip;
}

constexpr int kClassU = 4;  // class nodes per wave and step

// blockIdx.y = group; the class's nodes are spread over the waves of the grid.
template <typename Src, typename F>
__global__ __launch_bounds__(SA_BLOCK) void ls_class_kernel(LsProblem<Src, F> P, const int32_t* __restrict__ nodes,
                                                            int64_t count, uint64_t seed, uint64_t pass,
                                                            uint64_t thresh, int32_t all, int32_t* __restrict__ flags) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.y;
  const int64_t nw = (int64_t)gridDim.x * SA_WAVES_PER_BLOCK;
  bool any_flip = false;
  for (int64_t u0 = (int64_t)blockIdx.x * SA_WAVES_PER_BLOCK + wave; u0 < count; u0 += nw * kClassU)
    any_flip |= class_step<false, Src, F, kClassU>(P, nodes, count, u0, nw, g, seed, pass, thresh, all, lane);
  if (any_flip) flags[g * 64 + lane] = 1;
}

// ---------------------------------------------------------------------------------------------
// The whole search in one cooperative launch. The grid (one 1024-thread workgroup per CU) walks the
// passes of the schedule (main passes, then the final rounds) and, in each, the colour classes in the
// pass's order, with a grid barrier after every class; between passes workgroup 0 reads the
// per-candidate flip flags, advances the schedule and writes the next pass's class order. Every
// workgroup reads the schedule after a barrier, so all of them take the same branches and meet at the
// same barriers. The barrier's wait is bounded in time: past the bound it records an error and every
// workgroup returns.
//
// Everything the workgroups exchange (spins, fields, flags, schedule, order) is accessed with device-
// scope atomics (the field updates already are), so a barrier needs only each wave's own operations
// complete (s_waitcnt) before it arrives: no cache write-back or invalidation.
//
// ctl (int32, zeroed by the caller; 128-byte lines of 32 entries): line 0 the schedule ([0] phase,
// [1] main pass index, [2] final round, [3] passes in the round, [4] main passes, [5] all passes,
// [6] fixed point reached, [7] error: 1 another workgroup's, 2 barrier timeout); line 1 the barrier's
// top arrival count, line 2 its generation, lines 3..10 the arrival counts of the eight workgroup
// groups (b mod 8); from kCtlHead the class order, then the candidates still active in the main phase
// (kp), then the flip flags of the pass (kp).
// ---------------------------------------------------------------------------------------------
constexpr int kCtlTop = 32, kCtlGen = 64, kCtlSub = 96;
constexpr int kCtlHead = kCtlSub + 8 * 32;
constexpr int kSearchThreads = 1024;
constexpr int kSearchWaves = kSearchThreads / 64;
constexpr int kMaxClasses = 1024;  // class-order keys are ranked in LDS (8 KB)
constexpr int32_t kPhaseMain = 0, kPhaseFinal = 1, kPhaseDone = 2;
constexpr uint64_t kOrderSalt = 0x0DE5C1A55C0FFEE1ull;
constexpr uint64_t kFinalPassBase = 1ull << 40;

struct LsSchedule {
  const int32_t* __restrict__ class_nodes;
  const int32_t* __restrict__ class_off;  // C + 1 offsets into class_nodes
  int32_t C;
  uint64_t seed;
  int32_t max_iterations, num_rounds, final_iters;
  uint64_t thresh_main, thresh_final;
  int32_t all_main, all_final;
  uint64_t timeout_ticks;  // of wall_clock64
};

__device__ inline int32_t ctl_load(const int32_t* p) { return ls_load<true>(p); }
__device__ inline void ctl_store(int32_t* p, int32_t v) { ls_store<true>(p, v); }
__device__ inline int32_t ctl_add(int32_t* p, int32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// This wave's memory operations complete, and the compiler keeps them on their side.
__device__ inline void ls_drain() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Grid barrier over gridDim.x workgroups (all resident: cooperative launch), two-level: a workgroup
// arrives at its group's count; the last of a group arrives at the top count; the last there opens
// the next generation. False when the search stopped on an error (then the caller returns).
__device__ bool ls_grid_sync(int32_t* ctl, uint64_t timeout_ticks) {
  __shared__ int32_t err;
  ls_drain();
  __syncthreads();
  if (threadIdx.x == 0) {
    const int32_t g = ctl_load(ctl + kCtlGen);
    const int grp = (int)(blockIdx.x & 7);
    const int32_t members = (int32_t)((gridDim.x - grp + 7) / 8);
    const int32_t ngroups = gridDim.x < 8 ? (int32_t)gridDim.x : 8;
    int32_t* sub = ctl + kCtlSub + 32 * grp;
    bool opened = false;
    if (ctl_add(sub, 1) + 1 == members) {
      ctl_store(sub, 0);
      ls_drain();
      if (ctl_add(ctl + kCtlTop, 1) + 1 == ngroups) {
        ctl_store(ctl + kCtlTop, 0);
        ls_drain();
        ctl_store(ctl + kCtlGen, g + 1);
        opened = true;
      }
    }
    if (!opened) {
      const uint64_t t0 = wall_clock64();
      while (ctl_load(ctl + kCtlGen) == g) {
        if (ctl_load(ctl + 7) != 0) break;
        if (wall_clock64() - t0 > timeout_ticks) {
          ctl_store(ctl + 7, 2);
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
    }
    err = ctl_load(ctl + 7);
  }
  __syncthreads();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return err == 0;
}

// Position of every class in pass q's order (ascending (key, class)), by workgroup 0.
__device__ void ls_write_order(const LsSchedule& S, uint64_t q, int32_t* order) {
  __shared__ uint64_t keys[kMaxClasses];
  for (int c = threadIdx.x; c < S.C; c += blockDim.x) keys[c] = mix3(S.seed ^ kOrderSalt, q, (uint64_t)c);
  __syncthreads();
  for (int c = threadIdx.x; c < S.C; c += blockDim.x) {
    const uint64_t kc = keys[c];
    int rank = 0;
    for (int d = 0; d < S.C; ++d) {
      const uint64_t kd = keys[d];
      rank += (kd < kc || (kd == kc && d < c)) ? 1 : 0;
    }
    ctl_store(&order[rank], c);
  }
  __syncthreads();
}

// End of a pass (workgroup 0): the schedule step of the host loop it replaces, the flags cleared, and
// the next pass's class order.
__device__ void ls_control(const LsSchedule& S, int32_t k, int32_t kp, int32_t* ctl) {
  int32_t* order = ctl + kCtlHead;
  int32_t* active = order + S.C;
  int32_t* flags = active + kp;
  const int32_t phase = ctl_load(ctl + 0);
  int anyf = 0, anya = 0;
  for (int c = threadIdx.x; c < k; c += blockDim.x) {
    const int32_t f = ctl_load(&flags[c]) != 0;
    anyf |= f;
    if (phase == kPhaseMain) {
      const int32_t a = active[c] & f;
      active[c] = a;
      anya |= a;
    }
  }
  anyf = __syncthreads_or(anyf);
  anya = __syncthreads_or(anya);
  for (int c = threadIdx.x; c < kp; c += blockDim.x) ctl_store(&flags[c], 0);
  __shared__ int32_t next;
  if (threadIdx.x == 0) {
    int32_t ph = phase, it = ctl_load(ctl + 1), r = ctl_load(ctl + 2), iters = ctl_load(ctl + 3);
    int32_t passes = ctl_load(ctl + 4), total = ctl_load(ctl + 5), fixed = ctl_load(ctl + 6);
    const bool finals = S.num_rounds > 0 && S.final_iters > 0;
    if (ph == kPhaseMain) {
      ++passes;
      ++it;
      if (!anya || it >= S.max_iterations) {
        ph = finals ? kPhaseFinal : kPhaseDone;
        r = 0;
        iters = 0;
      }
    } else {
      ++iters;
      bool end = false;
      if (!anyf) {
        fixed = S.all_final;  // a pass over every node without a flip is a fixed point of all later passes
        end = true;
      } else if (iters >= S.final_iters) {
        end = true;
      }
      if (end) {
        ++r;
        iters = 0;
        if (r >= S.num_rounds || fixed) ph = kPhaseDone;
      }
    }
    ++total;
    ctl_store(ctl + 0, ph);
    ctl_store(ctl + 1, it);
    ctl_store(ctl + 2, r);
    ctl_store(ctl + 3, iters);
    ctl_store(ctl + 4, passes);
    ctl_store(ctl + 5, total);
    ctl_store(ctl + 6, fixed);
    next = ph == kPhaseDone ? -1 : total;
  }
  __syncthreads();
  if (next >= 0) ls_write_order(S, (uint64_t)next, order);
}

template <typename Src, typename F>
__global__ __launch_bounds__(kSearchThreads) void ls_search_kernel(LsProblem<Src, F> P, LsSchedule S, int32_t* ctl) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int32_t kp = P.groups * 64;
  int32_t* order = ctl + kCtlHead;
  int32_t* active = order + S.C;
  int32_t* flags = active + kp;
  if (blockIdx.x == 0) {
    for (int c = threadIdx.x; c < kp; c += blockDim.x) active[c] = c < P.k ? 1 : 0;
    if (threadIdx.x == 0)
      ctl_store(ctl + 0, S.max_iterations > 0 ? kPhaseMain
                                              : (S.num_rounds > 0 && S.final_iters > 0 ? kPhaseFinal : kPhaseDone));
    ls_write_order(S, 0, order);
  }
  if (!ls_grid_sync(ctl, S.timeout_ticks)) return;
  const int64_t nw = (int64_t)gridDim.x * kSearchWaves;
  const int64_t gw = (int64_t)blockIdx.x * kSearchWaves + wave;
  const int64_t max_passes = (int64_t)S.max_iterations + (int64_t)S.num_rounds * S.final_iters;
  for (int64_t guard = 0; guard <= max_passes; ++guard) {
    const int32_t phase = ctl_load(ctl + 0);
    if (phase == kPhaseDone) break;
    const bool main = phase == kPhaseMain;
    const uint64_t sel_id = main ? (uint64_t)ctl_load(ctl + 1) : kFinalPassBase + (uint64_t)ctl_load(ctl + 2);
    const uint64_t thresh = main ? S.thresh_main : S.thresh_final;
    const int32_t all = main ? S.all_main : S.all_final;
    // a barrier after every class, empty or not: the schedule and the order read above are rewritten
    // only after the first of them
    for (int pos = 0; pos < S.C; ++pos) {
      const int32_t cls = ctl_load(&order[pos]);
      const int32_t lo = S.class_off[cls], count = S.class_off[cls + 1] - lo;
      const int32_t* nodes = S.class_nodes + lo;
      for (int g = 0; g < P.groups; ++g) {
        bool any_flip = false;
        for (int64_t u0 = gw; u0 < count; u0 += nw * kClassU)
          any_flip |= class_step<true, Src, F, kClassU>(P, nodes, count, u0, nw, g, S.seed, sel_id, thresh, all, lane);
        if (any_flip) ctl_store(&flags[g * 64 + lane], 1);
      }
      if (!ls_grid_sync(ctl, S.timeout_ticks)) return;
    }
    if (blockIdx.x == 0) ls_control(S, P.k, kp, ctl);
    if (!ls_grid_sync(ctl, S.timeout_ticks)) return;
  }
}

__global__ __launch_bounds__(SA_BLOCK) void ls_energy_kernel(int64_t n, int32_t k, int32_t groups,
                                                             const int8_t* __restrict__ spins,
                                                             const int32_t* __restrict__ fields,
                                                             int64_t* __restrict__ e2) {
  __shared__ int64_t red[SA_WAVES_PER_BLOCK][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.y;
  const int c = g * 64 + lane;
  const int64_t kp = (int64_t)groups * 64;
  const int64_t nw = (int64_t)gridDim.x * SA_WAVES_PER_BLOCK;
  int64_t acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * SA_WAVES_PER_BLOCK + wave; i < n; i += nw)
    acc += (int64_t)spins[i * kp + c] * (int64_t)fields[i * kp + c];
  red[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && c < k) {
    const int64_t v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    if (v != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&e2[c]), (unsigned long long)v);
  }
}

// Sum over the nodes of s_ic h_ic in FP64, in a fixed order: workgroup b sums its strided rows (waves in
// a fixed order) into part[b * kp + c]; ls_energy_f64_finish adds the kEnergyGrid partials in order.
constexpr int kEnergyGrid = 512;

__global__ __launch_bounds__(SA_BLOCK) void ls_energy_f64_kernel(int64_t n, int32_t groups,
                                                                 const int8_t* __restrict__ spins,
                                                                 const double* __restrict__ fields,
                                                                 double* __restrict__ part) {
  __shared__ double red[SA_WAVES_PER_BLOCK][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.y;
  const int c = g * 64 + lane;
  const int64_t kp = (int64_t)groups * 64;
  const int64_t nw = (int64_t)gridDim.x * SA_WAVES_PER_BLOCK;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * SA_WAVES_PER_BLOCK + wave; i < n; i += nw)
    acc += (double)spins[i * kp + c] * fields[i * kp + c];
  red[wave][lane] = acc;
  __syncthreads();
  if (wave == 0) part[(int64_t)blockIdx.x * kp + c] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

__global__ __launch_bounds__(SA_BLOCK) void ls_energy_f64_finish(int32_t k, int64_t kp, int32_t nparts,
                                                                 const double* __restrict__ part,
                                                                 double* __restrict__ e2) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  double acc = 0.0;
  for (int b = 0; b < nparts; ++b) acc += part[(int64_t)b * kp + c];
  e2[c] = acc;
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void ls_extract_kernel(int64_t n, int32_t groups, int32_t c,
                                                              const int8_t* __restrict__ spins, T* __restrict__ out) {
  const int64_t kp = (int64_t)groups * 64;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (T)spins[i * kp + c];
}

// ---------------------------------------------------------------------------------------------
// Greedy colouring over a worklist (see the header comment of sa_colour_start). Priority order:
// degree, then hash(seed, i), then i, all descending; the hash is computed only on degree ties.
// ---------------------------------------------------------------------------------------------
// hi caches hash(seed, i): computed on the first degree tie only (hi_ready). kIndexTies: ties of degree
// go to the smaller index instead (largest-first order, local_search_guards/common.py:136-159).
template <bool kIndexTies>
__device__ inline bool outranks(uint64_t seed, int32_t j, int32_t dj, int32_t i, int32_t di, uint64_t& hi,
                                bool& hi_ready) {
  if (dj != di) return dj > di;
  if constexpr (kIndexTies) return j < i;
  if (!hi_ready) {
    hi = mix3(seed, 0xC0102ull, (uint64_t)i);
    hi_ready = true;
  }
  const uint64_t hj = mix3(seed, 0xC0102ull, (uint64_t)j);
  if (hj != hi) return hj > hi;
  return j > i;
}

__global__ __launch_bounds__(SA_BLOCK) void colour_start_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                                int32_t* __restrict__ colour,
                                                                int32_t* __restrict__ list0, int32_t* __restrict__ resume,
                                                                int32_t* __restrict__ cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    colour[i] = -1;
    list0[i] = (int32_t)i;
    resume[i] = indptr[i];
  }
  if (blockIdx.x == 0 && threadIdx.x < 3) cnt[threadIdx.x] = threadIdx.x == 0 ? (int32_t)n : 0;
}

// Round r: the nodes of list r % 2 (cnt[r % 3] of them). A node whose higher-priority neighbours are
// all coloured takes the smallest colour none of them has (its lower-priority neighbours wait for it,
// so this is its colour in the sequential greedy order, whatever this round's timing); the others go
// to list (r + 1) % 2, counted in cnt[(r + 1) % 3], with one atomic per wave. cnt[(r + 2) % 3] is
// cleared for round r + 1's successor. A waiting node keeps in resume[i] the neighbour that blocked
// it: the neighbours before it were coloured or of lower priority, and a coloured node stays coloured,
// so the next round starts its check there (every adjacency entry is checked about once overall).
// From round 1 on every listed node has such a blocker, which outranks it for good (priorities are
// fixed): while the blocker is uncoloured the node only re-reads its colour.
constexpr int32_t kColourHeavy = 128;  // nodes with more neighbours are checked and coloured by a whole wave

// OR of v over the wave's 64 lanes (every lane receives it).
__device__ inline uint64_t wave_or(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

// One round for a node with many neighbours, by the whole wave (the arguments are uniform): the first
// neighbour from resume[i] on that blocks it (uncoloured and of higher priority) becomes resume[i] and
// the node waits (true); otherwise it takes the smallest colour none of its neighbours has (false).
// The same outcome as the one-lane scan, with 64 neighbours per step.
template <bool kIndexTies>
__device__ inline bool colour_node_wave(const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                        uint64_t seed, int32_t* colour, int32_t* __restrict__ resume, int32_t i,
                                        int32_t r, int lane) {
  const int32_t e0 = indptr[i], e1 = indptr[i + 1];
  const int32_t di = e1 - e0;
  int32_t start = resume[i];
  if (r > 0) {
    const int32_t jb = indices[start];
    if (__hip_atomic_load(&colour[jb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0) return true;
    ++start;
  }
  uint64_t hi = 0;
  bool hi_ready = false;
  // 256 neighbours per step, four per lane with their loads in flight together (indices clamped to the
  // list), the four groups of 64 checked in order
  for (int32_t b = start; b < e1; b += 256) {
    int32_t j[4], cj[4], dj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) j[u] = indices[min(b + u * 64 + lane, e1 - 1)];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      cj[u] = __hip_atomic_load(&colour[j[u]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      dj[u] = indptr[j[u] + 1] - indptr[j[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int32_t e = b + u * 64 + lane;
      const bool blk =
          e < e1 && j[u] != i && cj[u] < 0 && outranks<kIndexTies>(seed, j[u], dj[u], i, di, hi, hi_ready);
      const uint64_t bm = __ballot(blk);
      if (bm) {
        if (lane == 0) resume[i] = b + u * 64 + (int32_t)__builtin_ctzll(bm);
        return true;
      }
    }
  }
  int32_t col = -1;
  for (int32_t base = 0; col < 0; base += 64) {
    uint64_t used = 0;
    for (int32_t e = e0 + lane; e < e1; e += 256) {
      int32_t j[4], cj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) j[u] = indices[min(e + u * 64, e1 - 1)];
#pragma unroll
      for (int u = 0; u < 4; ++u) cj[u] = __hip_atomic_load(&colour[j[u]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (e + u * 64 < e1 && j[u] != i && cj[u] >= base && cj[u] < base + 64) used |= 1ull << (cj[u] - base);
      }
    }
    used = wave_or(used);
    if (~used) col = base + (int32_t)__builtin_ctzll(~used);
  }
  if (lane == 0) __hip_atomic_store(&colour[i], col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return false;
}

template <bool kIndexTies>
__global__ __launch_bounds__(SA_BLOCK) void colour_list_kernel(const int32_t* __restrict__ indptr,
                                                               const int32_t* __restrict__ indices, uint64_t seed,
                                                               int32_t* colour, const int32_t* __restrict__ in,
                                                               int32_t* __restrict__ out, int32_t* __restrict__ resume,
                                                               int32_t* cnt, int32_t r) {
  const int64_t count = cnt[r % 3];
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt[(r + 2) % 3] = 0;
  const int lane = threadIdx.x & 63;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); t0 < count; t0 += stride) {
    const int64_t t = t0 + lane;
    bool pending = false, heavy = false;
    int32_t i = 0;
    if (t < count) {
      i = in[t];
      const int32_t e0 = indptr[i], e1 = indptr[i + 1];
      const int32_t di = e1 - e0;
      heavy = di > kColourHeavy;
      if (!heavy) {
        int32_t e = resume[i];
        if (r > 0) {
          const int32_t jb = indices[e];
          if (__hip_atomic_load(&colour[jb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0) pending = true;
          else ++e;
        }
        uint64_t hi = 0;
        bool hi_ready = false;
        // four neighbours' loads in flight at a time (indices clamped to the list), checked in order
        for (; e < e1 && !pending; e += 4) {
          int32_t j[4], cj[4], dj[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) j[u] = indices[min(e + u, e1 - 1)];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            cj[u] = __hip_atomic_load(&colour[j[u]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            dj[u] = indptr[j[u] + 1] - indptr[j[u]];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (!pending && e + u < e1 && j[u] != i && cj[u] < 0 &&
                outranks<kIndexTies>(seed, j[u], dj[u], i, di, hi, hi_ready)) {
              pending = true;
              resume[i] = e + u;
            }
          }
        }
        if (!pending) {
          int32_t col = -1;
          for (int32_t base = 0; col < 0; base += 64) {
            uint64_t used = 0;
            for (int32_t e = e0; e < e1; e += 4) {
              int32_t j[4], cj[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) j[u] = indices[min(e + u, e1 - 1)];
#pragma unroll
              for (int u = 0; u < 4; ++u)
                cj[u] = __hip_atomic_load(&colour[j[u]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                if (e + u < e1 && j[u] != i && cj[u] >= base && cj[u] < base + 64) used |= 1ull << (cj[u] - base);
              }
            }
            if (~used) col = base + __builtin_ctzll(~used);
          }
          __hip_atomic_store(&colour[i], col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    // the wave's heavy nodes, one at a time, by all 64 lanes
    uint64_t hm = __ballot(heavy);
    while (hm) {
      const int l = __builtin_ctzll(hm);
      hm &= hm - 1;
      const int32_t hn = __shfl(i, l);
      const bool hp = colour_node_wave<kIndexTies>(indptr, indices, seed, colour, resume, hn, r, lane);
      if (lane == l) pending = hp;
    }
    const uint64_t m = __ballot(pending);
    if (m) {
      int32_t at = 0;
      if (lane == 0) at = atomicAdd(&cnt[(r + 1) % 3], (int32_t)__popcll(m));
      at = __shfl(at, 0);
      if (pending) out[at + (int32_t)__popcll(m & ((1ull << lane) - 1))] = i;
    }
  }
}

template <bool kIndexTies>
static int colour_rounds(const sa_csr* A, uint64_t seed, int32_t* colour, int32_t* lists, int32_t* counts,
                         int32_t first_round, int32_t rounds, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && A->indices && colour && lists && counts, "bad arguments");
  SA_REQUIRE(first_round >= 0 && rounds >= 0, "bad round numbers");
  const int64_t n = A->n;
  const int G = grid_for(n, SA_BLOCK * 4, SA_RED_GRID);
  for (int32_t q = 0; q < rounds; ++q) {
    const int32_t r = first_round + q;
    hipLaunchKernelGGL(colour_list_kernel<kIndexTies>, dim3(G), dim3(SA_BLOCK), 0, stream, A->indptr, A->indices,
                       seed, colour, (const int32_t*)(lists + (r & 1) * n), lists + ((r + 1) & 1) * n,
                       lists + 2 * n, counts, r);
    const int rc = check_launch("colour_rounds");
    if (rc) return rc;
  }
  return 0;
}

constexpr uint64_t kSearchTimeoutSec = 30;  // longest wait at one grid barrier before the search gives up

// The persistent search: one cooperative launch (every workgroup resident), one 1024-thread workgroup
// per CU.
template <typename Src, typename F>
static int ls_search_launch(const LsProblem<Src, F>& P, const int32_t* class_nodes, const int32_t* class_off,
                            int32_t num_classes, uint64_t seed, int32_t max_iterations, double subset_fraction,
                            int32_t num_rounds, int32_t final_sweep_iters, double final_sweep_fraction, int32_t* ctl,
                            hipStream_t stream, const char* what) {
  SA_REQUIRE(class_nodes && class_off && ctl, "bad arguments");
  SA_REQUIRE(P.k >= 1 && P.groups >= 1 && P.k <= 64 * P.groups && P.groups <= 1024, "bad candidate count");
  SA_REQUIRE(num_classes >= 1 && num_classes <= kMaxClasses, "num_classes must be in [1, 1024]");
  SA_REQUIRE(max_iterations >= 0 && num_rounds >= 0 && final_sweep_iters >= 0, "negative pass counts");
  SA_REQUIRE(subset_fraction > 0.0 && subset_fraction <= 1.0 && final_sweep_fraction > 0.0 &&
                 final_sweep_fraction <= 1.0,
             "fractions must be in (0, 1]");
  int dev = 0, coop = 0, cus = 0, rate = 0, per = 0;
  const void* kf = reinterpret_cast<const void*>(&ls_search_kernel<Src, F>);
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ||
      hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, dev) ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kf, kSearchThreads, 0)) {
    set_error(std::string(what) + ": device query failed");
    return SA_E_LAUNCH;
  }
  if (!coop || cus < 1 || rate <= 0 || per < 1) {
    set_error(std::string(what) + ": the device cannot run the cooperative search kernel");
    return SA_E_LAUNCH;
  }
  LsSchedule S{};
  S.class_nodes = class_nodes;
  S.class_off = class_off;
  S.C = num_classes;
  S.seed = seed;
  S.max_iterations = max_iterations;
  S.num_rounds = num_rounds;
  S.final_iters = final_sweep_iters;
  S.all_main = subset_fraction >= 1.0 ? 1 : 0;
  S.all_final = final_sweep_fraction >= 1.0 ? 1 : 0;
  S.thresh_main = (uint64_t)(subset_fraction * 9007199254740992.0);
  S.thresh_final = (uint64_t)(final_sweep_fraction * 9007199254740992.0);
  S.timeout_ticks = (uint64_t)rate * 1000ull * kSearchTimeoutSec;
  LsProblem<Src, F> Pa = P;
  void* args[] = {reinterpret_cast<void*>(&Pa), reinterpret_cast<void*>(&S), reinterpret_cast<void*>(&ctl)};
  // one workgroup per CU (16 of its 32 wave slots, 8 KB of LDS): every workgroup stays resident while
  // the others wait at a barrier
  const hipError_t e = hipLaunchCooperativeKernel(kf, dim3(cus), dim3(kSearchThreads), args, 0, stream);
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return SA_E_LAUNCH;
  }
  return check_launch(what);
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_energy_bits_i64(const sa_csr* A, int32_t k, const uint64_t* cand_bits, int64_t nwords, int64_t* e2,
                       hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->jint && cand_bits && e2 && k >= 1 && k <= 65535, "bad arguments");
  SA_REQUIRE(nwords * 64 >= A->n, "nwords too small for n");
  hipLaunchKernelGGL(energy_bits_kernel, dim3(grid_for(A->n, SA_BLOCK * 4, 2048), k), dim3(SA_BLOCK), 0, stream,
                     A->n, A->indptr, A->indices, A->jint, cand_bits, nwords, e2);
  return check_launch("energy_bits_i64");
}

int sa_bits_transpose(int64_t n, int32_t k, const uint64_t* cand_bits, int64_t nwords, int32_t groups, uint64_t* sb,
                      hipStream_t stream) {
  SA_REQUIRE(n > 0 && k >= 1 && groups >= 1 && k <= 64 * groups && cand_bits && sb, "bad arguments");
  SA_REQUIRE(nwords * 64 >= n, "nwords too small for n");
  hipLaunchKernelGGL(bits_transpose_kernel, dim3(grid_for(n * groups, SA_BLOCK, 4096)), dim3(SA_BLOCK), 0, stream, n,
                     k, cand_bits, nwords, groups, sb);
  return check_launch("bits_transpose");
}

int sa_ls_init(const sa_csr* A, int32_t k, int32_t groups, const uint64_t* sb, int8_t* spins, int32_t* fields,
               int64_t* e2, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->jint && sb && spins && fields && e2, "bad arguments");
  SA_REQUIRE(k >= 1 && groups >= 1 && k <= 64 * groups && groups <= 1024, "bad candidate count");
  hipLaunchKernelGGL((ls_init_kernel<int32_t, int32_t>), dim3(grid_for(A->n, SA_WAVES_PER_BLOCK * 8, 4096), groups),
                     dim3(SA_BLOCK), 0, stream, A->n, A->indptr, A->indices, A->jint, k, groups, sb, spins, fields, e2);
  return check_launch("ls_init");
}

int sa_ls_class_pass(const sa_csr* A, int32_t k, int32_t groups, const int32_t* class_nodes, int64_t count,
                     uint64_t seed, int64_t pass, double frac, int8_t* spins, int32_t* fields, int32_t* flags,
                     hipStream_t stream) {
  SA_REQUIRE(A && A->jint && class_nodes && spins && fields && flags && pass >= 0, "bad arguments");
  SA_REQUIRE(k >= 1 && groups >= 1 && k <= 64 * groups && groups <= 1024, "bad candidate count");
  SA_REQUIRE(frac > 0.0 && frac <= 1.0, "frac must be in (0, 1]");
  if (count <= 0) return SA_OK;
  const int32_t all = frac >= 1.0 ? 1 : 0;
  const uint64_t thresh = (uint64_t)(frac * 9007199254740992.0);
  const LsProblem<int32_t, int32_t> P{A->indptr, A->indices, A->jint, k, groups, spins, fields};
  hipLaunchKernelGGL((ls_class_kernel<int32_t, int32_t>), dim3(grid_for(count, SA_WAVES_PER_BLOCK * 2, 8192), groups),
                     dim3(SA_BLOCK), 0, stream, P, class_nodes, count, seed, (uint64_t)pass, thresh, all, flags);
  return check_launch("ls_class_pass");
}

int sa_ls_init_f64(const sa_csr* A, const float* fvals, int32_t k, int32_t groups, const uint64_t* sb,
                   int8_t* spins, double* fields, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && fvals && sb && spins && fields, "bad arguments");
  SA_REQUIRE(k >= 1 && groups >= 1 && k <= 64 * groups && groups <= 1024, "bad candidate count");
  hipLaunchKernelGGL((ls_init_kernel<float, double>), dim3(grid_for(A->n, SA_WAVES_PER_BLOCK * 8, 4096), groups),
                     dim3(SA_BLOCK), 0, stream, A->n, A->indptr, A->indices, fvals, k, groups, sb, spins, fields,
                     (int64_t*)nullptr);
  return check_launch("ls_init_f64");
}

int sa_ls_class_pass_f64(const sa_csr* A, const float* fvals, int32_t k, int32_t groups, const int32_t* class_nodes,
                         int64_t count, uint64_t seed, int64_t pass, double frac, int8_t* spins, double* fields,
                         int32_t* flags, hipStream_t stream) {
  SA_REQUIRE(A && fvals && class_nodes && spins && fields && flags && pass >= 0, "bad arguments");
  SA_REQUIRE(k >= 1 && groups >= 1 && k <= 64 * groups && groups <= 1024, "bad candidate count");
  SA_REQUIRE(frac > 0.0 && frac <= 1.0, "frac must be in (0, 1]");
  if (count <= 0) return SA_OK;
  const int32_t all = frac >= 1.0 ? 1 : 0;
  const uint64_t thresh = (uint64_t)(frac * 9007199254740992.0);
  const LsProblem<float, double> P{A->indptr, A->indices, fvals, k, groups, spins, fields};
  hipLaunchKernelGGL((ls_class_kernel<float, double>), dim3(grid_for(count, SA_WAVES_PER_BLOCK * 2, 8192), groups),
                     dim3(SA_BLOCK), 0, stream, P, class_nodes, count, seed, (uint64_t)pass, thresh, all, flags);
  return check_launch("ls_class_pass_f64");
}

size_t sa_ls_search_ctl_ints(int32_t num_classes, int32_t groups) {
  if (num_classes < 1 || groups < 1) return 0;
  return (size_t)kCtlHead + (size_t)num_classes + 2 * (size_t)groups * 64;
}

int sa_ls_search(const sa_csr* A, int32_t k, int32_t groups, const int32_t* class_nodes, const int32_t* class_off,
                 int32_t num_classes, uint64_t seed, int32_t max_iterations, double subset_fraction, int32_t num_rounds,
                 int32_t final_sweep_iters, double final_sweep_fraction, int8_t* spins, int32_t* fields, int32_t* ctl,
                 hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->jint && spins && fields, "bad arguments");
  const LsProblem<int32_t, int32_t> P{A->indptr, A->indices, A->jint, k, groups, spins, fields};
  return ls_search_launch(P, class_nodes, class_off, num_classes, seed, max_iterations, subset_fraction, num_rounds,
                          final_sweep_iters, final_sweep_fraction, ctl, stream, "ls_search");
}

int sa_ls_search_f64(const sa_csr* A, const float* fvals, int32_t k, int32_t groups, const int32_t* class_nodes,
                     const int32_t* class_off, int32_t num_classes, uint64_t seed, int32_t max_iterations,
                     double subset_fraction, int32_t num_rounds, int32_t final_sweep_iters,
                     double final_sweep_fraction, int8_t* spins, double* fields, int32_t* ctl, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && fvals && spins && fields, "bad arguments");
  const LsProblem<float, double> P{A->indptr, A->indices, fvals, k, groups, spins, fields};
  return ls_search_launch(P, class_nodes, class_off, num_classes, seed, max_iterations, subset_fraction, num_rounds,
                          final_sweep_iters, final_sweep_fraction, ctl, stream, "ls_search_f64");
}

size_t sa_ls_energy_f64_scratch(int32_t groups) { return (size_t)kEnergyGrid * (size_t)groups * 64 * sizeof(double); }

int sa_ls_energy_f64(int64_t n, int32_t k, int32_t groups, const int8_t* spins, const double* fields, double* part,
                     double* e2, hipStream_t stream) {
  SA_REQUIRE(n > 0 && spins && fields && part && e2 && k >= 1 && groups >= 1 && k <= 64 * groups, "bad arguments");
  hipLaunchKernelGGL(ls_energy_f64_kernel, dim3(kEnergyGrid, groups), dim3(SA_BLOCK), 0, stream, n, groups, spins,
                     fields, part);
  int rc = check_launch("ls_energy_f64");
  if (rc) return rc;
  hipLaunchKernelGGL(ls_energy_f64_finish, dim3((k + SA_BLOCK - 1) / SA_BLOCK), dim3(SA_BLOCK), 0, stream, k,
                     (int64_t)groups * 64, kEnergyGrid, part, e2);
  return check_launch("ls_energy_f64_finish");
}

int sa_ls_energy(int64_t n, int32_t k, int32_t groups, const int8_t* spins, const int32_t* fields, int64_t* e2,
                 hipStream_t stream) {
  SA_REQUIRE(n > 0 && spins && fields && e2 && k >= 1 && groups >= 1 && k <= 64 * groups, "bad arguments");
  hipLaunchKernelGGL(ls_energy_kernel, dim3(grid_for(n, SA_WAVES_PER_BLOCK * 16, 2048), groups), dim3(SA_BLOCK), 0,
                     stream, n, k, groups, spins, fields, e2);
  return check_launch("ls_energy");
}

int sa_ls_extract_f32(int64_t n, int32_t groups, int32_t c, const int8_t* spins, float* out, hipStream_t stream) {
  SA_REQUIRE(n > 0 && spins && out && c >= 0 && c < 64 * groups, "bad arguments");
  hipLaunchKernelGGL(ls_extract_kernel<float>, dim3(grid_for(n, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK), 0, stream, n,
                     groups, c, spins, out);
  return check_launch("ls_extract_f32");
}

int sa_ls_extract_f64(int64_t n, int32_t groups, int32_t c, const int8_t* spins, double* out, hipStream_t stream) {
  SA_REQUIRE(n > 0 && spins && out && c >= 0 && c < 64 * groups, "bad arguments");
  hipLaunchKernelGGL(ls_extract_kernel<double>, dim3(grid_for(n, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK), 0, stream, n,
                     groups, c, spins, out);
  return check_launch("ls_extract_f64");
}

int sa_colour_start(const sa_csr* A, int32_t* colour, int32_t* lists, int32_t* counts, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && colour && lists && counts, "bad arguments");
  hipLaunchKernelGGL(colour_start_kernel, dim3(grid_for(A->n, SA_BLOCK, 4096)), dim3(SA_BLOCK), 0, stream, A->n,
                     A->indptr, colour, lists, lists + 2 * A->n, counts);
  return check_launch("colour_start");
}

int sa_colour_rounds(const sa_csr* A, uint64_t seed, int32_t* colour, int32_t* lists, int32_t* counts,
                     int32_t first_round, int32_t rounds, hipStream_t stream) {
  return colour_rounds<false>(A, seed, colour, lists, counts, first_round, rounds, stream);
}

int sa_colour_rounds_lf(const sa_csr* A, int32_t* colour, int32_t* lists, int32_t* counts, int32_t first_round,
                        int32_t rounds, hipStream_t stream) {
  return colour_rounds<true>(A, 0, colour, lists, counts, first_round, rounds, stream);
}

}  // extern "C"
