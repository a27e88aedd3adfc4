// LDS-blocked SpMV for gfx950: the gathers of the fused scaled SpMV served from LDS.
//
// On a graph without locality every gather of g costs an L2 cache-line request, and the CSR kernels
// are bound by the number of such requests a CU can keep in flight. This layout trades them for
// streaming: workgroup w = rb * 8 + p runs on XCD p (workgroups reach the XCDs round-robin), owns
// row block rb (brows rows) of column panel p, and walks the panel's column blocks (bcols columns)
// in order. For each block it copies that slice of g into LDS with 16-byte loads (the panel's slice
// of g stays in the XCD's L2 across the row blocks), then every entry of the tile gathers from LDS
// and adds into the row sums, which also live in LDS for the whole walk. The row sums go to
// partial[p * pld + r] like the XCD-panel kernel's, and the same consumers combine the panels.
//
// A row's entries in one tile (a run) sit in one 64-entry chunk, with the cont bit set on all but
// the last; a wave loads one chunk, reduces each run onto its first lane (a segmented reduction over
// the cont mask, log2 of the longest run steps) and that lane alone adds the run's sum to its row.
// No two lanes of a tile write one row, and tiles are separated by barriers, so the sums need no
// atomics and run in a fixed order: results are bitwise reproducible. Padding entries (local row =
// brows) add into a spare slot that is never read.
#include "sa_spmv.hpp"

namespace sa {

constexpr int kBThreads = 1024;                // 16 waves: one workgroup per CU (all of LDS)
constexpr int kBWaves = kBThreads / 64;
constexpr int kBTable = 376;                   // tile-bound table entries, at the end of LDS (fills it)
constexpr int kBMaxTiles = kBTable - 8;        // column blocks per panel
constexpr int kBPre = 4;                       // chunks of the next tile each wave loads ahead

template <typename T, bool kSigned>
__device__ inline T block_entry(uint32_t e, T val, const T* xb) {
  const T x = xb[(e >> 15) & 0x7FFF];
  if constexpr (kSigned) {
    return (e & 0x80000000u) ? -x : x;
  } else {
    return val * x;
  }
}

// v of lane + 1 (0 on lane 63): a DPP wave_shl:1 move, no LDS round trip.
__device__ inline float next_lane(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}
__device__ inline double next_lane(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, 0x130, 0xf, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), 0x130, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// NC chunks of one tile (on[u]: chunk u exists), in the order the wave takes them: lane values v
// (already gathered) are reduced run by run onto each run's first lane (a segmented reduction over the
// cont mask: lane l continues into l + 1; runs of two take one step, longer ones log2 of their
// length), and that lane adds the run sum into its row. A row's entries in one tile sit in one chunk,
// or, for a run longer than 64, in pieces in chunks that only this wave takes; no other wave touches
// the row while the tile is summed, so a plain read-modify-write per chunk, in chunk order, suffices
// (an LDS atomic add is several times slower here).
template <typename T, int NC>
__device__ inline void block_chunks_add(const uint32_t* e, T* v, const bool* on, T* yacc, int lane) {
  uint64_t m[NC];
#pragma unroll
  for (int u = 0; u < NC; ++u) m[u] = __ballot((e[u] >> 30) & 1u);
#pragma unroll
  for (int u = 0; u < NC; ++u) {
    if (on[u] && m[u]) {
      const T o = next_lane(v[u]);
      if ((m[u] >> lane) & 1ull) v[u] += o;
      uint64_t c = m[u] & (m[u] >> 1);
      for (int d = 2; c != 0; d <<= 1) {
        const T w = __shfl_down(v[u], d, 64);
        if ((c >> lane) & 1ull) v[u] += w;
        c &= c >> d;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NC; ++u) {
    const bool head = on[u] && (lane == 0 || !((m[u] >> (lane - 1)) & 1ull));
    if (head) {
      T* y = &yacc[e[u] & 0x7FFFu];
      *y += v[u];
    }
  }
}

// LDS split of one workgroup: row sums for up to kRows rows (and the padding slot), the column block
// (kCols entries) and the tile-bound table. One workgroup per CU.
template <typename T>
struct BlockCfg;
template <>
struct BlockCfg<float> {
  static constexpr int kRows = 26240, kCols = 14336;
};
template <>
struct BlockCfg<double> {
  static constexpr int kRows = 13312, kCols = 6976;
};
constexpr int kBXv = 4;  // 16-byte loads per thread for one column block: kCols <= kBXv * 16 KB

// Workgroup barrier for the walk: this wave's LDS operations complete (lgkmcnt 0) and at most N
// vector-memory operations still in flight (vmcnt N), then s_barrier. The wavefront-scope fences only
// keep the compiler from moving memory operations across the barrier.
template <int N>
__device__ inline void walk_barrier() {
  static_assert(N >= 0 && N < 64, "vmcnt out of range");
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4));
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The walk over the column blocks is software-pipelined: while tile c is summed, the entries of tile
// c + 1 are loaded into registers, and after it the column block of tile c + 2 (two register sets of
// blocks alternate); a block goes to LDS between the two barriers of its tile. The walk is unrolled by two so that the two register sets of
// entries never move between registers (a move would make their loads complete early). Every load is
// issued unconditionally (out-of-range lanes read a valid stand-in address and drop the value), so
// the loads in flight are counted exactly and each wait covers only the loads it needs.
template <typename T, bool kSigned>
__global__ __launch_bounds__(kBThreads) void spmv_block_kernel(SpmvParams<T> p) {
  typedef typename vec16<T>::type V;
  constexpr int kRows = BlockCfg<T>::kRows, kCols = BlockCfg<T>::kCols;
  constexpr int L = vec16<T>::L;
  __shared__ __attribute__((aligned(16))) T yacc[kRows + 4];
  __shared__ __attribute__((aligned(16))) T xb[kCols];
  __shared__ int32_t tb[kBTable];
  if (p.gate && *p.gate != SA_RUNNING) return;
  const int panel = (int)(blockIdx.x & 7);
  const int rb = (int)(blockIdx.x >> 3);
  const int RB = p.brows, CB = p.bcols;
  const int64_t r0 = p.rbound ? (int64_t)p.rbound[rb] : (int64_t)rb * RB;
  if (r0 >= p.n) return;
  const int64_t r1 = p.rbound ? (int64_t)p.rbound[rb + 1] : r0 + RB;
  const int rows = (int)((p.n < r1 ? p.n : r1) - r0);  // at most brows (checked when the layout is built)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  for (int i = tid; i <= RB; i += kBThreads) yacc[i] = (T)0;

  const int64_t c0 = p.pbound[panel];
  const int64_t width = (int64_t)p.pbound[panel + 1] - c0;
  const int nc = (int)min((width + CB - 1) / CB, (int64_t)min(p.bnc, kBMaxTiles));
  const int64_t tbase = ((int64_t)panel * p.bnr + rb) * p.bnc;
  const int64_t tlast = (int64_t)8 * p.bnr * p.bnc;  // bchunk has tlast + 1 entries
  const uint32_t* __restrict__ ent = p.bent;
  const T* __restrict__ bv = (const T*)p.bvals;

  // The workgroup's tile bounds in LDS (read back as uniform values). The walk reads up to
  // bound(nc + 2) (steps run to c = nc and read up to bound(c + 2)); entries past the workgroup's own
  // tiles hold later tiles' bounds, clamped to the last, so every bound is a valid chunk offset.
  for (int i = tid; i < nc + 8; i += kBThreads) tb[i] = p.bchunk[min(tbase + i, tlast)];
  __syncthreads();
  auto bound = [&](int c) -> int64_t { return __builtin_amdgcn_readfirstlane(tb[min(c, nc + 7)]); };
  auto block_count = [&](int c) -> int {
    const int64_t rest = width - (int64_t)c * CB;
    return (int)(rest < CB ? rest : CB);  // <= 0 past the panel
  };

  V xa[kBXv], xq[kBXv];  // column blocks c and c + 1 in flight, alternating
  auto load_block = [&](int c, V* xr) {
    const int64_t lo = c0 + (int64_t)c * CB;
    const int cnt = block_count(c);
#pragma unroll
    for (int u = 0; u < kBXv; ++u) {
      const int i = (u * kBThreads + tid) * L;
      xr[u] = *reinterpret_cast<const V*>(i + L <= cnt ? p.g + lo + i : p.g);
    }
  };
  auto store_block = [&](int c, const V* xr) {
    const int64_t lo = c0 + (int64_t)c * CB;
    const int cnt = block_count(c);
#pragma unroll
    for (int u = 0; u < kBXv; ++u) {
      const int i = (u * kBThreads + tid) * L;
      if (i + L <= cnt) {
        *reinterpret_cast<V*>(xb + i) = xr[u];
      } else if (i < cnt) {  // the ragged end of the last panel
        for (int q = 0; q < L; ++q) xb[i + q] = (i + q < cnt) ? p.g[lo + i + q] : (T)0;
      }
    }
  };
  auto fetch = [&](int64_t k0, int64_t k1, uint32_t* e, T* v) {
#pragma unroll
    for (int u = 0; u < kBPre; ++u) {
      const int64_t k = k0 + wave + (int64_t)u * kBWaves;
      const int64_t kk = k < k1 ? k : 0;  // chunk 0 always exists; its entries are dropped
      e[u] = stream_load(&ent[kk * 64 + lane]);
      if constexpr (!kSigned) v[u] = stream_load(&bv[kk * 64 + lane]);
      else v[u] = (T)0;
    }
  };
  // Tile c (summed only if c < nc; a step past the panel keeps the barriers uniform): its entries are
  // in cur and its column block in X; the entries of tile c + 1 are loaded into nxt and the column
  // block of tile c + 2 into X.
  auto step = [&](int c, V* X, const uint32_t* cur, const T* curv, uint32_t* nxt, T* nxtv) {
    const bool live = c < nc;
    const int64_t k0 = live ? bound(c) : 0, k1 = live ? bound(c + 1) : 0;
    const int64_t f0 = bound(c + 1), f1 = bound(c + 2);
    __syncthreads();  // tile c - 1 is summed: xb is free
    store_block(c, X);
    __syncthreads();
    fetch(f0, f1, nxt, nxtv);
    T v[kBPre];
    bool on[kBPre];
#pragma unroll
    for (int u = 0; u < kBPre; ++u) {
      v[u] = block_entry<T, kSigned>(cur[u], curv[u], xb);
      on[u] = k0 + wave + (int64_t)u * kBWaves < k1;
    }
    block_chunks_add<T, kBPre>(cur, v, on, yacc, lane);
    for (int64_t k = k0 + wave + (int64_t)kBPre * kBWaves; k < k1; k += kBWaves) {
      uint32_t e[1] = {stream_load(&ent[k * 64 + lane])};
      T w = (T)0;
      if constexpr (!kSigned) w = stream_load(&bv[k * 64 + lane]);
      T x[1] = {block_entry<T, kSigned>(e[0], w, xb)};
      const bool one[1] = {true};
      block_chunks_add<T, 1>(e, x, one, yacc, lane);
    }
    // the column block two tiles ahead, issued after the summing: its loads then do not delay this
    // tile's and have a whole tile to land before their store
    load_block(c + 2, X);
  };

  uint32_t eA[kBPre], eB[kBPre];
  T vA[kBPre], vB[kBPre];
  if (nc > 0) {
    load_block(0, xa);
    fetch(bound(0), bound(1), eA, vA);
    load_block(1, xq);
  }
  for (int c = 0; c < nc; c += 2) {
    step(c, xa, eA, vA, eB, vB);
    step(c + 1, xq, eB, vB, eA, vA);
  }
  __syncthreads();
  const T scale = kSigned ? p.pmag : (T)1;
  T* out = p.partial + (int64_t)panel * p.pld + r0;
  for (int i = tid; i < rows; i += kBThreads) out[i] = scale * yacc[i];
}

template <typename T>
int launch_block_spmv(const SpmvParams<T>& p, hipStream_t stream) {
  if (p.npanels != 8 || !p.pbound || !p.bchunk || !p.bent || (!p.bvals && !p.psigned) || !p.partial) {
    set_error("spmv: incomplete LDS-blocked layout");
    return SA_E_ARG;
  }
  if (p.brows < 64 || p.brows % 64 || p.brows > BlockCfg<T>::kRows || p.bcols < 64 || p.bcols % 64 ||
      p.bcols > BlockCfg<T>::kCols || p.bnr < 1 || p.bnc < 1 || p.bnc > kBMaxTiles ||
      (!p.rbound && (int64_t)p.bnr * p.brows < p.n) || p.n < 64) {
    set_error("spmv: LDS-blocked layout sizes out of range");
    return SA_E_ARG;
  }
  if (!aligned16(p.g)) {
    set_error("spmv: the LDS-blocked layout needs a 16-byte aligned gathered vector");
    return SA_E_ARG;
  }
  const dim3 grid((unsigned)(8 * p.bnr)), block(kBThreads);
  if (p.psigned) hipLaunchKernelGGL((spmv_block_kernel<T, true>), grid, block, 0, stream, p);
  else hipLaunchKernelGGL((spmv_block_kernel<T, false>), grid, block, 0, stream, p);
  return check_launch("spmv_block");
}

template int launch_block_spmv<float>(const SpmvParams<float>&, hipStream_t);
template int launch_block_spmv<double>(const SpmvParams<double>&, hipStream_t);

// ---- layout construction ---------------------------------------------------------------------
// Size class of a run of L entries: 0 for 1, else ceil(log2 L) (slot 2^k); 7 for L > 64.
__device__ inline int run_class(int L) { return L <= 1 ? 0 : L > 64 ? 7 : 32 - __clz(L - 1); }

struct BlockGeom {
  int64_t n;
  int32_t brows, bcols, bnr, bnc;
  const int32_t* rbound;  // NULL or bnr + 1 row-block boundaries
};

// Row block of row r: r / brows, or the block [rbound[rb], rbound[rb + 1]) that holds r.
__device__ inline int row_block(const BlockGeom& G, int64_t r) {
  if (!G.rbound) return (int)(r / G.brows);
  int lo = 0, hi = G.bnr - 1;  // the last rb with rbound[rb] <= r
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (G.rbound[mid] <= r) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Calls f(t, k, e0, L, lr, cb_lo) for each run of row r: entries [e0, e0 + L) in tile t, class k.
template <class F>
__device__ inline void for_each_run(const sa_csr& A, const int32_t* bnd, const BlockGeom& G, int64_t r, F f) {
  const int rb = row_block(G, r);
  const int lr = (int)(r - (G.rbound ? (int64_t)G.rbound[rb] : (int64_t)rb * G.brows));
  const int32_t e_end = A.indptr[r + 1];
  int pnl = 0;
  int32_t e = A.indptr[r];
  while (e < e_end) {
    const int32_t j = A.indices[e];
    while (pnl < 7 && j >= bnd[pnl + 1]) ++pnl;
    const int cb = (j - bnd[pnl]) / G.bcols;
    const int32_t lo = bnd[pnl] + cb * G.bcols;
    const int32_t hi = min(lo + G.bcols, bnd[pnl + 1]);
    int32_t e1 = e + 1;
    while (e1 < e_end && A.indices[e1] < hi) ++e1;
    const int64_t t = ((int64_t)pnl * G.bnr + rb) * G.bnc + cb;
    f(t, run_class(e1 - e), e, e1 - e, lr, lo);
    e = e1;
  }
}

// Per-tile counters in the caller's runs array, kRunStride int32 per tile: [0, 7) runs per size class,
// [8, 15) the fill's slot cursors per class, [16, 48) singleton runs per row bank (local row mod 32),
// [48, 80) the fill's cursors per row bank, [80, 96) 64-entry pieces of the long runs (more than 64
// entries) per chunk column (local row mod 16), [96, 112) the fill's cursors per chunk column.
constexpr int kRunStride = 128;
constexpr int kRunCur = 8, kBankCnt = 16, kBankCur = 48, kColCnt = 80, kColCur = 96;

// Entries of the long region at the head of a tile: 16 chunk columns (chunk r * 16 + c is in column
// c; the kernel gives chunk k of a tile to wave k mod 16, so a column is one wave's, in row order) as
// tall as the tallest column's pieces.
__device__ inline int64_t long_slots(const int32_t* runs, int64_t t) {
  int32_t h = 0;
  for (int c = 0; c < 16; ++c) h = max(h, runs[kRunStride * t + kColCnt + c]);
  return (int64_t)h * 16 * 64;
}

// Where the singletons start: after the long region and the slots of the runs of two to 64 entries
// (classes 1..6, largest first), rounded up to 32 so that they start on a half-wave boundary.
__device__ inline int64_t multi_slots(const int32_t* runs, int64_t t) {
  int64_t s = 0;
  for (int k = 1; k <= 6; ++k) s += (int64_t)runs[kRunStride * t + k] << k;
  return long_slots(runs, t) + ((s + 31) & ~(int64_t)31);
}

__global__ __launch_bounds__(SA_BLOCK) void block_count_kernel(sa_csr A, const int32_t* __restrict__ pbound,
                                                               BlockGeom G, int32_t* __restrict__ runs,
                                                               int32_t* __restrict__ flag) {
  __shared__ int32_t bnd[9];
  if (threadIdx.x < 9) bnd[threadIdx.x] = pbound[threadIdx.x];
  __syncthreads();
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < G.n; r += (int64_t)gridDim.x * blockDim.x) {
    for_each_run(A, bnd, G, r, [&](int64_t t, int k, int32_t, int32_t L, int lr, int32_t) {
      if (lr >= G.brows) atomicOr(flag, 16);  // a row block taller than brows: the plan is refused
      if (k > 6) {
        atomicAdd(&runs[kRunStride * t + kColCnt + (lr & 15)], (L + 63) >> 6);
      } else {
        atomicAdd(&runs[kRunStride * t + k], 1);
        if (k == 0) atomicAdd(&runs[kRunStride * t + kBankCnt + (lr & 31)], 1);
      }
    });
  }
}

__device__ inline int64_t tile_slots(const int32_t* runs, int64_t t) {
  return multi_slots(runs, t) + runs[kRunStride * t];
}

// bchunk[t] = sum over earlier tiles of ceil(slots / 64) (one workgroup, in passes of SA_BLOCK tiles).
__global__ __launch_bounds__(SA_BLOCK) void block_scan_kernel(int64_t ntiles, const int32_t* __restrict__ runs,
                                                              int32_t* __restrict__ bchunk,
                                                              int32_t* __restrict__ flag) {
  __shared__ int64_t wsum[SA_WAVES_PER_BLOCK];
  int64_t carry = 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t base = 0; base < ntiles; base += SA_BLOCK) {
    const int64_t t = base + threadIdx.x;
    const int64_t ch = t < ntiles ? (tile_slots(runs, t) + 63) >> 6 : 0;
    int64_t incl = ch;
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int64_t before = carry;
    for (int w = 0; w < wave; ++w) before += wsum[w];
    int64_t total = carry;
    for (int w = 0; w < SA_WAVES_PER_BLOCK; ++w) total += wsum[w];
    const int64_t excl = before + incl - ch;
    if (t < ntiles) {
      if (excl + ch >= (1ll << 31) / 64) atomicOr(flag, 4);
      bchunk[t] = (int32_t)excl;
    }
    carry = total;
    __syncthreads();
  }
  if (threadIdx.x == 0) bchunk[ntiles] = (int32_t)(carry < (1ll << 31) ? carry : 0);
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void block_pad_kernel(int64_t nent, uint32_t pad, uint32_t* __restrict__ bent,
                                                             T* __restrict__ bvals) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nent; i += (int64_t)gridDim.x * blockDim.x) {
    bent[i] = pad;
    if (bvals) bvals[i] = (T)0;
  }
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void block_fill_kernel(sa_csr A, const int32_t* __restrict__ pbound, BlockGeom G,
                                                              int32_t* __restrict__ runs,
                                                              const int32_t* __restrict__ bchunk, int64_t nent,
                                                              uint32_t* __restrict__ bent, T* __restrict__ bvals,
                                                              T mag, int32_t* __restrict__ flag) {
  __shared__ int32_t bnd[9];
  if (threadIdx.x < 9) bnd[threadIdx.x] = pbound[threadIdx.x];
  __syncthreads();
  const T* vals = (const T*)A.vals;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < G.n; r += (int64_t)gridDim.x * blockDim.x) {
    for_each_run(A, bnd, G, r, [&](int64_t t, int k, int32_t e0, int32_t L, int lr, int32_t lo) {
      const int32_t* rt = runs + kRunStride * t;
      const int64_t tbase = (int64_t)bchunk[t] * 64, tend = (int64_t)bchunk[t + 1] * 64;
      if (k > 6) {
        // a long run: pieces of 64 entries stacked in chunk column lr mod 16 of the long region, each
        // piece one run for the kernel (its last entry ends the run)
        const int c = lr & 15;
        const int32_t P = (L + 63) >> 6;
        const int32_t r0 = atomicAdd(&runs[kRunStride * t + kColCur + c], P);
        for (int32_t q = 0; q < P; ++q) {
          const int64_t pos = tbase + ((int64_t)(r0 + q) * 16 + c) * 64;
          if (pos + 64 > nent || pos + 64 > tend || r0 + q >= rt[kColCnt + c]) {
            atomicOr(flag, 8);  // the plan does not match this matrix
            return;
          }
          for (int i = q * 64; i < L && i < (q + 1) * 64; ++i) {
            const T v = vals[e0 + i];
            const uint32_t col = (uint32_t)(A.indices[e0 + i] - lo);
            const bool cont = i + 1 < L && i + 1 < (q + 1) * 64;
            uint32_t code = (col << 15) | (uint32_t)lr | (cont ? (1u << 30) : 0u);
            if (mag > (T)0) {
              const T a = v < (T)0 ? -v : v;
              if (a != mag) atomicOr(flag, 2);
              if (v < (T)0) code |= 0x80000000u;
            } else {
              bvals[pos + (i - q * 64)] = v;
            }
            bent[pos + (i - q * 64)] = code;
          }
        }
        return;
      }
      int64_t off;
      if (k > 0) {
        // slots of class k start after the long region and the larger classes' slots
        off = long_slots(runs, t);
        for (int q = 6; q > k; --q) off += (int64_t)rt[q] << q;
        off += (int64_t)atomicAdd(&runs[kRunStride * t + kRunCur + k], 1) << k;
      } else {
        // singletons: rank r in row bank b goes to position r-th round over the banks, b-th among
        // the banks that still have an r-th singleton, so each 32 consecutive singletons (one
        // half-wave: the region starts on a multiple of 32) hit distinct row banks until the banks
        // with fewer singletons run out
        const int b = lr & 31;
        const int32_t rank = atomicAdd(&runs[kRunStride * t + kBankCur + b], 1);
        off = multi_slots(runs, t);
        for (int q = 0; q < 32; ++q) {
          const int32_t c = rt[kBankCnt + q];
          off += min(c, rank) + ((q < b && c > rank) ? 1 : 0);
        }
      }
      const int64_t pos = tbase + off;
      if (pos + L > nent || pos + L > tend) {
        atomicOr(flag, 8);  // the plan does not match this matrix
        return;
      }
      for (int i = 0; i < L; ++i) {
        const T v = vals[e0 + i];
        const uint32_t col = (uint32_t)(A.indices[e0 + i] - lo);
        uint32_t code = (col << 15) | (uint32_t)lr | (i + 1 < L ? (1u << 30) : 0u);
        if (mag > (T)0) {
          const T a = v < (T)0 ? -v : v;
          if (a != mag) atomicOr(flag, 2);
          if (v < (T)0) code |= 0x80000000u;
        } else {
          bvals[pos + i] = v;
        }
        bent[pos + i] = code;
      }
    });
  }
}

static int block_geom(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bcols, int32_t bnr, int32_t bnc,
                      BlockGeom& G) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && A->indices && A->vals && pbound, "bad arguments");
  SA_REQUIRE(brows >= 64 && brows % 64 == 0 && brows < 32768, "brows must be a multiple of 64 below 32768");
  SA_REQUIRE(bcols >= 64 && bcols % 64 == 0 && bcols <= 32768, "bcols must be a multiple of 64 up to 32768");
  SA_REQUIRE(bnr >= 1 && bnc >= 1 && (A->rbound || (int64_t)bnr * brows >= A->n),
             "bnr row blocks must cover the rows");
  SA_REQUIRE((int64_t)8 * bnr * bnc < (1ll << 31), "too many tiles");
  G.rbound = A->rbound;
  G.n = A->n;
  G.brows = brows;
  G.bcols = bcols;
  G.bnr = bnr;
  G.bnc = bnc;
  return SA_OK;
}

template <typename T>
static int block_fill(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bcols, int32_t bnr, int32_t bnc,
                      int32_t* runs, const int32_t* bchunk, int64_t nent, uint32_t* bent, T* bvals, double mag,
                      int32_t* flag, hipStream_t stream) {
  BlockGeom G;
  int rc = block_geom(A, pbound, brows, bcols, bnr, bnc, G);
  if (rc) return rc;
  SA_REQUIRE(runs && bchunk && bent && flag && (mag > 0.0 || bvals) && nent >= 0 && nent % 64 == 0, "bad arguments");
  hipLaunchKernelGGL(block_pad_kernel<T>, dim3(grid_for(nent, SA_BLOCK * 4, 8192)), dim3(SA_BLOCK), 0, stream, nent,
                     (uint32_t)brows, bent, mag > 0.0 ? (T*)nullptr : bvals);
  rc = check_launch("block_pad");
  if (rc) return rc;
  hipLaunchKernelGGL(block_fill_kernel<T>, dim3(grid_for(A->n, SA_BLOCK, 8192)), dim3(SA_BLOCK), 0, stream, *A, pbound,
                     G, runs, bchunk, nent, bent, bvals, (T)(mag > 0.0 ? mag : 0.0), flag);
  return check_launch("block_fill");
}

}  // namespace sa

extern "C" {

int sa_block_plan(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bcols, int32_t bnr, int32_t bnc,
                  int32_t* runs, int32_t* bchunk, int32_t* flag, hipStream_t stream) {
  sa::BlockGeom G;
  int rc = sa::block_geom(A, pbound, brows, bcols, bnr, bnc, G);
  if (rc) return rc;
  SA_REQUIRE(runs && bchunk && flag, "bad arguments");
  hipLaunchKernelGGL(sa::block_count_kernel, dim3(sa::grid_for(A->n, SA_BLOCK, 8192)), dim3(SA_BLOCK), 0, stream, *A,
                     pbound, G, runs, flag);
  rc = sa::check_launch("block_count");
  if (rc) return rc;
  hipLaunchKernelGGL(sa::block_scan_kernel, dim3(1), dim3(SA_BLOCK), 0, stream, (int64_t)8 * bnr * bnc, runs, bchunk,
                     flag);
  return sa::check_launch("block_scan");
}

int sa_block_fill_f32(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bcols, int32_t bnr, int32_t bnc,
                      int32_t* runs, const int32_t* bchunk, int64_t nent, uint32_t* bent, float* bvals, double mag,
                      int32_t* flag, hipStream_t stream) {
  return sa::block_fill<float>(A, pbound, brows, bcols, bnr, bnc, runs, bchunk, nent, bent, bvals, mag, flag, stream);
}

int sa_block_fill_f64(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bcols, int32_t bnr, int32_t bnc,
                      int32_t* runs, const int32_t* bchunk, int64_t nent, uint32_t* bent, double* bvals, double mag,
                      int32_t* flag, hipStream_t stream) {
  return sa::block_fill<double>(A, pbound, brows, bcols, bnr, bnc, runs, bchunk, nent, bent, bvals, mag, flag, stream);
}

}  // extern "C"
