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
constexpr int kBLdsBytes = 163840;
constexpr int kBPre = 4;                       // chunks of the next tile each wave loads ahead
constexpr int kBMaxXv = 8;                     // 16-byte loads per thread for one column block

template <typename T>
__device__ inline int block_xoff(int brows) {
  // row sums [0, brows] (brows is the padding slot), then the column block, 16-byte aligned
  const int bytes = (brows + 1) * (int)sizeof(T);
  return ((bytes + 15) & ~15) / (int)sizeof(T);
}

template <typename T, bool kSigned>
__device__ inline T block_entry(uint32_t e, T val, const T* xb) {
  const T x = xb[(e >> 15) & 0x7FFF];
  if constexpr (kSigned) {
    return (e & 0x80000000u) ? -x : x;
  } else {
    return val * x;
  }
}

// One chunk: lane values v (already gathered), runs reduced onto their first lane, which adds the
// run sum into its row.
template <typename T>
__device__ inline void block_chunk_add(uint32_t e, T v, T* yacc, int lane) {
  const uint64_t m = __ballot((e >> 30) & 1u);  // lane l continues into lane l + 1
  uint64_t c = m;
  for (int d = 1; c != 0; d <<= 1) {
    const T o = __shfl_down(v, d, 64);
    if ((c >> lane) & 1ull) v += o;
    c &= c >> d;
  }
  const bool head = lane == 0 || !((m >> (lane - 1)) & 1ull);
  if (head) {
    const int r = (int)(e & 0x7FFFu);
    yacc[r] += v;
  }
}

template <typename T, bool kSigned>
__global__ __launch_bounds__(kBThreads) void spmv_block_kernel(SpmvParams<T> p) {
  __shared__ __attribute__((aligned(16))) char lds[kBLdsBytes];
  if (p.gate && *p.gate != SA_RUNNING) return;
  typedef typename vec16<T>::type V;
  constexpr int L = vec16<T>::L;
  const int panel = (int)(blockIdx.x & 7);
  const int rb = (int)(blockIdx.x >> 3);
  const int RB = p.brows, CB = p.bcols;
  const int64_t r0 = (int64_t)rb * RB;
  if (r0 >= p.n) return;
  const int rows = (int)((p.n - r0) < RB ? (p.n - r0) : RB);
  T* yacc = reinterpret_cast<T*>(lds);
  T* xb = yacc + block_xoff<T>(RB);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  for (int i = tid; i <= RB; i += kBThreads) yacc[i] = (T)0;

  const int64_t c0 = p.pbound[panel];
  const int64_t width = (int64_t)p.pbound[panel + 1] - c0;
  const int nc = (int)min((width + CB - 1) / CB, (int64_t)p.bnc);
  const int64_t tbase = ((int64_t)panel * p.bnr + rb) * p.bnc;
  const uint32_t* __restrict__ ent = p.bent;
  const T* __restrict__ bv = (const T*)p.bvals;

  // column block c of g into registers (elements beyond the block are not loaded)
  V xr[kBMaxXv];
  auto load_block = [&](int c) {
    const int64_t lo = c0 + (int64_t)c * CB;
    const int cnt = (int)((width - (int64_t)c * CB) < CB ? (width - (int64_t)c * CB) : CB);
#pragma unroll
    for (int u = 0; u < kBMaxXv; ++u) {
      const int i = (u * kBThreads + tid) * L;
      if (i + L <= cnt) {
        xr[u] = *reinterpret_cast<const V*>(p.g + lo + i);
      } else if (i < cnt) {
        V t;
#pragma unroll
        for (int q = 0; q < L; ++q) t[q] = (i + q < cnt) ? p.g[lo + i + q] : (T)0;
        xr[u] = t;
      }
    }
  };
  auto store_block = [&](int c) {
    const int cnt = (int)((width - (int64_t)c * CB) < CB ? (width - (int64_t)c * CB) : CB);
#pragma unroll
    for (int u = 0; u < kBMaxXv; ++u) {
      const int i = (u * kBThreads + tid) * L;
      if (i < cnt) *reinterpret_cast<V*>(xb + i) = xr[u];
    }
  };

  // entries of the first kBPre chunks this wave takes in a tile
  uint32_t ce[kBPre], ne[kBPre];
  T cv[kBPre], nv[kBPre];
  auto fetch = [&](int64_t t, uint32_t* e, T* v) {
    const int64_t k0 = p.bchunk[t], k1 = p.bchunk[t + 1];
#pragma unroll
    for (int u = 0; u < kBPre; ++u) {
      const int64_t k = k0 + wave + (int64_t)u * kBWaves;
      e[u] = 0u;
      v[u] = (T)0;
      if (k < k1) {
        e[u] = stream_load(&ent[k * 64 + lane]);
        if constexpr (!kSigned) v[u] = stream_load(&bv[k * 64 + lane]);
      }
    }
  };

  if (nc > 0) {
    load_block(0);
    fetch(tbase, ce, cv);
  }
  for (int c = 0; c < nc; ++c) {
    const int64_t t = tbase + c;
    const int64_t k0 = p.bchunk[t], k1 = p.bchunk[t + 1];
    __syncthreads();  // the previous tile's gathers are done with xb (and yacc is zeroed)
    store_block(c);
    __syncthreads();
    if (c + 1 < nc) {
      load_block(c + 1);
      fetch(t + 1, ne, nv);
    }
#pragma unroll
    for (int u = 0; u < kBPre; ++u) {
      const int64_t k = k0 + wave + (int64_t)u * kBWaves;
      if (k < k1) block_chunk_add<T>(ce[u], block_entry<T, kSigned>(ce[u], cv[u], xb), yacc, lane);
    }
    for (int64_t k = k0 + wave + (int64_t)kBPre * kBWaves; k < k1; k += kBWaves) {
      const uint32_t e = stream_load(&ent[k * 64 + lane]);
      T v = (T)0;
      if constexpr (!kSigned) v = stream_load(&bv[k * 64 + lane]);
      block_chunk_add<T>(e, block_entry<T, kSigned>(e, v, xb), yacc, lane);
    }
#pragma unroll
    for (int u = 0; u < kBPre; ++u) {
      ce[u] = ne[u];
      cv[u] = nv[u];
    }
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
  const int xoff = (((p.brows + 1) * (int)sizeof(T) + 15) & ~15);
  if (p.brows < 64 || p.brows % 64 || p.brows >= 32768 || p.bcols < 64 || p.bcols % 64 || p.bcols > 32768 ||
      p.bcols > kBMaxXv * kBThreads * vec16<T>::L || xoff + p.bcols * (int)sizeof(T) > kBLdsBytes || p.bnr < 1 ||
      p.bnc < 1 || (int64_t)p.bnr * p.brows < p.n) {
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
};

// Calls f(t, k, e0, L, lr, cb_lo) for each run of row r: entries [e0, e0 + L) in tile t, class k.
template <class F>
__device__ inline void for_each_run(const sa_csr& A, const int32_t* bnd, const BlockGeom& G, int64_t r, F f) {
  const int rb = (int)(r / G.brows);
  const int lr = (int)(r - (int64_t)rb * G.brows);
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

__global__ __launch_bounds__(SA_BLOCK) void block_count_kernel(sa_csr A, const int32_t* __restrict__ pbound,
                                                               BlockGeom G, int32_t* __restrict__ runs,
                                                               int32_t* __restrict__ flag) {
  __shared__ int32_t bnd[9];
  if (threadIdx.x < 9) bnd[threadIdx.x] = pbound[threadIdx.x];
  __syncthreads();
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < G.n; r += (int64_t)gridDim.x * blockDim.x) {
    for_each_run(A, bnd, G, r, [&](int64_t t, int k, int32_t, int32_t, int, int32_t) {
      if (k > 6) atomicOr(flag, 1);
      else atomicAdd(&runs[16 * t + k], 1);
    });
  }
}

__device__ inline int64_t tile_slots(const int32_t* runs, int64_t t) {
  int64_t s = 0;
  for (int k = 0; k <= 6; ++k) s += (int64_t)runs[16 * t + k] << k;
  return s;
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
      if (k > 6) return;
      // slots of class k start after the larger classes' slots
      int64_t off = 0;
      for (int q = 6; q > k; --q) off += (int64_t)runs[16 * t + q] << q;
      const int32_t slot = atomicAdd(&runs[16 * t + 8 + k], 1);
      const int64_t pos = (int64_t)bchunk[t] * 64 + off + ((int64_t)slot << k);
      if (pos + L > nent || pos + L > (int64_t)bchunk[t + 1] * 64) {
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
  SA_REQUIRE(bnr >= 1 && bnc >= 1 && (int64_t)bnr * brows >= A->n, "bnr row blocks must cover the rows");
  SA_REQUIRE((int64_t)8 * bnr * bnc < (1ll << 31), "too many tiles");
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
