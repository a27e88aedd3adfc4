// Fused scaled CSR SpMV for gfx950.
//
// The matrix is cut into tiles of tile_nnz stored entries; a workgroup owns the rows whose first
// entry lies in its tile. Phase 1 streams the tile's column indices and values (non-temporal,
// coalesced), eight entries per thread with all eight gathers of g in flight at once, and keeps the
// exact FP64 products in LDS. Phase 2 sums each row's products with P lanes per row (P from the
// tile's row count) and a shuffle reduction. A row that runs past the end of its tile (only the last
// row of a tile can) is summed by the whole workgroup. Sums run in a fixed order, so results are
// bitwise reproducible; row sums are FP64 even in the FP32 form.
//
// Random gathers are the bound of this kernel on graphs without locality (a 4-byte gather costs a
// cache-line access). Workgroups reach the 8 XCDs round-robin, so in the XCD-panel layout each XCD
// only gathers from its own column panel, which then stays resident in that XCD's 4 MB L2; the
// per-panel row sums are combined by the consumer in panel order.
#include "sa_spmv.hpp"

#include <mutex>

#include "sa_reduce.hpp"

namespace sa {

constexpr int kUnroll = 8;

template <typename T>
struct PlainStore {
  const T* s;
  const T* x;
  double a, bb;
  T* y;
  __device__ void operator()(int32_t r, double acc) const {
    const double sr = s ? (double)s[r] : 1.0;
    double v = bb * sr * acc;
    if (x) v += a * (double)x[r];
    y[r] = (T)v;
  }
};

template <typename T>
struct PanelStore {
  T* part;  // partial + p * n
  __device__ void operator()(int32_t r, double acc) const { part[r] = (T)acc; }
};

// Entry decoding: plain (index, value) pairs, or sign-packed indices (bit 31 = sign, magnitude mag)
// when every coupling has one magnitude, which halves the stream of a panel.
template <typename T, bool kSigned>
__device__ inline void load_entry(const int32_t* __restrict__ indices, const T* __restrict__ vals, T mag, int32_t e,
                                  int32_t& c, T& v) {
  const int32_t raw = stream_load(&indices[e]);
  if constexpr (kSigned) {
    c = raw & 0x7FFFFFFF;
    v = raw < 0 ? -mag : mag;
  } else {
    c = raw;
    v = stream_load(&vals[e]);
  }
}

// Products of entries [e0, e0 + cnt) into prod[0, cnt), eight gathers in flight per thread.
template <typename T, bool kSigned>
__device__ inline void stage_products(const int32_t* __restrict__ indices, const T* __restrict__ vals, T mag,
                                      const T* __restrict__ g, int32_t e0, int32_t cnt, double* prod) {
  for (int32_t base = 0; base < cnt; base += SA_BLOCK * kUnroll) {
    int32_t c[kUnroll];
    T v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int32_t e = base + u * SA_BLOCK + (int32_t)threadIdx.x;
      c[u] = 0;
      v[u] = (T)0;
      if (e < cnt) load_entry<T, kSigned>(indices, vals, mag, e0 + e, c[u], v[u]);
    }
    T gv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int32_t e = base + u * SA_BLOCK + (int32_t)threadIdx.x;
      gv[u] = e < cnt ? g[c[u]] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int32_t e = base + u * SA_BLOCK + (int32_t)threadIdx.x;
      if (e < cnt) prod[e] = (double)v[u] * (double)gv[u];
    }
  }
}

// One tile: rows [r0, r1) of the CSR (indptr, indices, vals); store(r, row sum) for each row.
template <typename T, int TILE, bool kSigned, class Store>
__device__ inline void tile_body(const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                 const T* __restrict__ vals, T mag, const T* __restrict__ g, int32_t r0, int32_t r1,
                                 int lanes_fixed, double* prod, double* red, const Store& store) {
  const int32_t e0 = indptr[r0];
  const int32_t e_last = indptr[r1 - 1];
  const int32_t e1 = indptr[r1];
  const bool long_last = (e1 - e0) > TILE;
  const int32_t rs = long_last ? r1 - 1 : r1;
  const int32_t es = long_last ? e_last : e1;

  stage_products<T, kSigned>(indices, vals, mag, g, e0, es - e0, prod);
  __syncthreads();

  const int32_t nrows = rs - r0;
  if (nrows > 0) {
    int P = lanes_fixed;
    if (P <= 0) {
      const int q = SA_BLOCK / nrows;
      P = q >= 64 ? 64 : q >= 32 ? 32 : q >= 16 ? 16 : q >= 8 ? 8 : q >= 4 ? 4 : q >= 2 ? 2 : 1;
    }
    const int groups = SA_BLOCK / P;
    const int gid = threadIdx.x / P, lig = threadIdx.x % P;
    for (int32_t rb = r0; rb < rs; rb += groups) {
      const int32_t r = rb + gid;
      double acc = 0.0;
      if (r < rs) {
        const int32_t b0 = indptr[r] - e0, b1 = indptr[r + 1] - e0;
        for (int32_t e = b0 + lig; e < b1; e += P) acc += prod[e];
      }
      for (int o = P >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if (lig == 0 && r < rs) store(r, acc);
    }
  }

  if (long_last) {
    const int32_t r = r1 - 1;
    double acc = 0.0;
    for (int32_t base = e_last; base < e1; base += SA_BLOCK * kUnroll) {
      int32_t c[kUnroll];
      T v[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int32_t e = base + u * SA_BLOCK + (int32_t)threadIdx.x;
        c[u] = 0;
        v[u] = (T)0;
        if (e < e1) load_entry<T, kSigned>(indices, vals, mag, e, c[u], v[u]);
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int32_t e = base + u * SA_BLOCK + (int32_t)threadIdx.x;
        if (e < e1) acc += (double)v[u] * (double)g[c[u]];
      }
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) store(r, acc);
  }
}

template <typename T, int TILE>
__global__ __launch_bounds__(SA_BLOCK) void spmv_tile_kernel(SpmvParams<T> p) {
  __shared__ double prod[TILE];
  __shared__ double red[SA_WAVES_PER_BLOCK];
  double bb = p.b;
  if (p.gate && *p.gate != SA_RUNNING) return;
  if (p.lz) bb *= p.lz->sigma[p.lz->j];
  const int64_t t = p.tile_list ? (int64_t)p.tile_list[blockIdx.x] : (int64_t)blockIdx.x;
  if (p.tile_inst &&
      *reinterpret_cast<const int32_t*>(p.gate_base + (int64_t)p.tile_inst[t] * p.gate_stride) != SA_RUNNING)
    return;
  const int32_t r0 = p.tile_rows[t], r1 = p.tile_rows[t + 1];
  if (r0 >= r1) return;
  PlainStore<T> st{p.s, p.x, p.a, bb, p.y};
  tile_body<T, TILE, false>(p.indptr, p.indices, p.vals, (T)0, p.g, r0, r1, p.lanes_fixed, prod, red, st);
}

// Workgroup w -> XCD w % 8; panel (w / (8 maxT)) * 8 + w % 8, tile (w / 8) % maxT.
template <typename T, int TILE, bool kSigned>
__global__ __launch_bounds__(SA_BLOCK) void spmv_panel_kernel(SpmvParams<T> p) {
  __shared__ double prod[TILE];
  __shared__ double red[SA_WAVES_PER_BLOCK];
  if (p.gate && *p.gate != SA_RUNNING) return;
  const int64_t w = blockIdx.x;
  const int64_t q = w >> 3;
  const int panel = (int)((q / p.pmax_tiles) * 8 + (w & 7));
  const int64_t t = q % p.pmax_tiles;
  if (panel >= p.npanels) return;
  const int64_t toff = p.ptile_off[panel];
  const int64_t nt = p.ptile_off[panel + 1] - toff - 1;
  if (t >= nt) return;
  const int32_t r0 = p.ptile_rows[toff + t], r1 = p.ptile_rows[toff + t + 1];
  if (r0 >= r1) return;
  PanelStore<T> st{p.partial + (int64_t)panel * p.pld};
  tile_body<T, TILE, kSigned>(p.pindptr + (int64_t)panel * (p.n + 1), p.pindices, p.pvals, p.pmag, p.g, r0, r1,
                              p.lanes_fixed, prod, red, st);
}

// The same with 16-byte loads and stores (4 floats or 2 doubles per lane), every row summed in the same
// order: bit-identical to combine_kernel. Needs 16-byte aligned vectors and pld a multiple of the vector.
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void combine_vec_kernel(int64_t n, int32_t npanels,
                                                               const T* __restrict__ partial, int64_t pld,
                                                               const T* __restrict__ s, const T* __restrict__ x,
                                                               double a, double b, T* __restrict__ y,
                                                               const int32_t* gate) {
  typedef typename vec16<T>::type V;
  constexpr int L = vec16<T>::L;
  if (gate && *gate != SA_RUNNING) return;
  const int64_t nv = n / L;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nv; q += stride) {
    const int64_t r = q * L;
    double acc[L];
#pragma unroll
    for (int k = 0; k < L; ++k) acc[k] = 0.0;
    int p = 0;
    for (; p + 4 <= npanels; p += 4) {
      V v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(reinterpret_cast<const V*>(&partial[(int64_t)(p + u) * pld + r]));
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int k = 0; k < L; ++k) acc[k] += (double)v[u][k];
    }
    for (; p < npanels; ++p) {
      const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(&partial[(int64_t)p * pld + r]));
#pragma unroll
      for (int k = 0; k < L; ++k) acc[k] += (double)v[k];
    }
    V sv, xv, out;
    if (s) sv = *reinterpret_cast<const V*>(&s[r]);
    if (x) xv = *reinterpret_cast<const V*>(&x[r]);
#pragma unroll
    for (int k = 0; k < L; ++k) {
      double t = b * (s ? (double)sv[k] : 1.0) * acc[k];
      if (x) t += a * (double)xv[k];
      out[k] = (T)t;
    }
    *reinterpret_cast<V*>(&y[r]) = out;
  }
  for (int64_t r = nv * L + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += stride) {
    double acc = 0.0;
    for (int p = 0; p < npanels; ++p) acc += (double)partial[(int64_t)p * pld + r];
    double t = b * (s ? (double)s[r] : 1.0) * acc;
    if (x) t += a * (double)x[r];
    y[r] = (T)t;
  }
}

// y_r = a x_r + b s_r sum_p partial[p][r] (panel order fixed).
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void combine_kernel(int64_t n, int32_t npanels, const T* __restrict__ partial,
                                                           int64_t pld, const T* __restrict__ s,
                                                           const T* __restrict__ x, double a, double b,
                                                           T* __restrict__ y, const int32_t* gate) {
  if (gate && *gate != SA_RUNNING) return;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int p = 0; p < npanels; ++p) acc += (double)partial[(int64_t)p * pld + r];
    double v = b * (s ? (double)s[r] : 1.0) * acc;
    if (x) v += a * (double)x[r];
    y[r] = (T)v;
  }
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void prescale_kernel(int64_t n, const T* __restrict__ s,
                                                            const T* __restrict__ x, T* __restrict__ z) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    z[r] = s ? s[r] * x[r] : x[r];
}

// ---- panel layout construction --------------------------------------------------------------
__global__ __launch_bounds__(SA_BLOCK) void panel_count_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                               const int32_t* __restrict__ indices, int32_t npanels,
                                                               const int32_t* __restrict__ pbound,
                                                               int32_t* __restrict__ counts) {
  __shared__ int32_t bnd[65];
  for (int i = threadIdx.x; i <= npanels; i += blockDim.x) bnd[i] = pbound[i];
  __syncthreads();
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int p = 0, run = 0;
    for (int32_t e = indptr[r]; e < indptr[r + 1]; ++e) {
      const int32_t c = indices[e];
      while (p < npanels - 1 && c >= bnd[p + 1]) {
        counts[(int64_t)p * n + r] = run;
        run = 0;
        ++p;
      }
      ++run;
    }
    for (; p < npanels; ++p) {
      counts[(int64_t)p * n + r] = run;
      run = 0;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void panel_scatter_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                                 const int32_t* __restrict__ indices,
                                                                 const T* __restrict__ vals, int32_t npanels,
                                                                 const int32_t* __restrict__ pbound,
                                                                 const int32_t* __restrict__ pindptr,
                                                                 int32_t* __restrict__ pindices,
                                                                 T* __restrict__ pvals) {
  __shared__ int32_t bnd[65];
  for (int i = threadIdx.x; i <= npanels; i += blockDim.x) bnd[i] = pbound[i];
  __syncthreads();
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int p = 0;
    int32_t dst = pindptr[r];
    for (int32_t e = indptr[r]; e < indptr[r + 1]; ++e) {
      const int32_t c = indices[e];
      while (p < npanels - 1 && c >= bnd[p + 1]) {
        ++p;
        dst = pindptr[(int64_t)p * (n + 1) + r];
      }
      pindices[dst] = c;
      pvals[dst] = vals[e];
      ++dst;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void panel_sign_fold_kernel(int64_t nnz, int32_t* __restrict__ pindices,
                                                                   const T* __restrict__ pvals, T mag,
                                                                   int32_t* __restrict__ mismatch) {
  bool bad = false;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const T v = pvals[e];
    const bool neg = v < (T)0;
    bad |= (neg ? -v : v) != mag;
    pindices[e] = pindices[e] | (neg ? (int32_t)0x80000000 : 0);
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(mismatch, 1);
}

__global__ __launch_bounds__(SA_BLOCK) void panel_tiles_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                               int32_t tile_nnz, int32_t* __restrict__ tile_rows,
                                                               int64_t ntiles) {
  const int64_t base = indptr[0];
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t <= ntiles; t += (int64_t)gridDim.x * blockDim.x) {
    if (t == ntiles) {
      tile_rows[t] = (int32_t)n;
      continue;
    }
    const int64_t target = base + t * (int64_t)tile_nnz;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)indptr[mid] >= target) hi = mid; else lo = mid + 1;
    }
    tile_rows[t] = (int32_t)lo;
  }
}

// ---- launchers ------------------------------------------------------------------------------
template <typename T>
SpmvParams<T> params_from(const sa_csr* A) {
  SpmvParams<T> p;
  p.n = A->n;
  p.indptr = A->indptr;
  p.indices = A->indices;
  p.vals = (const T*)A->vals;
  p.tile_rows = A->tile_rows;
  p.ntiles = A->ntiles;
  p.lanes_fixed = A->lanes_per_row;
  p.g = nullptr;
  p.s = nullptr;
  p.x = nullptr;
  p.a = 0.0;
  p.b = 1.0;
  p.lz = nullptr;
  p.gate = nullptr;
  p.tile_inst = nullptr;
  p.gate_base = nullptr;
  p.gate_stride = 0;
  p.tile_list = nullptr;
  p.ntile_list = 0;
  p.y = nullptr;
  p.npanels = A->npanels;
  p.pindptr = A->pindptr;
  p.pindices = A->pindices;
  p.pvals = (const T*)A->pvals;
  p.ptile_rows = A->ptile_rows;
  p.ptile_off = A->ptile_off;
  p.pmax_tiles = A->pmax_tiles;
  p.partial = (T*)A->partial;
  p.pld = A->pld > 0 ? A->pld : A->n;
  p.psigned = A->psigned;
  p.pmag = (T)A->pmag;
  p.bmode = A->bmode;
  p.brows = A->brows;
  p.bcols = A->bcols;
  p.bnc = A->bnc;
  p.bnr = A->bnr;
  p.pbound = A->pbound;
  p.bchunk = A->bchunk;
  p.bent = A->bent;
  p.bvals = A->bvals;
  p.rbound = A->rbound;
  p.imode = A->imode;
  p.irows = A->irows;
  p.ilanes = A->ilanes;
  p.ipasses = A->ipasses;
  p.isigned = A->isigned;
  p.inwg = A->inwg;
  p.iinst = A->iinst;
  p.irow = A->irow;
  p.ibase = A->ibase;
  p.ireal = A->ireal;
  p.iend = A->iend;
  p.isplit = A->isplit;
  p.ient = A->ient;
  p.imag = (const T*)A->imag;
  p.bcbase = A->bcbase;
  p.pbent = A->pbent;
  p.pbrow = A->pbrow;
  p.pbtoff = A->pbtoff;
  p.pbrtoff = A->pbrtoff;
  p.pbcontrib = (float*)A->pbcontrib;
  p.pbnc = A->pbnc;
  return p;
}

template <typename T>
int launch_spmv(const SpmvParams<T>& p, int32_t tile_nnz, hipStream_t stream) {
  const dim3 block(SA_BLOCK);
  if (p.imode == 1 && p.npanels == 0) return launch_inst_spmv<T>(p, stream);
  if (p.imode == 2 && p.npanels == 0) {
    if constexpr (std::is_same<T, float>::value) {
      return launch_istream_spmv(p, stream);
    } else {
      set_error("spmv: the instance-streamed layout is FP32 only");
      return SA_E_ARG;
    }
  }
  if (p.imode != 0 && p.imode != 1 && p.imode != 2) {
    set_error("spmv: unknown instance layout mode");
    return SA_E_ARG;
  }
  if (p.bmode == 1) return launch_block_spmv<T>(p, stream);
  if (p.bmode == 2 || p.bmode == 3 || p.bmode == 4) {
    if constexpr (std::is_same<T, float>::value) {
      if (p.bmode == 4) return launch_rstream_spmv(p, stream);
      return p.bmode == 2 ? launch_cstream_spmv(p, stream) : launch_pblock_spmv(p, stream);
    } else {
      set_error("spmv: the column-streamed and two-pass layouts are FP32 only");
      return SA_E_ARG;
    }
  }
  if (p.bmode != 0) {
    set_error("spmv: unknown layout mode");
    return SA_E_ARG;
  }
  if (p.npanels > 0) {
    if (p.npanels % 8 != 0 || p.npanels > 64 || p.pmax_tiles < 1 || !p.pindptr || !p.pindices ||
        (!p.pvals && !p.psigned) || !p.ptile_rows || !p.ptile_off || !p.partial) {
      set_error("spmv: incomplete XCD panel layout");
      return SA_E_ARG;
    }
    const int64_t nblocks = (int64_t)p.npanels * p.pmax_tiles;
    if (nblocks >= (1ll << 31)) {
      set_error("spmv: panel grid too large");
      return SA_E_ARG;
    }
    const dim3 grid((unsigned)nblocks);
#define SA_PANEL_CASE(TN)                                                                          \
  case TN:                                                                                         \
    if (p.psigned) hipLaunchKernelGGL((spmv_panel_kernel<T, TN, true>), grid, block, 0, stream, p);  \
    else hipLaunchKernelGGL((spmv_panel_kernel<T, TN, false>), grid, block, 0, stream, p);           \
    break;
    switch (tile_nnz) {
      SA_PANEL_CASE(512)
      SA_PANEL_CASE(1024)
      SA_PANEL_CASE(2048)
      SA_PANEL_CASE(4096)
      default: set_error("spmv: tile_nnz must be 512, 1024, 2048 or 4096"); return SA_E_ARG;
    }
#undef SA_PANEL_CASE
    return check_launch("spmv_panel");
  }
  if (p.ntiles < 1 || p.ntiles >= (1ll << 31) || (p.tile_list && (p.ntile_list < 0 || p.ntile_list > p.ntiles))) {
    set_error("spmv: ntiles out of range");
    return SA_E_ARG;
  }
  if (p.tile_list && p.ntile_list == 0) return SA_OK;  // no tile to run
  const dim3 grid((unsigned)(p.tile_list ? p.ntile_list : p.ntiles));
  switch (tile_nnz) {
    case 512: hipLaunchKernelGGL((spmv_tile_kernel<T, 512>), grid, block, 0, stream, p); break;
    case 1024: hipLaunchKernelGGL((spmv_tile_kernel<T, 1024>), grid, block, 0, stream, p); break;
    case 2048: hipLaunchKernelGGL((spmv_tile_kernel<T, 2048>), grid, block, 0, stream, p); break;
    case 4096: hipLaunchKernelGGL((spmv_tile_kernel<T, 4096>), grid, block, 0, stream, p); break;
    default: set_error("spmv: tile_nnz must be 512, 1024, 2048 or 4096"); return SA_E_ARG;
  }
  return check_launch("spmv");
}

template <typename T>
int launch_prescale(int64_t n, const T* s, const T* x, T* z, hipStream_t stream) {
  hipLaunchKernelGGL(prescale_kernel<T>, dim3(grid_for(n, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK), 0, stream, n, s, x,
                     z);
  return check_launch("prescale");
}

template SpmvParams<float> params_from<float>(const sa_csr*);
template SpmvParams<double> params_from<double>(const sa_csr*);
template int launch_spmv<float>(const SpmvParams<float>&, int32_t, hipStream_t);
template int launch_spmv<double>(const SpmvParams<double>&, int32_t, hipStream_t);
template int launch_prescale<float>(int64_t, const float*, const float*, float*, hipStream_t);
template int launch_prescale<double>(int64_t, const double*, const double*, double*, hipStream_t);

// y = a x + b s (.) (J g): the tile kernel, or the panel kernel plus the combine; skipped unless
// gate is NULL or *gate == SA_RUNNING.
template <typename T>
int spmv_apply(const sa_csr* A, const T* s, const T* g, double a, const T* x, double b, T* y,
               const int32_t* gate, hipStream_t stream) {
  SpmvParams<T> p = params_from<T>(A);
  p.g = g;
  p.s = s;
  p.x = (a != 0.0) ? x : nullptr;
  p.a = a;
  p.b = b;
  p.y = y;
  p.gate = gate;
  int rc = launch_spmv<T>(p, A->tile_nnz, stream);
  if (rc || A->npanels == 0) return rc;
  constexpr int L = vec16<T>::L;
  if (aligned16(A->partial) && p.pld % L == 0 && aligned16(y) && (!s || aligned16(s)) && (!p.x || aligned16(p.x))) {
    hipLaunchKernelGGL(combine_vec_kernel<T>, dim3(grid_for(A->n, SA_BLOCK * L * 2, 4096)), dim3(SA_BLOCK), 0, stream,
                       A->n, A->npanels, (const T*)A->partial, p.pld, s, p.x, a, b, y, gate);
  } else {
    hipLaunchKernelGGL(combine_kernel<T>, dim3(grid_for(A->n, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK), 0, stream, A->n,
                       A->npanels, (const T*)A->partial, p.pld, s, p.x, a, b, y, gate);
  }
  return check_launch("spmv_combine");
}

template int spmv_apply<float>(const sa_csr*, const float*, const float*, double, const float*, double, float*,
                               const int32_t*, hipStream_t);
template int spmv_apply<double>(const sa_csr*, const double*, const double*, double, const double*, double, double*,
                                const int32_t*, hipStream_t);

template <typename T>
static int spmv_scaled(const sa_csr* A, const T* s, const T* x, double a, double b, T* z, T* y, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && A->indices && A->vals && A->tile_rows && x && z && y,
             "bad arguments");
  SA_REQUIRE(A->nnz >= 0 && A->nnz < (1ll << 31), "nnz must fit int32 indices");
  int rc = launch_prescale<T>(A->n, s, x, z, stream);
  if (rc) return rc;
  return spmv_apply<T>(A, s, z, a, x, b, y, nullptr, stream);
}

template <typename T>
static int spmv_gather(const sa_csr* A, const T* s, const T* g, double a, const T* x, double b, T* y,
                       hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && A->indices && A->vals && A->tile_rows && g && y, "bad arguments");
  SA_REQUIRE(a == 0.0 || x, "x is required when a != 0");
  return spmv_apply<T>(A, s, g, a, x, b, y, nullptr, stream);
}

template <typename T>
static int panel_scatter(const sa_csr* A, int32_t npanels, const int32_t* pbound, int32_t* pindptr,
                         int32_t* pindices, T* pvals, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && A->indices && A->vals && pbound && pindptr && pindices && pvals,
             "bad arguments");
  SA_REQUIRE(npanels >= 1 && npanels <= 64, "npanels must be in 1..64");
  hipLaunchKernelGGL(panel_scatter_kernel<T>, dim3(grid_for(A->n, SA_BLOCK, 8192)), dim3(SA_BLOCK), 0, stream, A->n,
                     A->indptr, A->indices, (const T*)A->vals, npanels, pbound, pindptr, pindices, pvals);
  return check_launch("panel_scatter");
}

// ---------------------------------------------------------------------------------------------
// Allocation-free variant of the fused FP32 SpMV: the prescaled vector s (.) x, which
// sa_spmv_scaled_f32 takes as the caller's scratch z, lives in a file-scope device array instead. Each device that
// loads the library holds its own copy until the process ends. Its device address is looked up once
// per device (hipGetSymbolAddress with that device current) on the first call there; host code only
// ever passes that looked-up address, never the array's name.
// ---------------------------------------------------------------------------------------------
constexpr size_t kStaticScratchBytes = (size_t)32 << 20;   // n <= 2^23 FP32 entries
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[kMaxDevices];

static int static_scratch(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("spmv_scaled_static: no current HIP device (or device index >= 64)");
    return SA_E_LAUNCH;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("spmv_scaled_static: hipGetSymbolAddress of the static scratch failed: ") +
                hipGetErrorString(e));
      return SA_E_LAUNCH;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SA_OK;
}

template <typename T>
static int spmv_scaled_static(const sa_csr* A, const T* s, const T* x, double a, double b, T* y,
                              hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0, "bad arguments");
  SA_REQUIRE((size_t)A->n * sizeof(T) <= kStaticScratchBytes,
             "n exceeds the static scratch of the allocation-free SpMV (32 MiB: n <= 2^23); "
             "use sa_spmv_scaled_f32 with a caller buffer");
  void* z = nullptr;
  const int rc = static_scratch(&z);
  if (rc) return rc;
  return spmv_scaled<T>(A, s, x, a, b, static_cast<T*>(z), y, stream);
}

}  // namespace sa

extern "C" {

int sa_spmv_scaled_static_f32(const sa_csr* A, const float* s, const float* x, double a, double b, float* y,
                              hipStream_t stream) {
  return sa::spmv_scaled_static<float>(A, s, x, a, b, y, stream);
}

int sa_spmv_scaled_f32(const sa_csr* A, const float* s, const float* x, double a, double b, float* z, float* y,
                       hipStream_t stream) {
  return sa::spmv_scaled<float>(A, s, x, a, b, z, y, stream);
}

int sa_spmv_scaled_f64(const sa_csr* A, const double* s, const double* x, double a, double b, double* z, double* y,
                       hipStream_t stream) {
  return sa::spmv_scaled<double>(A, s, x, a, b, z, y, stream);
}

int sa_spmv_gather_f32(const sa_csr* A, const float* s, const float* g, double a, const float* x, double b, float* y,
                       hipStream_t stream) {
  return sa::spmv_gather<float>(A, s, g, a, x, b, y, stream);
}

int sa_spmv_gather_f64(const sa_csr* A, const double* s, const double* g, double a, const double* x, double b,
                       double* y, hipStream_t stream) {
  return sa::spmv_gather<double>(A, s, g, a, x, b, y, stream);
}

int sa_panel_count(const sa_csr* A, int32_t npanels, const int32_t* pbound, int32_t* counts, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->indptr && A->indices && pbound && counts, "bad arguments");
  SA_REQUIRE(npanels >= 1 && npanels <= 64, "npanels must be in 1..64");
  hipLaunchKernelGGL(sa::panel_count_kernel, dim3(sa::grid_for(A->n, SA_BLOCK, 8192)), dim3(SA_BLOCK), 0, stream,
                     A->n, A->indptr, A->indices, npanels, pbound, counts);
  return sa::check_launch("panel_count");
}

int sa_panel_scatter_f32(const sa_csr* A, int32_t npanels, const int32_t* pbound, int32_t* pindptr, int32_t* pindices,
                         float* pvals, hipStream_t stream) {
  return sa::panel_scatter<float>(A, npanels, pbound, pindptr, pindices, pvals, stream);
}

int sa_panel_scatter_f64(const sa_csr* A, int32_t npanels, const int32_t* pbound, int32_t* pindptr, int32_t* pindices,
                         double* pvals, hipStream_t stream) {
  return sa::panel_scatter<double>(A, npanels, pbound, pindptr, pindices, pvals, stream);
}

int sa_panel_sign_fold_f32(int64_t nnz, int32_t* pindices, const float* pvals, double mag, int32_t* mismatch,
                           hipStream_t stream) {
  SA_REQUIRE(nnz >= 0 && (nnz == 0 || (pindices && pvals)) && mismatch && mag > 0.0, "bad arguments");
  if (nnz == 0) return SA_OK;
  hipLaunchKernelGGL(sa::panel_sign_fold_kernel<float>, dim3(sa::grid_for(nnz, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK), 0,
                     stream, nnz, pindices, pvals, (float)mag, mismatch);
  return sa::check_launch("panel_sign_fold_f32");
}

int sa_panel_sign_fold_f64(int64_t nnz, int32_t* pindices, const double* pvals, double mag, int32_t* mismatch,
                           hipStream_t stream) {
  SA_REQUIRE(nnz >= 0 && (nnz == 0 || (pindices && pvals)) && mismatch && mag > 0.0, "bad arguments");
  if (nnz == 0) return SA_OK;
  hipLaunchKernelGGL(sa::panel_sign_fold_kernel<double>, dim3(sa::grid_for(nnz, SA_BLOCK * 4, 4096)), dim3(SA_BLOCK),
                     0, stream, nnz, pindices, pvals, mag, mismatch);
  return sa::check_launch("panel_sign_fold_f64");
}

int sa_panel_tiles(int64_t n, const int32_t* pindptr_p, int32_t tile_nnz, int32_t* tile_rows, int64_t ntiles,
                   hipStream_t stream) {
  SA_REQUIRE(n > 0 && pindptr_p && tile_rows && ntiles >= 1, "bad arguments");
  SA_REQUIRE(tile_nnz == 512 || tile_nnz == 1024 || tile_nnz == 2048 || tile_nnz == 4096,
             "tile_nnz must be 512, 1024, 2048 or 4096");
  hipLaunchKernelGGL(sa::panel_tiles_kernel, dim3(sa::grid_for(ntiles + 1, SA_BLOCK, 4096)), dim3(SA_BLOCK), 0,
                     stream, n, pindptr_p, tile_nnz, tile_rows, ntiles);
  return sa::check_launch("panel_tiles");
}

}  // extern "C"
