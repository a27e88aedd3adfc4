// Column-streamed SpMV for gfx950: the gathers of the fused scaled SpMV served from L2 in column order.
//
// Workgroup w = rb * 8 + p runs on XCD p (workgroups reach the XCDs round-robin) and sums the entries
// of row block rb that lie in column panel p; the panel's slice of the gathered vector g stays in the
// XCD's L2. Its entries are sorted by column (then row) and cut into 256-entry chunks that the 16
// waves take in turn, so the 64 lanes of one gather read about 330 consecutive columns of g (ten or so
// cache lines at C4's density) instead of 64 random lines, and every wave walks the panel left to
// right at about the same pace as the others. There is no staging of g and no barrier inside the
// walk: each wave keeps six chunks (four with values; 24 KiB of entries per CU) and their gathers in flight.
//
// The row sums live in LDS in FP64 and take atomic adds (ds_add_f64) from any wave. Every addend is an
// FP32 value (+-g_j, or the FP32 product J_rj g_j), and a sum of FP32 values in FP64 is exact while the
// row's addends in one panel span fewer than 29 - log2(count) binades, so the sums, and the FP32
// partials rounded from them, do not depend on the order of the atomic adds: the results are bitwise
// reproducible for every input whose addends stay within that range (beyond it an FP64 sum may differ
// in its last bit between runs, and its FP32 rounding almost never). The FP32 partials go to
// partial[p * pld + r] like the other panel layouts', and the same consumers combine the panels.
//
// The row-owned form (bmode = 4) is the same walk with one workgroup per row block over every column:
// the layout is built with a single panel (panel 0 = all columns), the row blocks come in whole rounds
// of one workgroup per CU, and each workgroup's waves walk its columns left to right, so the 32 CUs of
// an XCD sweep the gathered vector through their L2 at about the same place. Its row sums are final:
// the epilogue writes y = a x + b s (.) (row sums) directly, with no panel partials and no combine (the
// panel form writes and the combine reads back 8 partial vectors; the row-owned form gathers at a
// lower column density, fewer rows per workgroup, and reads g once per XCD).
#include <cstring>  // rocprim's texture cache iterator calls memset
#include <rocprim/rocprim.hpp>

#include "sa_spmv.hpp"

namespace sa {

constexpr int kCsThreads = 1024;           // 16 waves: one workgroup per CU (the row sums fill LDS)
constexpr int kCsWaves = kCsThreads / 64;
#ifndef SA_CS_UNROLL_SIGNED
#define SA_CS_UNROLL_SIGNED 6
#endif
constexpr int kCsUnrollSigned = SA_CS_UNROLL_SIGNED;  // chunks each wave has in flight (sign-packed entries)
constexpr int kCsUnrollValued = 4;         // with a value per entry (more registers per chunk)
constexpr int kCsRows = 20416;             // FP64 row sums + the padding slot: 163336 bytes of LDS
constexpr int kCsPad = kCsRows;            // local row of padding entries
constexpr int kCsChunk = 256;
constexpr int kCsKeyWg = 40;               // key bits: workgroup << 40 | panel column << 15 | local row
constexpr int kCsColBits = 25;

typedef uint32_t cs_u4 __attribute__((ext_vector_type(4)));
typedef float cs_f4 __attribute__((ext_vector_type(4)));

// kOwned: workgroup = row block over every column (bmode 4), final rows; else workgroup = (row block,
// panel) on that panel's XCD (bmode 2), panel partials.
template <bool kSigned, bool kOwned>
__global__ __launch_bounds__(kCsThreads) void spmv_cstream_kernel(SpmvParams<float> p) {
  constexpr int kCsUnroll = kSigned ? kCsUnrollSigned : kCsUnrollValued;
  __shared__ double acc[kCsRows + 1];
  if (p.gate && *p.gate != SA_RUNNING) return;
  const int panel = kOwned ? 0 : (int)(blockIdx.x & 7);
  const int rb = kOwned ? (int)blockIdx.x : (int)(blockIdx.x >> 3);
  const int64_t seg = kOwned ? (int64_t)8 * blockIdx.x : (int64_t)blockIdx.x;  // its chunks: segment seg
  const int64_t r0 = p.rbound ? (int64_t)p.rbound[rb] : (int64_t)rb * p.brows;
  if (r0 >= p.n) return;
  const int64_t r1 = p.rbound ? (int64_t)p.rbound[rb + 1] : r0 + p.brows;
  const int rows = (int)((p.n < r1 ? p.n : r1) - r0);  // at most brows <= kCsRows (checked at launch)
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < rows; i += kCsThreads) acc[i] = 0.0;  // the padding slot is never read
  __syncthreads();

  const int64_t k0 = p.bchunk[seg], k1 = p.bchunk[seg + 1];
  const float* __restrict__ gp = p.g + p.pbound[panel];
  const cs_u4* __restrict__ ent = reinterpret_cast<const cs_u4*>(p.bent);
  const cs_f4* __restrict__ val = reinterpret_cast<const cs_f4*>(p.bvals);
  cs_u4 e[kCsUnroll];
  cs_f4 w[kCsUnroll];
  int32_t cb[kCsUnroll];
  // chunks k, k + 16, ... of this wave; one past the end reads chunk k0 (it exists) and is dropped
  auto fetch = [&](int64_t k) {
#pragma unroll
    for (int u = 0; u < kCsUnroll; ++u) {
      const int64_t kk = k + (int64_t)u * kCsWaves;
      const int64_t ks = kk < k1 ? kk : k0;
      e[u] = __builtin_nontemporal_load(&ent[ks * 64 + lane]);
      if constexpr (!kSigned) w[u] = __builtin_nontemporal_load(&val[ks * 64 + lane]);
      cb[u] = p.bcbase[ks];
    }
  };
  int64_t k = k0 + wave;
  if (k < k1) fetch(k);
  for (; k < k1; k += (int64_t)kCsUnroll * kCsWaves) {
    float x[kCsUnroll][4];
    cs_u4 ec[kCsUnroll];
    cs_f4 wc[kCsUnroll];
#pragma unroll
    for (int u = 0; u < kCsUnroll; ++u) {
      ec[u] = e[u];
      if constexpr (!kSigned) wc[u] = w[u];
#pragma unroll
      for (int j = 0; j < 4; ++j) x[u][j] = gp[cb[u] + (int)(ec[u][j] & 0xFFFFu)];
    }
    // the next chunks' entries are issued after this round's gathers, so waiting for the gathers
    // below does not wait for them
    if (k + (int64_t)kCsUnroll * kCsWaves < k1) fetch(k + (int64_t)kCsUnroll * kCsWaves);
#pragma unroll
    for (int u = 0; u < kCsUnroll; ++u) {
      if (k + (int64_t)u * kCsWaves < k1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float v;
          if constexpr (kSigned) v = (ec[u][j] & 0x80000000u) ? -x[u][j] : x[u][j];
          else v = wc[u][j] * x[u][j];
          atomicAdd(&acc[(ec[u][j] >> 16) & 0x7FFFu], (double)v);
        }
      }
    }
  }
  __syncthreads();
  const double scale = kSigned ? (double)p.pmag : 1.0;
  if constexpr (kOwned) {
    double bb = p.b;
    if (p.lz) bb *= p.lz->sigma[p.lz->j];
    for (int i = tid; i < rows; i += kCsThreads) {
      const int64_t r = r0 + i;
      const double sr = p.s ? (double)p.s[r] : 1.0;
      double v = bb * sr * (scale * acc[i]);
      if (p.x) v += p.a * (double)p.x[r];
      p.y[r] = (float)v;
    }
  } else {
    float* out = p.partial + (int64_t)panel * p.pld + r0;
    for (int i = tid; i < rows; i += kCsThreads) out[i] = (float)(scale * acc[i]);
  }
}

// Instance-streamed SpMV of a block-diagonal batch (imode = 2): the same walk per workgroup of one
// instance's rows. Workgroup w (tile_list[blockIdx.x] when a list is given) owns rows [irow[w],
// min(irow[w] + irows, iend[iinst[w]])) and the chunks [bchunk[8 w], bchunk[8 w + 1]) of the
// column-streamed layout built with those row blocks and one panel (every column in panel 0), entries
// with values. An instance's columns are its rows' (block diagonal), so a row block's columns are
// dense (tens of entries per column for instances of a few thousand nodes) and a 64-lane gather
// reads one or two lines. The workgroup does nothing unless its instance is running, and the epilogue
// is the plain CSR kernel's: y = a x + b s (.) (row sums); padding rows get y = a x. kSigned: every
// coupling of an instance has one magnitude (imag[instance]); the entries carry the sign bit, no values.
template <bool kSigned>
__global__ __launch_bounds__(kCsThreads) void spmv_istream_kernel(SpmvParams<float> p) {
  constexpr int kU = kSigned ? kCsUnrollSigned : kCsUnrollValued;
  __shared__ double acc[kCsRows + 1];
  if (p.gate && *p.gate != SA_RUNNING) return;
  double bb = p.b;
  if (p.lz) bb *= p.lz->sigma[p.lz->j];
  const int64_t w = p.tile_list ? (int64_t)p.tile_list[blockIdx.x] : (int64_t)blockIdx.x;
  const int inst = p.iinst[w];
  if (p.gate_base && *reinterpret_cast<const int32_t*>(p.gate_base + (int64_t)inst * p.gate_stride) != SA_RUNNING)
    return;
  const int64_t r0 = p.irow[w];
  const int64_t end = p.iend[inst];
  const int rows = (int)((r0 + p.irows < end ? r0 + p.irows : end) - r0);  // at most irows <= kCsRows
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < rows; i += kCsThreads) acc[i] = 0.0;
  __syncthreads();
  const int64_t k0 = p.bchunk[8 * w], k1 = p.bchunk[8 * w + 1];
  const float* __restrict__ gp = p.g;
  const cs_u4* __restrict__ ent = reinterpret_cast<const cs_u4*>(p.bent);
  const cs_f4* __restrict__ val = reinterpret_cast<const cs_f4*>(p.bvals);
  cs_u4 e[kU];
  cs_f4 wv[kU];
  int32_t cb[kU];
  auto fetch = [&](int64_t k) {
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t kk = k + (int64_t)u * kCsWaves;
      const int64_t ks = kk < k1 ? kk : k0;
      e[u] = __builtin_nontemporal_load(&ent[ks * 64 + lane]);
      if constexpr (!kSigned) wv[u] = __builtin_nontemporal_load(&val[ks * 64 + lane]);
      cb[u] = p.bcbase[ks];
    }
  };
  int64_t k = k0 + wave;
  if (k < k1) fetch(k);
  for (; k < k1; k += (int64_t)kU * kCsWaves) {
    float x[kU][4];
    cs_u4 ec[kU];
    cs_f4 wc[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      ec[u] = e[u];
      if constexpr (!kSigned) wc[u] = wv[u];
#pragma unroll
      for (int j = 0; j < 4; ++j) x[u][j] = gp[cb[u] + (int)(ec[u][j] & 0xFFFFu)];
    }
    if (k + (int64_t)kU * kCsWaves < k1) fetch(k + (int64_t)kU * kCsWaves);
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (k + (int64_t)u * kCsWaves < k1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float v;
          if constexpr (kSigned) v = (ec[u][j] & 0x80000000u) ? -x[u][j] : x[u][j];
          else v = wc[u][j] * x[u][j];
          atomicAdd(&acc[(ec[u][j] >> 16) & 0x7FFFu], (double)v);
        }
      }
    }
  }
  __syncthreads();
  const double scale = kSigned ? (double)p.imag[inst] : 1.0;
  for (int i = tid; i < rows; i += kCsThreads) {
    const int64_t r = r0 + i;
    double v = bb * (p.s ? (double)p.s[r] : 1.0) * (scale * acc[i]);
    if (p.x) v += p.a * (double)p.x[r];
    p.y[r] = (float)v;
  }
}

int launch_istream_spmv(const SpmvParams<float>& p, hipStream_t stream) {
  if (!p.iinst || !p.irow || !p.iend || !p.bchunk || !p.bent || !p.bcbase || !p.y ||
      (p.isigned ? !p.imag : !p.bvals)) {
    set_error("spmv: incomplete instance-streamed layout");
    return SA_E_ARG;
  }
  if (p.irows < 64 || p.irows > kCsRows || p.inwg < 1 || p.inwg >= (1ll << 28)) {
    set_error("spmv: instance-streamed layout sizes out of range");
    return SA_E_ARG;
  }
  if (!aligned16(p.bent) || (!p.isigned && !aligned16(p.bvals))) {
    set_error("spmv: the instance-streamed layout needs 16-byte aligned entries and values");
    return SA_E_ARG;
  }
  if (p.tile_list && (p.ntile_list < 0 || p.ntile_list > p.inwg)) {
    set_error("spmv: tile list longer than the layout's workgroups");
    return SA_E_ARG;
  }
  const int64_t nwg = p.tile_list ? p.ntile_list : p.inwg;
  if (nwg == 0) return SA_OK;
  if (p.isigned) hipLaunchKernelGGL(spmv_istream_kernel<true>, dim3((unsigned)nwg), dim3(kCsThreads), 0, stream, p);
  else hipLaunchKernelGGL(spmv_istream_kernel<false>, dim3((unsigned)nwg), dim3(kCsThreads), 0, stream, p);
  return check_launch("spmv_istream");
}

int launch_cstream_spmv(const SpmvParams<float>& p, hipStream_t stream) {
  if (p.npanels != 8 || !p.pbound || !p.bchunk || !p.bent || !p.bcbase || (!p.bvals && !p.psigned) || !p.partial) {
    set_error("spmv: incomplete column-streamed layout");
    return SA_E_ARG;
  }
  if (p.brows < 64 || p.brows > kCsRows || p.bnr < 1 || (!p.rbound && (int64_t)p.bnr * p.brows < p.n) ||
      (int64_t)8 * p.bnr >= (1ll << 31)) {
    set_error("spmv: column-streamed layout sizes out of range");
    return SA_E_ARG;
  }
  if (!aligned16(p.bent) || (p.bvals && !aligned16(p.bvals))) {
    set_error("spmv: the column-streamed layout needs 16-byte aligned entries and values");
    return SA_E_ARG;
  }
  const dim3 grid((unsigned)(8 * p.bnr)), block(kCsThreads);
  if (p.psigned) hipLaunchKernelGGL((spmv_cstream_kernel<true, false>), grid, block, 0, stream, p);
  else hipLaunchKernelGGL((spmv_cstream_kernel<false, false>), grid, block, 0, stream, p);
  return check_launch("spmv_cstream");
}

int launch_rstream_spmv(const SpmvParams<float>& p, hipStream_t stream) {
  if (p.npanels != 0 || !p.pbound || !p.bchunk || !p.bent || !p.bcbase || (!p.bvals && !p.psigned) || !p.y) {
    set_error("spmv: incomplete row-owned column-streamed layout");
    return SA_E_ARG;
  }
  if (p.brows < 64 || p.brows > kCsRows || p.bnr < 1 || (!p.rbound && (int64_t)p.bnr * p.brows < p.n) ||
      (int64_t)8 * p.bnr >= (1ll << 31)) {
    set_error("spmv: row-owned column-streamed layout sizes out of range");
    return SA_E_ARG;
  }
  if (!aligned16(p.bent) || (p.bvals && !aligned16(p.bvals))) {
    set_error("spmv: the column-streamed layout needs 16-byte aligned entries and values");
    return SA_E_ARG;
  }
  const dim3 grid((unsigned)p.bnr), block(kCsThreads);
  if (p.psigned) hipLaunchKernelGGL((spmv_cstream_kernel<true, true>), grid, block, 0, stream, p);
  else hipLaunchKernelGGL((spmv_cstream_kernel<false, true>), grid, block, 0, stream, p);
  return check_launch("spmv_rstream");
}

// ---- layout construction ---------------------------------------------------------------------
// Row block of row r: r / brows, or the block [rbound[rb], rbound[rb + 1]) that holds r.
__device__ inline int cs_row_block(const int32_t* rbound, int32_t brows, int32_t bnr, int64_t r) {
  if (!rbound) return (int)(r / brows);
  int lo = 0, hi = bnr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rbound[mid] <= r) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One key per stored entry: workgroup (row block * 8 + panel) << 40 | column in the panel << 15 |
// local row, and the entry's index (one thread per row; a row's columns are sorted, so its panel only
// moves forward).
__global__ __launch_bounds__(SA_BLOCK) void cs_key_kernel(sa_csr A, const int32_t* __restrict__ pbound, int32_t brows,
                                                          int32_t bnr, uint64_t* __restrict__ keys,
                                                          int32_t* __restrict__ idx, int32_t* __restrict__ flag) {
  __shared__ int32_t bnd[9];
  if (threadIdx.x < 9) bnd[threadIdx.x] = pbound[threadIdx.x];
  __syncthreads();
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < A.n; r += (int64_t)gridDim.x * blockDim.x) {
    const int rb = cs_row_block(A.rbound, brows, bnr, r);
    const int64_t lr = r - (A.rbound ? (int64_t)A.rbound[rb] : (int64_t)rb * brows);
    if (lr >= brows || lr >= kCsRows) {
      atomicOr(flag, 1);
      continue;
    }
    int pnl = 0;
    for (int32_t e = A.indptr[r]; e < A.indptr[r + 1]; ++e) {
      const int32_t j = A.indices[e];
      while (pnl < 7 && j >= bnd[pnl + 1]) ++pnl;
      const int64_t cr = (int64_t)j - bnd[pnl];
      if (cr >= (1ll << kCsColBits)) atomicOr(flag, 2);
      keys[e] = ((uint64_t)(rb * 8 + pnl) << kCsKeyWg) | ((uint64_t)cr << 15) | (uint64_t)lr;
      idx[e] = e;
    }
  }
}

// start[w] = the first of the sorted keys whose segment (key >> shift) is w; start[nseg] = nnz.
__global__ __launch_bounds__(SA_BLOCK) void segment_start_kernel(int64_t nnz, const uint64_t* __restrict__ keys,
                                                                 int shift, int64_t nseg,
                                                                 int32_t* __restrict__ start) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t w = (int64_t)(keys[i] >> shift);
    const int64_t wp = i ? (int64_t)(keys[i - 1] >> shift) : -1;
    for (int64_t q = wp + 1; q <= w && q <= nseg; ++q) start[q] = (int32_t)i;
    if (i == nnz - 1)
      for (int64_t q = w + 1; q <= nseg; ++q) start[q] = (int32_t)nnz;
  }
}

__global__ __launch_bounds__(SA_BLOCK) void cs_fill_value_kernel(int64_t count, int32_t v, int32_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = v;
}

__global__ __launch_bounds__(SA_BLOCK) void cs_pad_kernel(int64_t nent, uint32_t* __restrict__ bent,
                                                          float* __restrict__ bvals) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nent; i += (int64_t)gridDim.x * blockDim.x) {
    bent[i] = (uint32_t)kCsPad << 16;
    if (bvals) bvals[i] = 0.f;
  }
}

// Sorted entry i of workgroup w is entry q = i - start[w] of the workgroup's stream: chunk
// bchunk[w] + q / 256, position q % 256 = 64 j + l stored at 4 l + j. The first entry of each chunk
// gives its base column.
__device__ inline bool cs_place(int64_t i, const uint64_t* keys, const int32_t* start, const int32_t* bchunk,
                                int64_t& chunk, int& r, int64_t& col) {
  const uint64_t key = keys[i];
  const int64_t w = (int64_t)(key >> kCsKeyWg);
  const int64_t q = i - start[w];
  chunk = (int64_t)bchunk[w] + q / kCsChunk;
  r = (int)(q % kCsChunk);
  col = (int64_t)((key >> 15) & ((1ull << kCsColBits) - 1));
  return chunk < bchunk[w + 1];
}

__global__ __launch_bounds__(SA_BLOCK) void cs_base_kernel(int64_t nnz, const uint64_t* __restrict__ keys,
                                                           const int32_t* __restrict__ start,
                                                           const int32_t* __restrict__ bchunk,
                                                           int32_t* __restrict__ bcbase, int32_t* __restrict__ flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t chunk, col;
    int r;
    if (!cs_place(i, keys, start, bchunk, chunk, r, col)) {
      atomicOr(flag, 16);  // start / bchunk do not come from these keys
      continue;
    }
    if (r == 0) bcbase[chunk] = (int32_t)col;
  }
}

__global__ __launch_bounds__(SA_BLOCK) void cs_fill_kernel(int64_t nnz, const uint64_t* __restrict__ keys,
                                                           const int32_t* __restrict__ idx,
                                                           const int32_t* __restrict__ start,
                                                           const int32_t* __restrict__ bchunk,
                                                           const int32_t* __restrict__ bcbase,
                                                           const float* __restrict__ vals, float mag,
                                                           uint32_t* __restrict__ bent, float* __restrict__ bvals,
                                                           int32_t* __restrict__ flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t chunk, col;
    int r;
    if (!cs_place(i, keys, start, bchunk, chunk, r, col)) continue;  // flagged by cs_base_kernel
    const int64_t delta = col - bcbase[chunk];
    if (delta < 0 || delta > 0xFFFF) {
      atomicOr(flag, 4);
      continue;
    }
    const int64_t addr = chunk * kCsChunk + (r & 63) * 4 + (r >> 6);
    const float v = vals[idx[i]];
    uint32_t code = ((uint32_t)(keys[i] & 0x7FFFu) << 16) | (uint32_t)delta;
    if (mag > 0.f) {
      if ((v < 0.f ? -v : v) != mag) atomicOr(flag, 8);
      if (v < 0.f) code |= 0x80000000u;
    } else {
      bvals[addr] = v;
    }
    bent[addr] = code;
  }
}

int key_bits(int64_t count) {
  int b = 1;
  while ((1ll << b) < count) ++b;
  return b;
}

static int sort_bytes(int64_t nnz, int end_bit, size_t& bytes) {
  bytes = 0;
  const hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                                 (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)nnz, 0u,
                                                 (unsigned)end_bit, (hipStream_t)0);
  if (e != hipSuccess) {
    set_error(std::string("layout build: radix sort size query failed: ") + hipGetErrorString(e));
    return SA_E_LAUNCH;
  }
  return SA_OK;
}

int sort_entry_keys(const uint64_t* keys, const int32_t* idx, uint64_t* keys_out, int32_t* idx_out, int64_t nnz,
                    int end_bit, void* tmp, size_t tmp_bytes, int shift, int64_t nseg, int32_t* start,
                    hipStream_t stream) {
  size_t need = 0;
  int rc = sort_bytes(nnz, end_bit, need);
  if (rc) return rc;
  if (!tmp || tmp_bytes < need) {
    set_error("layout build: sort scratch smaller than sa_cstream_sort_bytes");
    return SA_E_ARG;
  }
  if (nnz == 0) {
    hipLaunchKernelGGL(cs_fill_value_kernel, dim3(grid_for(nseg + 1, SA_BLOCK, 1024)), dim3(SA_BLOCK), 0, stream,
                       nseg + 1, 0, start);
    return check_launch("segment_start");
  }
  size_t bytes = tmp_bytes;
  const hipError_t e = rocprim::radix_sort_pairs(tmp, bytes, keys, keys_out, idx, idx_out, (size_t)nnz, 0u,
                                                 (unsigned)end_bit, stream);
  if (e != hipSuccess) {
    set_error(std::string("layout build: radix sort failed: ") + hipGetErrorString(e));
    return SA_E_LAUNCH;
  }
  hipLaunchKernelGGL(segment_start_kernel, dim3(grid_for(nnz, SA_BLOCK, 8192)), dim3(SA_BLOCK), 0, stream, nnz,
                     (const uint64_t*)keys_out, shift, nseg, start);
  return check_launch("segment_start");
}

}  // namespace sa

extern "C" {

int sa_cstream_max_rows(void) { return sa::kCsRows; }

int sa_cstream_pad_row(void) { return sa::kCsPad; }

int sa_cstream_sort_bytes(int64_t nnz, size_t* bytes) {
  SA_REQUIRE(bytes && nnz >= 0 && nnz < (1ll << 31), "bad arguments");
  size_t b = 0;
  const int rc = sa::sort_bytes(nnz, 64, b);
  if (rc) return rc;
  *bytes = b;
  return SA_OK;
}

int sa_cstream_plan(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bnr, uint64_t* keys, int32_t* idx,
                    uint64_t* keys_out, int32_t* idx_out, void* tmp, size_t tmp_bytes, int32_t* start, int32_t* flag,
                    hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->nnz >= 0 && A->indptr && (A->indices || A->nnz == 0) && pbound, "bad arguments");
  SA_REQUIRE(brows >= 64 && brows <= sa::kCsRows, "brows must be in [64, sa_cstream_max_rows()]");
  SA_REQUIRE(bnr >= 1 && (A->rbound || (int64_t)bnr * brows >= A->n) && (int64_t)8 * bnr < (1ll << 23),
             "bnr row blocks must cover the rows");
  SA_REQUIRE(keys && idx && keys_out && idx_out && start && flag, "bad arguments");
  const int64_t nwg = (int64_t)8 * bnr;
  if (A->nnz > 0) {
    hipLaunchKernelGGL(sa::cs_key_kernel, dim3(sa::grid_for(A->n, SA_BLOCK, 8192)), dim3(SA_BLOCK), 0, stream, *A,
                       pbound, brows, bnr, keys, idx, flag);
    const int rc = sa::check_launch("cstream_keys");
    if (rc) return rc;
  }
  return sa::sort_entry_keys(keys, idx, keys_out, idx_out, A->nnz, sa::kCsKeyWg + sa::key_bits(nwg + 1), tmp,
                             tmp_bytes, sa::kCsKeyWg, nwg, start, stream);
}

int sa_cstream_fill_f32(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bnr, const uint64_t* keys,
                        const int32_t* idx, const int32_t* start, const int32_t* bchunk, int64_t nchunks,
                        uint32_t* bent, float* bvals, int32_t* bcbase, double mag, int32_t* flag,
                        hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->nnz >= 0 && A->vals && pbound && brows >= 64 && bnr >= 1, "bad arguments");
  SA_REQUIRE(keys && idx && start && bchunk && bent && bcbase && flag && (mag > 0.0 || bvals) && nchunks >= 0 &&
                 nchunks * sa::kCsChunk < (1ll << 33),
             "bad arguments");
  const int64_t nent = nchunks * sa::kCsChunk;
  if (nent > 0) {
    hipLaunchKernelGGL(sa::cs_pad_kernel, dim3(sa::grid_for(nent, SA_BLOCK * 4, 8192)), dim3(SA_BLOCK), 0, stream,
                       nent, bent, mag > 0.0 ? (float*)nullptr : bvals);
    int rc = sa::check_launch("cstream_pad");
    if (rc) return rc;
  }
  if (A->nnz == 0) return SA_OK;
  const dim3 grid(sa::grid_for(A->nnz, SA_BLOCK, 8192)), block(SA_BLOCK);
  hipLaunchKernelGGL(sa::cs_base_kernel, grid, block, 0, stream, A->nnz, keys, start, bchunk, bcbase, flag);
  int rc = sa::check_launch("cstream_base");
  if (rc) return rc;
  hipLaunchKernelGGL(sa::cs_fill_kernel, grid, block, 0, stream, A->nnz, keys, idx, start, bchunk,
                     (const int32_t*)bcbase, (const float*)A->vals, (float)(mag > 0.0 ? mag : 0.0), bent, bvals, flag);
  return sa::check_launch("cstream_fill");
}

}  // extern "C"
