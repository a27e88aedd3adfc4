// Two-pass (propagation-blocked) SpMV for gfx950, for vectors far larger than an XCD's L2.
//
// When the gathered vector g (134 MB at n = 2^25) is far larger than the L2s, every layout that
// gathers g per entry pays a random Infinity-Cache or HBM access per stored coupling. This layout
// splits the product in two streaming passes over a fixed intermediate:
//   pass one: workgroup C copies column block C of g (at most 32768 columns, 128 KiB; the blocks are cut
//     to hold about equal entry counts, so hub columns sit in narrow ones) into LDS once and walks
//     the entries of those columns in stored order, writing each entry's product J_rj g_j to a
//     products array.
//     The products are written in row-block order (pass one's tiles each land in one contiguous run
//     of their row block's range), so that
//   pass two: workgroup R streams the contiguous products of its row block (column block by column
//     block, each tile running by row), sums each row's run in a wave, adds it into FP64 row sums in
//     LDS with one atomic add and applies the epilogue y = a x + b s (.) (row sums).
// The stream is 2 bytes per entry (local column and sign) in pass one, 4 bytes of product written and
// read back and 2 bytes of local row in pass two: 12 bytes per entry, all sequential, against one
// random line per entry for the gathers. The row sums are exact FP64 sums of FP32 products, so the
// atomic adds' order does not change them (the column-streamed layout's argument).
#include "sa_spmv.hpp"

namespace sa {

constexpr int kPbThreads = 1024;
constexpr int kPbColBits = 15;
constexpr int kPbCols = 1 << kPbColBits;   // column block: 128 KiB of g in LDS
constexpr int kPbRows = 20416;             // row block: FP64 row sums + the padding slot in LDS
constexpr int kPbKeyTile = 30;             // key: tile << 30 | local column << 15 | local row
constexpr int kPbUnroll = 16;              // entries per lane in flight (pass one)
constexpr int kPbUnroll2 = 16;             // entries per lane in flight (pass two)
constexpr int kPbMaxRowBlocks = 3840;      // pass one keeps two int32 per row block in LDS

typedef float pb_f4 __attribute__((ext_vector_type(4)));

// Pass one. The column block's tile starts (stored order) and the tiles' places in the products array
// (row-block order) sit in LDS beside the block of g.
template <bool kSigned>
__global__ __launch_bounds__(kPbThreads) void pb_expand_kernel(SpmvParams<float> p) {
  __shared__ __attribute__((aligned(16))) float xs[kPbCols];
  __shared__ int32_t ct[kPbMaxRowBlocks + 1];   // tile starts of this column block, stored order
  __shared__ int32_t cr[kPbMaxRowBlocks];       // the same tiles' starts in the products array
  if (p.gate && *p.gate != SA_RUNNING) return;
  const int C = (int)blockIdx.x;
  const int64_t c0 = p.pbound[C];
  const int64_t width = (int64_t)p.pbound[C + 1] - c0;
  const int cnt = (int)(width < kPbCols ? width : kPbCols);  // at most kPbCols by construction
  const int tid = threadIdx.x;
  const int nr = p.bnr;  // at most kPbMaxRowBlocks (checked at launch)
  for (int i = tid * 4; i < cnt; i += kPbThreads * 4) {
    if (i + 4 <= cnt) {
      *reinterpret_cast<pb_f4*>(&xs[i]) = *reinterpret_cast<const pb_f4*>(&p.g[c0 + i]);
    } else {
      for (int q = i; q < cnt; ++q) xs[q] = p.g[c0 + q];
    }
  }
  for (int r = tid; r <= nr; r += kPbThreads) {
    ct[r] = p.pbtoff[(int64_t)C * nr + r];
    if (r < nr) cr[r] = p.pbrtoff[(int64_t)r * p.pbnc + C];
  }
  __syncthreads();
  const int64_t e0 = ct[0], e1 = ct[nr];
  if (e0 >= e1) return;
  // Wave w takes the windows of 64 x kPbUnroll consecutive entries w, w + 16, ...; lane l the entries
  // l, l + 64, ... of its window, so that every load and every run of stores is contiguous across the
  // wave. The tile of a window's first entry comes from a binary search over ct (uniform), and each
  // lane moves forward from it past the tile starts at or below its entry (a window of a few hundred
  // entries crosses a few short tiles).
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint16_t* __restrict__ ent = p.pbent;
  const float* __restrict__ vals = (const float*)p.bvals;
  float* __restrict__ out = p.pbcontrib;
  constexpr int64_t kWin = 64 * kPbUnroll;
  for (int64_t w0 = e0 + (int64_t)wave * kWin; w0 < e1; w0 += (int64_t)(kPbThreads / 64) * kWin) {
    uint32_t e[kPbUnroll];
    float wv[kPbUnroll];
#pragma unroll
    for (int u = 0; u < kPbUnroll; ++u) {
      const int64_t i = w0 + 64 * u + lane;
      const int64_t is = i < e1 ? i : e0;  // past the end: an entry of this block, not written
      e[u] = __builtin_nontemporal_load(&ent[is]);
      if constexpr (!kSigned) wv[u] = __builtin_nontemporal_load(&vals[is]);
    }
    int lo = 0, hi = nr - 1;  // the last tile starting at or below w0
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (ct[mid] <= w0) lo = mid; else hi = mid - 1;
    }
    int R = lo;
#pragma unroll
    for (int u = 0; u < kPbUnroll; ++u) {
      const int64_t i = w0 + 64 * u + lane;
      const float x = xs[e[u] & (kPbCols - 1)];
      float v;
      if constexpr (kSigned) v = (e[u] & 0x8000u) ? -x : x;
      else v = wv[u] * x;
      if (i < e1) {
        while (i >= ct[R + 1]) ++R;
        out[(int64_t)cr[R] + (i - ct[R])] = v;
      }
    }
  }
}

// Pass two: workgroup R streams its products; wave w takes every 16th run of 64 x kPbUnroll2.
template <bool kSigned>
__global__ __launch_bounds__(kPbThreads) void pb_reduce_kernel(SpmvParams<float> p) {
  __shared__ double acc[kPbRows + 1];
  if (p.gate && *p.gate != SA_RUNNING) return;
  double bb = p.b;
  if (p.lz) bb *= p.lz->sigma[p.lz->j];
  const int R = (int)blockIdx.x;
  const int64_t r0 = p.rbound ? (int64_t)p.rbound[R] : (int64_t)R * p.brows;
  if (r0 >= p.n) return;
  const int64_t r1 = p.rbound ? (int64_t)p.rbound[R + 1] : r0 + p.brows;
  const int rows = (int)((p.n < r1 ? p.n : r1) - r0);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < rows; i += kPbThreads) acc[i] = 0.0;
  __syncthreads();
  const float* __restrict__ prod = p.pbcontrib;
  const uint16_t* __restrict__ prow = p.pbrow;
  // the row block's products are contiguous (column block by column block, each running by row)
  const int64_t a = p.pbrtoff[(int64_t)R * p.pbnc], b = p.pbrtoff[(int64_t)(R + 1) * p.pbnc];
  for (int64_t i0 = a + (int64_t)wave * 64 * kPbUnroll2; i0 < b; i0 += (int64_t)kPbThreads * kPbUnroll2) {
    float v[kPbUnroll2];
    uint32_t r[kPbUnroll2];
#pragma unroll
    for (int u = 0; u < kPbUnroll2; ++u) {
      const int64_t i = i0 + 64 * u + lane;
      const int64_t is = i < b ? i : a;
      v[u] = __builtin_nontemporal_load(&prod[is]);
      r[u] = __builtin_nontemporal_load(&prow[is]);
    }
#pragma unroll
    for (int u = 0; u < kPbUnroll2; ++u) {
      // a row's run (consecutive entries of one row, possibly across two column blocks' tiles) is
      // summed onto its first lane in FP64, which adds once
      const int64_t i = i0 + 64 * u + lane;
      const bool ok = i < b;
      const uint32_t rn = (uint32_t)__shfl_down((int)r[u], 1, 64);
      const uint64_t m = __ballot(ok && lane < 63 && i + 1 < b && rn == r[u]);
      double d = ok ? (double)v[u] : 0.0;
      if (m) {
        uint64_t c = m;
        for (int dd = 1; c != 0; dd <<= 1) {
          const double w = __shfl_down(d, dd, 64);
          if ((c >> lane) & 1ull) d += w;
          c &= c >> dd;
        }
      }
      const bool head = ok && (lane == 0 || !((m >> (lane - 1)) & 1ull));
      if (head) atomicAdd(&acc[r[u]], d);
    }
  }
  __syncthreads();
  const double scale = kSigned ? (double)p.pmag : 1.0;
  for (int i = tid; i < rows; i += kPbThreads) {
    const int64_t r = r0 + i;
    const double sr = p.s ? (double)p.s[r] : 1.0;
    double v = bb * sr * (scale * acc[i]);
    if (p.x) v += p.a * (double)p.x[r];
    p.y[r] = (float)v;
  }
}

int launch_pblock_spmv(const SpmvParams<float>& p, hipStream_t stream) {
  if (p.npanels != 0 || !p.pbent || !p.pbrow || !p.pbtoff || !p.pbcontrib || (!p.bvals && !p.psigned) || !p.y) {
    set_error("spmv: incomplete two-pass layout");
    return SA_E_ARG;
  }
  if (p.brows < 64 || p.brows > kPbRows || p.bnr < 1 || p.bnr > kPbMaxRowBlocks ||
      (!p.rbound && (int64_t)p.bnr * p.brows < p.n) || p.pbnc < 1 || !p.pbound || !p.pbrtoff ||
      (int64_t)p.pbnc * p.bnr >= (1ll << 31)) {
    set_error("spmv: two-pass layout sizes out of range");
    return SA_E_ARG;
  }
  if (!aligned16(p.g) || !aligned16(p.pbent) || !aligned16(p.pbcontrib) || (p.bvals && !aligned16(p.bvals))) {
    set_error("spmv: the two-pass layout needs 16-byte aligned vectors and entries");
    return SA_E_ARG;
  }
  const dim3 block(kPbThreads);
  if (p.psigned) hipLaunchKernelGGL(pb_expand_kernel<true>, dim3((unsigned)p.pbnc), block, 0, stream, p);
  else hipLaunchKernelGGL(pb_expand_kernel<false>, dim3((unsigned)p.pbnc), block, 0, stream, p);
  int rc = check_launch("spmv_pblock_expand");
  if (rc) return rc;
  if (p.psigned) hipLaunchKernelGGL(pb_reduce_kernel<true>, dim3((unsigned)p.bnr), block, 0, stream, p);
  else hipLaunchKernelGGL(pb_reduce_kernel<false>, dim3((unsigned)p.bnr), block, 0, stream, p);
  return check_launch("spmv_pblock_reduce");
}

// ---- layout construction ---------------------------------------------------------------------
__device__ inline int pb_row_block(const int32_t* rbound, int32_t brows, int32_t bnr, int64_t r) {
  if (!rbound) return (int)(r / brows);
  int lo = 0, hi = bnr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rbound[mid] <= r) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// key = (column block * bnr + row block) << 30 | local row << 15 | column in the block (thread per row):
// inside a tile the entries run by row, so pass two sums a row's run before its one atomic add.
// Column block of column j: cgran[j / 64] (the block boundaries are multiples of 64).
__global__ __launch_bounds__(SA_BLOCK) void pb_key_kernel(sa_csr A, int32_t brows, int32_t bnr,
                                                          const int32_t* __restrict__ cbound,
                                                          const int32_t* __restrict__ cgran,
                                                          uint64_t* __restrict__ keys, int32_t* __restrict__ idx,
                                                          int32_t* __restrict__ flag) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < A.n; r += (int64_t)gridDim.x * blockDim.x) {
    const int rb = pb_row_block(A.rbound, brows, bnr, r);
    const int64_t lr = r - (A.rbound ? (int64_t)A.rbound[rb] : (int64_t)rb * brows);
    if (lr >= brows || lr >= kPbRows) {
      atomicOr(flag, 1);
      continue;
    }
    for (int32_t e = A.indptr[r]; e < A.indptr[r + 1]; ++e) {
      const int64_t j = A.indices[e];
      const int32_t cb = cgran[j >> 6];
      const int64_t lc = j - cbound[cb];
      if (lc < 0 || lc >= kPbCols) {
        atomicOr(flag, 2);
        continue;
      }
      const uint64_t t = (uint64_t)cb * (uint64_t)bnr + (uint64_t)rb;
      keys[e] = (t << kPbKeyTile) | ((uint64_t)lr << 15) | (uint64_t)lc;
      idx[e] = e;
    }
  }
}

// Sorted entry i (stored order) of tile t = C * bnr + R, rank k in its tile: pbent[i] (and the value),
// and its local row at the tile's place in the products array, pbrow[rtoff[R * ncb + C] + k].
__global__ __launch_bounds__(SA_BLOCK) void pb_fill_kernel(int64_t nnz, const uint64_t* __restrict__ keys,
                                                           const int32_t* __restrict__ idx,
                                                           const int32_t* __restrict__ toff,
                                                           const int32_t* __restrict__ rtoff, int32_t bnr,
                                                           int32_t ncb, const float* __restrict__ vals, float mag,
                                                           uint16_t* __restrict__ pbent, uint16_t* __restrict__ pbrow,
                                                           float* __restrict__ bvals, int32_t* __restrict__ flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz + 8; i += (int64_t)gridDim.x * blockDim.x) {
    if (i >= nnz) {  // the padding group read by the last 16-byte load
      pbent[i] = 0;
      if (bvals) bvals[i] = 0.f;
      continue;
    }
    const uint64_t key = keys[i];
    const float v = vals[idx[i]];
    uint32_t code = (uint32_t)(key & (kPbCols - 1));
    if (mag > 0.f) {
      if ((v < 0.f ? -v : v) != mag) atomicOr(flag, 8);
      if (v < 0.f) code |= 0x8000u;
    } else {
      bvals[i] = v;
    }
    pbent[i] = (uint16_t)code;
    const int64_t t = (int64_t)(key >> kPbKeyTile);
    const int64_t R = t % bnr, C = t / bnr;
    pbrow[(int64_t)rtoff[R * ncb + C] + (i - toff[t])] = (uint16_t)((key >> 15) & 0x7FFFu);
  }
}

}  // namespace sa

extern "C" {

int sa_pblock_cols(void) { return sa::kPbCols; }

int sa_pblock_plan(const sa_csr* A, int32_t brows, int32_t bnr, const int32_t* cbound, int32_t ncb,
                   const int32_t* cgran, uint64_t* keys, int32_t* idx, uint64_t* keys_out, int32_t* idx_out, void* tmp,
                   size_t tmp_bytes, int32_t* toff, int32_t* flag, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->nnz >= 0 && A->indptr && (A->indices || A->nnz == 0), "bad arguments");
  SA_REQUIRE(brows >= 64 && brows <= sa::kPbRows, "brows must be in [64, 20416]");
  SA_REQUIRE(cbound && cgran && ncb >= 1, "column blocks required");
  const int64_t nc = ncb;
  SA_REQUIRE(bnr >= 1 && (A->rbound || (int64_t)bnr * brows >= A->n) && nc * bnr < (1ll << 31),
             "bnr row blocks must cover the rows");
  SA_REQUIRE(keys && idx && keys_out && idx_out && toff && flag, "bad arguments");
  const int64_t ntiles = nc * bnr;
  if (A->nnz > 0) {
    hipLaunchKernelGGL(sa::pb_key_kernel, dim3(sa::grid_for(A->n, SA_BLOCK, 8192)), dim3(SA_BLOCK), 0, stream, *A,
                       brows, bnr, cbound, cgran, keys, idx, flag);
    const int rc = sa::check_launch("pblock_keys");
    if (rc) return rc;
  }
  return sa::sort_entry_keys(keys, idx, keys_out, idx_out, A->nnz, sa::kPbKeyTile + sa::key_bits(ntiles + 1), tmp,
                             tmp_bytes, sa::kPbKeyTile, ntiles, toff, stream);
}

int sa_pblock_fill_f32(const sa_csr* A, int32_t brows, int32_t bnr, int32_t ncb, const uint64_t* keys,
                       const int32_t* idx, const int32_t* toff, const int32_t* rtoff, uint16_t* pbent,
                       uint16_t* pbrow, float* bvals, double mag, int32_t* flag, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->nnz >= 0 && A->vals && brows >= 64 && bnr >= 1 && ncb >= 1, "bad arguments");
  SA_REQUIRE(keys && idx && toff && rtoff && pbent && pbrow && flag && (mag > 0.0 || bvals), "bad arguments");
  hipLaunchKernelGGL(sa::pb_fill_kernel, dim3(sa::grid_for(A->nnz + 8, SA_BLOCK, 8192)), dim3(SA_BLOCK), 0, stream,
                     A->nnz, keys, idx, toff, rtoff, bnr, ncb, (const float*)A->vals, (float)(mag > 0.0 ? mag : 0.0),
                     pbent, pbrow, mag > 0.0 ? (float*)nullptr : bvals, flag);
  return sa::check_launch("pblock_fill");
}

}  // extern "C"
