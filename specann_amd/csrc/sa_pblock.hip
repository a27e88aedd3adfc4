// Two-pass (propagation-blocked) SpMV for gfx950, for vectors far larger than an XCD's L2.
//
// When the gathered vector g (134 MB at n = 2^25) is far larger than the L2s, every layout that
// gathers g per entry pays a random Infinity-Cache or HBM access per stored coupling. This layout
// splits the product in two streaming passes over a fixed intermediate:
//   pass one: workgroup C copies column block C of g (at most 32768 columns, 128 KiB; the blocks are cut
//     to hold about equal entry counts, so hub columns sit in narrow ones) into LDS once and walks
//     the entries of those columns in stored order, writing each entry's product J_rj g_j to the
//     next slot of a products array (coalesced: the products have the entries' order);
//   pass two: workgroup R walks the products of its row block, tile by tile (one tile per column
//     block, each a contiguous range, its entries running by row), sums each row's run in a wave,
//     adds it into FP64 row sums in LDS with one atomic add and applies the epilogue
//     y = a x + b s (.) (row sums).
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
constexpr int kPbUnroll = 8;               // 8-entry groups per thread in flight (pass one)
constexpr int kPbUnroll2 = 16;             // entries per lane in flight (pass two)

typedef uint16_t pb_u16x8 __attribute__((ext_vector_type(8)));
typedef float pb_f4 __attribute__((ext_vector_type(4)));

// Pass one. Groups of eight stored entries (16 bytes of pbent) per thread; the groups at the ends of
// the column block's range also hold entries of the neighbouring blocks, which are not written.
template <bool kSigned>
__global__ __launch_bounds__(kPbThreads) void pb_expand_kernel(SpmvParams<float> p) {
  __shared__ __attribute__((aligned(16))) float xs[kPbCols];
  if (p.gate && *p.gate != SA_RUNNING) return;
  const int C = (int)blockIdx.x;
  const int64_t c0 = p.pbound[C];
  const int64_t width = (int64_t)p.pbound[C + 1] - c0;
  const int cnt = (int)(width < kPbCols ? width : kPbCols);  // at most kPbCols by construction
  const int tid = threadIdx.x;
  for (int i = tid * 4; i < cnt; i += kPbThreads * 4) {
    if (i + 4 <= cnt) {
      *reinterpret_cast<pb_f4*>(&xs[i]) = *reinterpret_cast<const pb_f4*>(&p.g[c0 + i]);
    } else {
      for (int q = i; q < cnt; ++q) xs[q] = p.g[c0 + q];
    }
  }
  __syncthreads();
  const int64_t e0 = p.pbtoff[(int64_t)C * p.bnr], e1 = p.pbtoff[(int64_t)(C + 1) * p.bnr];
  if (e0 >= e1) return;
  const int64_t base = e0 & ~(int64_t)7;
  const pb_u16x8* __restrict__ ent = reinterpret_cast<const pb_u16x8*>(p.pbent);
  const float* __restrict__ vals = (const float*)p.bvals;
  float* __restrict__ out = p.pbcontrib;
  for (int64_t g0 = base + (int64_t)tid * 8; g0 < e1; g0 += (int64_t)kPbThreads * 8 * kPbUnroll) {
    pb_u16x8 e[kPbUnroll];
    pb_f4 w[kPbUnroll][2];
#pragma unroll
    for (int u = 0; u < kPbUnroll; ++u) {
      const int64_t g = g0 + (int64_t)u * kPbThreads * 8;
      const int64_t gs = g < e1 ? g : base;  // past the end: a group of this block, not written
      e[u] = __builtin_nontemporal_load(&ent[gs >> 3]);
      if constexpr (!kSigned) {
        w[u][0] = __builtin_nontemporal_load(reinterpret_cast<const pb_f4*>(&vals[gs]));
        w[u][1] = __builtin_nontemporal_load(reinterpret_cast<const pb_f4*>(&vals[gs + 4]));
      }
    }
#pragma unroll
    for (int u = 0; u < kPbUnroll; ++u) {
      const int64_t g = g0 + (int64_t)u * kPbThreads * 8;
      if (g >= e1) break;
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint32_t c = e[u][q];
        const float x = xs[c & (kPbCols - 1)];
        if constexpr (kSigned) v[q] = (c & 0x8000u) ? -x : x;
        else v[q] = w[u][q >> 2][q & 3] * x;
      }
      if (g >= e0 && g + 8 <= e1) {
        *reinterpret_cast<pb_f4*>(&out[g]) = pb_f4{v[0], v[1], v[2], v[3]};
        *reinterpret_cast<pb_f4*>(&out[g + 4]) = pb_f4{v[4], v[5], v[6], v[7]};
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (g + q >= e0 && g + q < e1) out[g + q] = v[q];
      }
    }
  }
}

// Pass two: workgroup R; wave w takes the tiles of column blocks w, w + 16, ..., 64 at a time.
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
  // The tiles are short (a few hundred entries), so a wave packs 64 of them into one index space:
  // lane l holds tile C = wave + 16 (k + l), its start and the exclusive prefix of the tile sizes, and
  // slot q of the batch belongs to the last lane whose prefix is <= q (found by a binary search over
  // the lanes with shuffles).
  constexpr int kWaves = kPbThreads / 64;
  for (int k = 0; k * kWaves + wave < p.pbnc; k += 64) {
    const int C = wave + kWaves * (k + lane);
    int64_t a = 0;
    int cnt = 0;
    if (C < p.pbnc) {
      const int64_t t = (int64_t)C * p.bnr + R;
      a = p.pbtoff[t];
      cnt = (int)(p.pbtoff[t + 1] - a);
    }
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    const int excl = incl - cnt;
    const int total = __shfl(incl, 63, 64);
    for (int q0 = 0; q0 < total; q0 += 64 * kPbUnroll2) {
      float v[kPbUnroll2];
      uint32_t r[kPbUnroll2];
#pragma unroll
      for (int u = 0; u < kPbUnroll2; ++u) {
        const int q = q0 + 64 * u + lane;
        int lo = 0;  // the last lane with excl <= q (q < total: a nonempty tile)
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
          const int e = __shfl(excl, lo + step, 64);
          if (e <= q) lo += step;
        }
        const int64_t base = __shfl(a, lo, 64);
        const int off = q - __shfl(excl, lo, 64);
        const int64_t i = q < total ? base + off : 0;
        v[u] = __builtin_nontemporal_load(&prod[i]);
        r[u] = __builtin_nontemporal_load(&prow[i]);
      }
#pragma unroll
      for (int u = 0; u < kPbUnroll2; ++u) {
        // runs of one row (a tile's entries run by row; a run may continue into the next tile, whose
        // local rows are the same rows) are summed onto their first lane in FP64, which adds once
        const bool ok = q0 + 64 * u + lane < total;
        const uint32_t rn = (uint32_t)__shfl_down((int)r[u], 1, 64);
        const uint64_t m = __ballot(ok && lane < 63 && q0 + 64 * u + lane + 1 < total && rn == r[u]);
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
  if (p.brows < 64 || p.brows > kPbRows || p.bnr < 1 || (!p.rbound && (int64_t)p.bnr * p.brows < p.n) ||
      p.pbnc < 1 || !p.pbound || (int64_t)p.pbnc * p.bnr >= (1ll << 31)) {
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

__global__ __launch_bounds__(SA_BLOCK) void pb_fill_kernel(int64_t nnz, const uint64_t* __restrict__ keys,
                                                           const int32_t* __restrict__ idx,
                                                           const float* __restrict__ vals, float mag,
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
    pbrow[i] = (uint16_t)((key >> 15) & 0x7FFFu);
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

int sa_pblock_fill_f32(const sa_csr* A, int32_t brows, int32_t bnr, const uint64_t* keys, const int32_t* idx,
                       uint16_t* pbent, uint16_t* pbrow, float* bvals, double mag, int32_t* flag, hipStream_t stream) {
  SA_REQUIRE(A && A->n > 0 && A->nnz >= 0 && A->vals && brows >= 64 && bnr >= 1, "bad arguments");
  SA_REQUIRE(keys && idx && pbent && pbrow && flag && (mag > 0.0 || bvals), "bad arguments");
  hipLaunchKernelGGL(sa::pb_fill_kernel, dim3(sa::grid_for(A->nnz + 8, SA_BLOCK, 8192)), dim3(SA_BLOCK), 0, stream,
                     A->nnz, keys, idx, (const float*)A->vals, (float)(mag > 0.0 ? mag : 0.0), pbent, pbrow,
                     mag > 0.0 ? (float*)nullptr : bvals, flag);
  return sa::check_launch("pblock_fill");
}

}  // extern "C"
