// Instance-staged SpMV for a block-diagonal batch: the gathers of the fused scaled SpMV served from LDS.
//
// Instance i of a packed batch couples only its own rows and columns [ibase[i], ireal[i]) (the batch
// sizes here are at most 2^16 rows). Workgroup w owns a run of irows rows of one instance; it copies
// the instance's slice of the gathered vector into LDS (kC entries at a time, one pass per kC columns)
// with 16-byte loads, then each group of `lanes` lanes sums one row from LDS, reading the row's entries
// of the current pass, [isplit[c - 1][r], isplit[c][r]), straight from the CSR. So every gather is an
// LDS read instead of an L2 request, and the staging traffic is the instance's slice once per
// workgroup (a few percent of the entries for irows in the thousands). A row's sum over the passes
// stays with the lane that owns it (in LDS, in FP64), and the fused epilogue y = a x + b s (.) sum is
// the plain CSR kernel's, so padding rows (no entries) get y = a x.
#include "sa_spmv.hpp"

namespace sa {

constexpr int kIThreads = 1024;          // one workgroup per CU: the staged slice takes most of LDS
constexpr int kIMaxRows = 2048;          // rows per workgroup (the FP64 row sums: 16 KB of LDS)
constexpr int kIMaxPasses = 4;
constexpr int kIUnroll = 4;              // entry loads per row in flight per lane
constexpr int kIRows = 4;                // rows per lane group in flight

template <typename T>
struct InstCfg {
  static constexpr int kC = 131072 / sizeof(T);   // staged columns per pass: 128 KB of LDS
};

// kIUnroll entries of a row segment [e0, e0 + len) from offset o (stride L): raw index words and values;
// positions past the segment load nothing and contribute 0.
template <typename T, bool kSigned, int L>
__device__ inline void inst_loads(const SpmvParams<T>& p, int32_t e0, int32_t len, int32_t o, int32_t* raw, T* v) {
#pragma unroll
  for (int u = 0; u < kIUnroll; ++u) {
    const int32_t k = o + u * L;
    raw[u] = -1;
    v[u] = (T)0;
    if (k < len) {
      if constexpr (kSigned) {
        raw[u] = stream_load(&p.ient[e0 + k]);
      } else {
        raw[u] = stream_load(&p.indices[e0 + k]);
        v[u] = stream_load(&p.vals[e0 + k]);
      }
    }
  }
}

template <typename T, bool kSigned, int L>
__device__ inline double inst_products(const int32_t* raw, const T* v, T mag, int32_t len, int32_t o,
                                       const T* xb, int64_t lo) {
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < kIUnroll; ++u) {
    if (o + u * L < len) {
      T w;
      int32_t col;
      if constexpr (kSigned) {
        col = raw[u] & 0x7FFFFFFF;
        w = raw[u] < 0 ? -mag : mag;
      } else {
        col = raw[u];
        w = v[u];
      }
      acc += (double)w * (double)xb[col - lo];
    }
  }
  return acc;
}

template <typename T, bool kSigned, int L>
__global__ __launch_bounds__(kIThreads) void spmv_inst_kernel(SpmvParams<T> p) {
  typedef typename vec16<T>::type V;
  constexpr int kC = InstCfg<T>::kC;
  constexpr int VL = vec16<T>::L;
  __shared__ __attribute__((aligned(16))) T xb[kC];
  __shared__ double ysum[kIMaxRows];
  __shared__ int32_t rb0[kIMaxRows], rb1[kIMaxRows];  // this pass's entry range of every row
  static_assert(sizeof(T) * kC + (sizeof(double) + 2 * sizeof(int32_t)) * kIMaxRows <= 163840, "LDS");
  if (p.gate && *p.gate != SA_RUNNING) return;
  double bb = p.b;
  if (p.lz) bb *= p.lz->sigma[p.lz->j];
  const int64_t w = p.tile_list ? (int64_t)p.tile_list[blockIdx.x] : (int64_t)blockIdx.x;
  const int inst = p.iinst[w];
  if (p.gate_base && *reinterpret_cast<const int32_t*>(p.gate_base + (int64_t)inst * p.gate_stride) != SA_RUNNING)
    return;
  const int64_t b0 = p.ibase[inst], real = p.ireal[inst], end = p.iend[inst];
  const int64_t r0 = p.irow[w];
  const int rows = (int)((r0 + p.irows < end ? r0 + p.irows : end) - r0);
  const int64_t width = real - b0;
  const int passes = width > 0 ? (int)((width + kC - 1) / kC) : 1;
  const T mag = kSigned ? p.imag[inst] : (T)0;
  const int tid = threadIdx.x;
  constexpr int G = kIThreads / L;  // row groups
  const int gid = tid / L, lig = tid % L;
  for (int c = 0; c < passes; ++c) {
    const int64_t lo = b0 + (int64_t)c * kC;
    const int64_t rest = width - (int64_t)c * kC;
    const int cnt = rest <= 0 ? 0 : (int)(rest < kC ? rest : kC);
    const bool first = c == 0, last = c == passes - 1;
    if (c > 0) __syncthreads();  // every row of the previous pass has read xb, rb0 and rb1
    // lo is a multiple of 64 entries, so the 16-byte loads are aligned; the ragged end goes by element
    for (int i = tid * VL; i < cnt; i += kIThreads * VL) {
      if (i + VL <= cnt) {
        *reinterpret_cast<V*>(xb + i) = *reinterpret_cast<const V*>(p.g + lo + i);
      } else {
        for (int q = i; q < cnt; ++q) xb[q] = p.g[lo + q];
      }
    }
    for (int i = tid; i < rows; i += kIThreads) {
      const int64_t r = r0 + i;
      rb0[i] = first ? p.indptr[r] : p.isplit[(int64_t)(c - 1) * p.n + r];
      rb1[i] = last ? p.indptr[r + 1] : p.isplit[(int64_t)c * p.n + r];
    }
    __syncthreads();
    // kIRows rows per group at a time, all their entry loads issued before the first gather
    for (int k = gid; k < rows; k += kIRows * G) {
      int32_t ea[kIRows], la[kIRows];
      int32_t len = 0;
#pragma unroll
      for (int h = 0; h < kIRows; ++h) {
        const int kk = k + h * G;
        ea[h] = kk < rows ? rb0[kk] : 0;
        la[h] = kk < rows ? rb1[kk] - ea[h] : 0;
        len = la[h] > len ? la[h] : len;
      }
      double acc[kIRows];
#pragma unroll
      for (int h = 0; h < kIRows; ++h) acc[h] = 0.0;
      for (int32_t o = lig; o < len; o += L * kIUnroll) {
        int32_t raw[kIRows][kIUnroll];
        T v[kIRows][kIUnroll];
#pragma unroll
        for (int h = 0; h < kIRows; ++h) inst_loads<T, kSigned, L>(p, ea[h], la[h], o, raw[h], v[h]);
#pragma unroll
        for (int h = 0; h < kIRows; ++h) acc[h] += inst_products<T, kSigned, L>(raw[h], v[h], mag, la[h], o, xb, lo);
      }
#pragma unroll
      for (int h = 0; h < kIRows; ++h) {
#pragma unroll
        for (int o = L >> 1; o > 0; o >>= 1) acc[h] += __shfl_xor(acc[h], o, 64);
      }
      if (lig == 0) {
#pragma unroll
        for (int h = 0; h < kIRows; ++h) {
          const int kk = k + h * G;
          if (kk >= rows) break;
          double a = acc[h];
          if (!first) a += ysum[kk];
          if (!last) {
            ysum[kk] = a;
          } else {
            const int64_t r = r0 + kk;
            const double sr = p.s ? (double)p.s[r] : 1.0;
            double y = bb * sr * a;
            if (p.x) y += p.a * (double)p.x[r];
            p.y[r] = (T)y;
          }
        }
      }
    }
  }
}

template <typename T>
int launch_inst_spmv(const SpmvParams<T>& p, hipStream_t stream) {
  if (!p.iinst || !p.irow || !p.ibase || !p.ireal || !p.iend || !p.indptr || p.inwg < 1 ||
      (p.isigned ? (!p.ient || !p.imag) : (!p.indices || !p.vals))) {
    set_error("spmv: incomplete instance-staged layout");
    return SA_E_ARG;
  }
  if (p.irows < 64 || p.irows > kIMaxRows || p.ipasses < 1 || p.ipasses > kIMaxPasses ||
      (p.ipasses > 1 && !p.isplit)) {
    set_error("spmv: instance-staged layout sizes out of range");
    return SA_E_ARG;
  }
  if (!aligned16(p.g)) {
    set_error("spmv: the instance-staged layout needs a 16-byte aligned gathered vector");
    return SA_E_ARG;
  }
  if (p.tile_list && (p.ntile_list < 0 || p.ntile_list > p.inwg)) {
    set_error("spmv: tile list longer than the layout's workgroups");
    return SA_E_ARG;
  }
  const int64_t nwg = p.tile_list ? p.ntile_list : p.inwg;
  if (nwg == 0) return SA_OK;
  const dim3 grid((unsigned)nwg), block(kIThreads);
#define SA_INST_CASE(LANES)                                                                          \
  case LANES:                                                                                        \
    if (p.isigned) hipLaunchKernelGGL((spmv_inst_kernel<T, true, LANES>), grid, block, 0, stream, p); \
    else hipLaunchKernelGGL((spmv_inst_kernel<T, false, LANES>), grid, block, 0, stream, p);          \
    break;
  switch (p.ilanes) {
    SA_INST_CASE(4)
    SA_INST_CASE(8)
    SA_INST_CASE(16)
    SA_INST_CASE(32)
    default: set_error("spmv: instance-staged lanes per row must be 4, 8, 16 or 32"); return SA_E_ARG;
  }
#undef SA_INST_CASE
  return check_launch("spmv_inst");
}

template int launch_inst_spmv<float>(const SpmvParams<float>&, hipStream_t);
template int launch_inst_spmv<double>(const SpmvParams<double>&, hipStream_t);

}  // namespace sa

extern "C" {

int sa_inst_columns_per_pass(int32_t elem_bytes) {
  if (elem_bytes == 4) return sa::InstCfg<float>::kC;
  if (elem_bytes == 8) return sa::InstCfg<double>::kC;
  return -1;
}

}  // extern "C"
