// The per-instance kernels of a batched sweep (instances packed block-diagonally, sa_batch): the
// alpha scaling with each instance's reductions, the rounding with each instance's sign, norm and
// signature, exact energies per (instance, candidate), and the assembly of the packed candidate
// columns of the local search. The batched Lanczos launches are in sa_lanczos.hip.
#include <cstddef>

#include "sa_reduce.hpp"

namespace sa {

// s = max(d^alpha, clamp)^(-1/2); per instance: sum of d^alpha (column 0) and max of d^(1-alpha)
// (column 1) over its real rows.
// alpha_each (NULL, or B values): instance b's own alpha, for a batch whose instances sit at
// different points of the sweep.
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void batch_alpha_scaling_kernel(BatchMap map, const T* __restrict__ d,
                                                                       T alpha, T one_minus_alpha,
                                                                       const double* __restrict__ alpha_each,
                                                                       T clamp, T* __restrict__ s) {
  __shared__ double red[SA_WAVES_PER_BLOCK];
  const Slice sl = map.slice();
  if (!sl.live) return;
  if (alpha_each) {
    const double a = alpha_each[sl.inst];
    alpha = (T)a;
    one_minus_alpha = (T)(1.0 - a);
  }
  double sum = 0.0, mx = -INFINITY;
  for (int64_t r = sl.lo + threadIdx.x; r < sl.hi; r += blockDim.x) {
    const T dr = d[r];
    const T dv = pow(dr, alpha);
    const T cl = dv > clamp ? dv : clamp;
    s[r] = (T)1 / sqrt(cl);
    sum += (double)dv;
    mx = fmax(mx, (double)pow(dr, one_minus_alpha));
  }
  double* part = partials_of(sl.work);
  const double bs = block_sum(sum, red);
  double bm = wave_max(mx);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = bm;
  __syncthreads();
  if (threadIdx.x == 0) {
    bm = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    part[2 * (int64_t)sl.blk] = bs;
    part[2 * (int64_t)sl.blk + 1] = bm;
  }
}

// Sign rounding of the Ritz vectors y (one per instance) with y_last = scale_b y, scale_b =
// sign_b / ||y_b|| from the instance's state (r[4] = ||y_b||, r[5] = y_b . y_last_b; sign_b = -1
// when align and r[5] < 0). One wavefront per 64-row word; the words of an instance start at its
// first row, the bits of padding rows are 0, and the signature hashes instance-local word indices
// (the same value the one-problem rounding gives for the instance alone).
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void batch_round_kernel(BatchMap map, const T* __restrict__ y,
                                                               T* __restrict__ y_last, uint64_t* __restrict__ bits,
                                                               int32_t align) {
  const Slice sl = map.slice();
  if (!sl.live) return;
  const sa_state* st = state_of(sl.work);
  const double nrm = st->r[4];
  const double sign = (align && st->r[5] < 0.0) ? -1.0 : 1.0;
  const double scale = nrm > 0.0 ? sign / nrm : 0.0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t w0 = sl.lo >> 6, w1 = (sl.hi + 63) >> 6;
  uint64_t hsum = 0;
  for (int64_t w = w0 + wave; w < w1; w += SA_WAVES_PER_BLOCK) {
    const int64_t r = w * 64 + lane;
    bool up = false;
    if (r < sl.hi) {
      const T v = (T)(scale * (double)y[r]);
      if (y_last) y_last[r] = v;
      up = v >= (T)0;
    }
    const uint64_t word = __ballot(up);
    if (lane == 0) {
      bits[w] = word;
      hsum += splitmix64(word ^ splitmix64((uint64_t)(w - (sl.base >> 6))));
    }
  }
  block_partials_u64(hsum, sl.work, sl.blk);
}

// dst = src on the real rows of the live instances.
template <typename T>
__global__ __launch_bounds__(SA_BLOCK) void batch_copy_kernel(BatchMap map, const T* __restrict__ src,
                                                              T* __restrict__ dst) {
  const Slice sl = map.slice();
  if (!sl.live) return;
  for (int64_t r = sl.lo + threadIdx.x; r < sl.hi; r += blockDim.x) dst[r] = src[r];
}

// e2[b][c] = sum_{i in b} s_i sum_j J_ij s_j for candidate bit vectors (exact, int64).
__global__ __launch_bounds__(SA_BLOCK) void batch_energy_bits_kernel(BatchMap map, const int32_t* __restrict__ indptr,
                                                                     const int32_t* __restrict__ indices,
                                                                     const int32_t* __restrict__ jint,
                                                                     const uint64_t* __restrict__ cand, int64_t nwords,
                                                                     int32_t k, int64_t* __restrict__ e2) {
  __shared__ int64_t red[SA_WAVES_PER_BLOCK];
  const Slice sl = map.slice();
  const int c = blockIdx.y;
  const uint64_t* __restrict__ b = cand + (int64_t)c * nwords;
  int64_t acc = 0;
  for (int64_t i = sl.lo + threadIdx.x; i < sl.hi; i += blockDim.x) {
    int64_t h = 0;
    for (int32_t e = indptr[i]; e < indptr[i + 1]; ++e) {
      const int32_t j = stream_load(&indices[e]);
      const int32_t sj = ((b[j >> 6] >> (j & 63)) & 1ull) ? 1 : -1;
      h += (int64_t)stream_load(&jint[e]) * sj;
    }
    const int64_t si = ((b[i >> 6] >> (i & 63)) & 1ull) ? 1 : -1;
    acc += si * h;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0 && acc != 0)
    atomicAdd(reinterpret_cast<unsigned long long*>(&e2[(int64_t)sl.inst * k + c]), (unsigned long long)acc);
}

// e2[b][c] = sum_{i in b} s_ic h_ic from the local search's node-major spins and exact fields.
__global__ __launch_bounds__(SA_BLOCK) void batch_ls_energy_kernel(BatchMap map, int32_t k, int32_t groups,
                                                                   const int8_t* __restrict__ spins,
                                                                   const int32_t* __restrict__ fields,
                                                                   int64_t* __restrict__ e2) {
  __shared__ int64_t red[SA_WAVES_PER_BLOCK][64];
  const Slice sl = map.slice();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.y;
  const int c = g * 64 + lane;
  const int64_t kp = (int64_t)groups * 64;
  int64_t acc = 0;
  for (int64_t i = sl.lo + wave; i < sl.hi; i += SA_WAVES_PER_BLOCK)
    acc += (int64_t)spins[i * kp + c] * (int64_t)fields[i * kp + c];
  red[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && c < k) {
    const int64_t v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    if (v != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&e2[(int64_t)sl.inst * k + c]), (unsigned long long)v);
  }
}

// out[c][w] = in[src[c][b(w)]][w]: packed candidate column c takes, for every instance b, the spin
// words of the rounded vector src[c][b] (words [word_off[b], word_off[b+1]) belong to instance b).
__global__ __launch_bounds__(SA_BLOCK) void bits_gather_kernel(int64_t nwords, int32_t k, int32_t B,
                                                               const int64_t* __restrict__ word_off,
                                                               const int32_t* __restrict__ src,
                                                               const uint64_t* __restrict__ in,
                                                               uint64_t* __restrict__ out) {
  const int c = blockIdx.y;
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < nwords; w += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = B - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (word_off[mid] <= w) lo = mid; else hi = mid - 1;
    }
    const int64_t row = src[(int64_t)c * B + lo];
    out[(int64_t)c * nwords + w] = in[row * nwords + w];
  }
}

template <typename T>
static int batch_alpha_scaling(const sa_batch* Bt, const T* d, double alpha, const double* alpha_each, double clamp,
                               T* s, hipStream_t stream) {
  SA_REQUIRE(batch_ok(Bt) && d && s, "bad arguments");
  hipLaunchKernelGGL(batch_alpha_scaling_kernel<T>, dim3((unsigned)Bt->nblocks), dim3(SA_BLOCK), 0, stream, bmap(Bt),
                     d, (T)alpha, (T)(1.0 - alpha), alpha_each, (T)clamp, s);
  int rc = check_launch("batch_alpha_scaling");
  if (rc) return rc;
  FinishSpec f{};
  f.cols = 2;
  f.only_if = -1;
  f.op[0] = FIN_SUM;
  f.op[1] = FIN_MAX;
  f.dst[0] = (int32_t)offsetof(sa_state, r[1]);
  f.dst[1] = (int32_t)offsetof(sa_state, r[2]);
  return launch_finish_batch(Bt, f, stream);
}

template <typename T>
static int batch_round(const sa_batch* Bt, const T* y, T* y_last, uint64_t* bits, int32_t align,
                       hipStream_t stream) {
  SA_REQUIRE(batch_ok(Bt) && y && bits, "bad arguments");
  hipLaunchKernelGGL(batch_round_kernel<T>, dim3((unsigned)Bt->nblocks), dim3(SA_BLOCK), 0, stream, bmap(Bt), y,
                     y_last, bits, align);
  int rc = check_launch("batch_round");
  if (rc) return rc;
  FinishSpec f{};
  f.cols = 1;
  f.only_if = -1;
  f.op[0] = FIN_SUM_U64;
  f.dst[0] = (int32_t)offsetof(sa_state, hash);
  return launch_finish_batch(Bt, f, stream);
}

template <typename T>
static int batch_copy(const sa_batch* Bt, const T* src, T* dst, hipStream_t stream) {
  SA_REQUIRE(batch_ok(Bt) && src && dst, "bad arguments");
  hipLaunchKernelGGL(batch_copy_kernel<T>, dim3((unsigned)Bt->nblocks), dim3(SA_BLOCK), 0, stream, bmap(Bt), src, dst);
  return check_launch("batch_copy");
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_batch_alpha_scaling_f32(const sa_batch* Bt, const float* d, double alpha, double clamp, float* s,
                               hipStream_t stream) {
  return batch_alpha_scaling<float>(Bt, d, alpha, nullptr, clamp, s, stream);
}

int sa_batch_alpha_scaling_f64(const sa_batch* Bt, const double* d, double alpha, double clamp, double* s,
                               hipStream_t stream) {
  return batch_alpha_scaling<double>(Bt, d, alpha, nullptr, clamp, s, stream);
}

int sa_batch_alpha_scaling_each_f32(const sa_batch* Bt, const float* d, const double* alpha, double clamp, float* s,
                                    hipStream_t stream) {
  SA_REQUIRE(alpha, "bad arguments");
  return batch_alpha_scaling<float>(Bt, d, 0.0, alpha, clamp, s, stream);
}

int sa_batch_alpha_scaling_each_f64(const sa_batch* Bt, const double* d, const double* alpha, double clamp,
                                    double* s, hipStream_t stream) {
  SA_REQUIRE(alpha, "bad arguments");
  return batch_alpha_scaling<double>(Bt, d, 0.0, alpha, clamp, s, stream);
}

int sa_batch_copy_f32(const sa_batch* Bt, const float* src, float* dst, hipStream_t stream) {
  return batch_copy<float>(Bt, src, dst, stream);
}

int sa_batch_copy_f64(const sa_batch* Bt, const double* src, double* dst, hipStream_t stream) {
  return batch_copy<double>(Bt, src, dst, stream);
}

int sa_batch_round_f32(const sa_batch* Bt, const float* y, float* y_last, uint64_t* bits, int32_t align,
                       hipStream_t stream) {
  return batch_round<float>(Bt, y, y_last, bits, align, stream);
}

int sa_batch_round_f64(const sa_batch* Bt, const double* y, double* y_last, uint64_t* bits, int32_t align,
                       hipStream_t stream) {
  return batch_round<double>(Bt, y, y_last, bits, align, stream);
}

int sa_batch_energy_bits_i64(const sa_csr* A, const sa_batch* Bt, int32_t k, const uint64_t* cand_bits,
                             int64_t nwords, int64_t* e2, hipStream_t stream) {
  SA_REQUIRE(A && A->jint && batch_ok(Bt) && cand_bits && e2 && k >= 1 && k <= 65535, "bad arguments");
  SA_REQUIRE(!Bt->inst_mask, "the batched energies take every instance (no mask)");
  SA_REQUIRE(nwords * 64 >= A->n, "nwords too small for n");
  hipLaunchKernelGGL(batch_energy_bits_kernel, dim3((unsigned)Bt->nblocks, k), dim3(SA_BLOCK), 0, stream, bmap(Bt),
                     A->indptr, A->indices, A->jint, cand_bits, nwords, k, e2);
  return check_launch("batch_energy_bits_i64");
}

int sa_batch_ls_energy(const sa_batch* Bt, int32_t k, int32_t groups, const int8_t* spins, const int32_t* fields,
                       int64_t* e2, hipStream_t stream) {
  SA_REQUIRE(batch_ok(Bt) && spins && fields && e2 && k >= 1 && groups >= 1 && k <= 64 * groups, "bad arguments");
  SA_REQUIRE(!Bt->inst_mask, "the batched energies take every instance (no mask)");
  hipLaunchKernelGGL(batch_ls_energy_kernel, dim3((unsigned)Bt->nblocks, groups), dim3(SA_BLOCK), 0, stream, bmap(Bt),
                     k, groups, spins, fields, e2);
  return check_launch("batch_ls_energy");
}

int sa_bits_gather(int64_t nwords, int32_t k, int32_t B, const int64_t* word_off, const int32_t* src,
                   const uint64_t* in, uint64_t* out, hipStream_t stream) {
  SA_REQUIRE(nwords > 0 && k >= 1 && k <= 65535 && B >= 1 && word_off && src && in && out, "bad arguments");
  hipLaunchKernelGGL(bits_gather_kernel, dim3(grid_for(nwords, SA_BLOCK, 1024), k), dim3(SA_BLOCK), 0, stream, nwords,
                     k, B, word_off, src, in, out);
  return check_launch("bits_gather");
}

}  // extern "C"
