// Fused scaled CSR SpMV: y_r = a * x_r + b * s_r * sum_k J_rk g_k, where g = s (.) x is gathered.
#pragma once

#include "sa_common.hpp"

namespace sa {

template <typename T>
struct SpmvParams {
  int64_t n;
  const int32_t* indptr;
  const int32_t* indices;
  const T* vals;
  const int32_t* tile_rows;
  int64_t ntiles;
  int32_t lanes_fixed;
  const T* g;                 // gathered vector (s (.) x)
  const T* s;                 // row scale of the epilogue; NULL means 1
  const T* x;                 // diagonal-term vector; NULL when a == 0
  double a, b;
  // NULL, or the Lanczos state: the launch then runs only while status == SA_RUNNING and b is
  // multiplied by sigma[j], the scale of the raw basis vector being applied.
  const sa_state* lz;
  // NULL, or a status word: the launch then runs only while *gate == SA_RUNNING (set to &lz->status
  // for the Lanczos step; the MINRES solve points it at its own status).
  const int32_t* gate;
  // NULL, or the instance of every tile of a block-diagonal batch (tiles do not cross instances):
  // a tile then runs only while its instance's status, the int32 at gate_base + inst * gate_stride,
  // is SA_RUNNING.
  const int32_t* tile_inst;
  const char* gate_base;
  int64_t gate_stride;
  // NULL, or the plain-layout tiles to run (ntile_list of them, each in [0, ntiles)): workgroup w runs
  // tile tile_list[w], so tiles left out cost no launch.
  const int32_t* tile_list;
  int64_t ntile_list;
  T* y;
  // XCD panel layout (npanels > 0): partial row sums per panel go to partial[p * n + r].
  int32_t npanels;
  const int32_t* pindptr;
  const int32_t* pindices;
  const T* pvals;
  const int32_t* ptile_rows;
  const int64_t* ptile_off;
  int64_t pmax_tiles;
  T* partial;
  int64_t pld;
  int32_t psigned;   // panel entries carry their sign in bit 31 of the index; magnitude pmag
  T pmag;
  // LDS-blocked layout (bmode = 1): tiles of brows x bcols per panel, entries in 64-entry chunks.
  int32_t bmode, brows, bcols, bnc, bnr;
  const int32_t* pbound;
  const int32_t* bchunk;
  const uint32_t* bent;
  const void* bvals;
  const int32_t* rbound;  // NULL (row block rb = rows [rb * brows, ...)) or bnr + 1 boundaries
  // Instance-staged layout of a block-diagonal batch (imode = 1; see sa_inst.hip).
  int32_t imode, irows, ilanes, ipasses, isigned;
  int64_t inwg;
  const int32_t* iinst;
  const int32_t* irow;
  const int64_t* ibase;
  const int64_t* ireal;
  const int64_t* iend;
  const int32_t* isplit;
  const int32_t* ient;
  const T* imag;
  // Column-streamed layout (bmode = 2): chunk base columns (bchunk, bent, bvals, pbound, rbound,
  // brows and bnr as above).
  const int32_t* bcbase;
  // Two-pass layout (bmode = 3): see sa_csr.
  const uint16_t* pbent;
  const uint16_t* pbrow;
  const int32_t* pbtoff;
  const int32_t* pbrtoff;
  float* pbcontrib;
  int32_t pbnc;
};

template <typename T>
SpmvParams<T> params_from(const sa_csr* A);

template <typename T>
int launch_spmv(const SpmvParams<T>& p, int32_t tile_nnz, hipStream_t stream);

template <typename T>
int launch_block_spmv(const SpmvParams<T>& p, hipStream_t stream);

template <typename T>
int launch_inst_spmv(const SpmvParams<T>& p, hipStream_t stream);

int launch_cstream_spmv(const SpmvParams<float>& p, hipStream_t stream);

int launch_rstream_spmv(const SpmvParams<float>& p, hipStream_t stream);

int launch_pblock_spmv(const SpmvParams<float>& p, hipStream_t stream);

int launch_istream_spmv(const SpmvParams<float>& p, hipStream_t stream);

// Layout builds: radix-sort (key, entry index) pairs (on the caller's scratch) and set start[q] to the
// first sorted key whose segment (key >> shift) is q, for q in [0, nseg]; bits needed to count to count.
int sort_entry_keys(const uint64_t* keys, const int32_t* idx, uint64_t* keys_out, int32_t* idx_out, int64_t nnz,
                    int end_bit, void* tmp, size_t tmp_bytes, int shift, int64_t nseg, int32_t* start,
                    hipStream_t stream);
int key_bits(int64_t count);

template <typename T>
int spmv_apply(const sa_csr* A, const T* s, const T* g, double a, const T* x, double b, T* y, const int32_t* gate,
               hipStream_t stream);

template <typename T>
int launch_prescale(int64_t n, const T* s, const T* x, T* z, hipStream_t stream);

}  // namespace sa
