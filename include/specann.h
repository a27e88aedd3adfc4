/*
 * libspecann: C ABI of the MI355X (gfx950) SpecAnn alpha-sweep hot path.
 *
 * Every function takes plain device pointers and sizes plus the HIP stream to launch on, returns 0
 * or a negative SA_E_* status, never throws, and allocates no device memory: the caller owns every
 * buffer (the reference's kernels likewise run on caller-owned CuPy arrays). Functions are stream
 * ordered. Reductions carry their partial sums in the caller's work buffer (sa_work_bytes), whose
 * first bytes are the sa_state block; two calls that share one work buffer must not overlap.
 *
 * Precision suffixes: _f32 = FP32 couplings and vectors (the reference's precision 32 with
 * fp32_spmv), _f64 = FP64 everywhere (the reference's SpecAnn-64). Lanczos coefficients, the Ritz
 * solve, dots and norms are FP64 in both.
 *
 * Reference citations are relative to the Vgueca/SpecAnn source tree.
 */
#ifndef SPECANN_H
#define SPECANN_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SA_OK 0
#define SA_E_ARG (-1)
#define SA_E_LAUNCH (-2)
#define SA_E_BREAKDOWN (-3)
#define SA_E_OOM (-4)
#define SA_E_NONINTEGER (-5)

/* Lanczos status values (sa_state.status). */
#define SA_RUNNING 0
#define SA_CONVERGED 1
#define SA_MAXITER 2
#define SA_RESTART 3
#define SA_REJECTED 4
#define SA_BROKEN (-1)

#define SA_MAX_NCV 128
#define SA_RED_GRID 512        /* workgroups of the reduction kernels */
#define SA_RED_COLS 8          /* basis vectors per dot-product chunk */
#define SA_MAX_KEEP 16         /* Ritz vectors a thick restart may keep */

/* Symmetric CSR of the couplings, device pointers. nnz < 2^31 (int32 indices). */
typedef struct sa_csr {
  int64_t n;
  int64_t nnz;
  const int32_t* indptr;    /* n + 1 */
  const int32_t* indices;   /* nnz, sorted per row */
  const void* vals;         /* nnz normalised couplings J / c: float (_f32) or double (_f64) */
  const int32_t* jint;      /* nnz integer couplings J (local search, exact energies); may be NULL */
  const int32_t* tile_rows; /* ntiles + 1: first row whose first stored entry is >= t * tile_nnz */
  int64_t ntiles;
  int32_t tile_nnz;         /* stored entries per SpMV tile: 512, 1024, 2048 or 4096 */
  int32_t lanes_per_row;    /* 0: choose per tile from its mean row length; else 1..64 */
  /* Optional XCD-panel layout (npanels = 0: plain CSR). The same couplings regrouped by column
   * panel p = [pbound[p], pbound[p+1]); workgroups reach the 8 XCDs round-robin, and workgroup w
   * works on panel (w / (8 * pmax_tiles)) * 8 + w % 8, so each XCD gathers only from its panel's
   * slice of the vector, which stays in that XCD's L2. Per-panel row sums land in `partial` and
   * are combined (in a fixed order) by the consumer. */
  int32_t npanels;          /* 0, or a multiple of 8 up to 64 */
  int32_t pad0;
  const int32_t* pindptr;   /* npanels x (n + 1): offsets into pindices / pvals */
  const int32_t* pindices;  /* nnz */
  const void* pvals;        /* nnz, T */
  const int32_t* ptile_rows;/* per panel: ntiles_p + 1 tile rows, concatenated */
  const int64_t* ptile_off; /* npanels + 1 offsets into ptile_rows */
  int64_t pmax_tiles;       /* max over panels of ntiles_p */
  void* partial;            /* npanels rows of n entries of T, leading dimension pld */
  int64_t pld;
  /* psigned = 1: every coupling has the magnitude pmag (J in {-1,+1} or J = -1 ...): pindices then
   * hold column | (sign bit << 31) and pvals is not read (sa_panel_sign_fold). */
  int32_t psigned;
  int32_t pad2;
  double pmag;
  /* Optional LDS-blocked layout (bmode = 1; needs npanels = 8, pbound, partial, pld, and psigned /
   * pmag as above). Panel p's columns are cut into blocks of bcols, rows into bnr blocks of brows;
   * tile t = (p * bnr + rb) * bnc + cb holds the entries of row block rb in column block cb of
   * panel p, in 64-entry chunks [bchunk[t], bchunk[t+1]). Workgroup w = rb * 8 + p (XCD w % 8 = p)
   * stages each column block of the gathered vector in LDS and sums its rows in LDS. An entry is
   * sign << 31 | cont << 30 | local column << 15 | local row (cont: the next entry has the same row;
   * a row's entries in one tile never straddle a chunk); padding entries have local row = brows.
   * rbound (optional): bnr + 1 row-block boundaries, multiples of 64 except rbound[bnr] = n, each
   * block at most brows rows (row blocks of variable height, balanced by entries when a few rows hold
   * most of them); NULL: row block rb is rows [rb * brows, (rb + 1) * brows). */
  int32_t bmode;
  int32_t brows;            /* rows per row block, a multiple of 64, < 32768 */
  int32_t bcols;            /* columns per column block, a multiple of 64, <= 65536 / sizeof(T); the
                             * row sums (brows + 1, 16-byte aligned) and the block take <= 162816 bytes */
  int32_t bnc;              /* column blocks per panel (the widest panel's count), at most 368 */
  int32_t bnr;              /* row blocks (bnr * brows >= n unless rbound is given) */
  int32_t pad3;
  const int32_t* pbound;    /* npanels + 1 column boundaries, multiples of 64 except pbound[npanels] = n */
  const int32_t* bchunk;    /* 8 * bnr * bnc + 1 tile starts, in chunks of 64 entries */
  const uint32_t* bent;     /* 64 * bchunk[8 * bnr * bnc] entries */
  const void* bvals;        /* entry values (T; 0 for padding), or NULL when psigned */
  const int32_t* rbound;    /* NULL, or bnr + 1 row-block boundaries (see above) */
  /* Optional instance-staged layout of a block-diagonal batch (imode = 1, npanels = 0; no counterpart
   * in the reference). Instance i couples only rows and columns [ibase[i], ireal[i]) (rows up to
   * iend[i] are padding without entries). Workgroup w owns rows [irow[w], min(irow[w] + irows,
   * iend[iinst[w]])) of instance iinst[w]; it stages the instance's slice of the gathered vector in
   * LDS, sa_inst_columns_per_pass(sizeof(T)) columns per pass (ipasses at most 4 over the batch),
   * and sums each row from LDS with ilanes lanes (4, 8, 16 or 32), reading the row's entries of pass
   * c in [isplit[(c-1) n + r], isplit[c n + r]) of the CSR (indptr bounds the first and last pass).
   * isigned = 1: every coupling of instance i has the magnitude imag[i] (T) and ient holds
   * column | sign << 31 for each entry; else indices and vals are read. */
  /* imode = 2 (FP32, no counterpart in the reference either): the instance-streamed layout, the
   * column-streamed chunks (bchunk, bent, bvals, bcbase) built with one row block per workgroup w,
   * rows [irow[w], min(irow[w] + irows, iend[iinst[w]])) with irows up to sa_cstream_max_rows(), and
   * every column in panel 0: workgroup w's chunks are [bchunk[8 w], bchunk[8 w + 1]). isigned = 1:
   * every coupling of instance i has the magnitude imag[i] and the entries carry the sign (bit 31),
   * bvals unused; else bvals holds the couplings. */
  int32_t imode;
  int32_t irows;            /* rows per workgroup, 64..2048 (imode 1) */
  int32_t ilanes;
  int32_t ipasses;
  int32_t isigned;
  int32_t pad4;
  int64_t inwg;             /* workgroups */
  const int32_t* iinst;     /* inwg */
  const int32_t* irow;      /* inwg */
  const int64_t* ibase;     /* B */
  const int64_t* ireal;     /* B */
  const int64_t* iend;      /* B */
  const int32_t* isplit;    /* (ipasses - 1) * n, or NULL when ipasses = 1 */
  const int32_t* ient;      /* nnz when isigned */
  const void* imag;         /* B values of T when isigned */
  /* Optional column-streamed layout (bmode = 2, FP32 only; needs npanels = 8, pbound, partial, pld,
   * and psigned / pmag as above; no counterpart in the reference). Workgroup w = rb * 8 + p sums the
   * entries of row block rb (brows rows, or [rbound[rb], rbound[rb + 1])) in panel p, sorted by
   * column then row, from the 256-entry chunks [bchunk[w], bchunk[w + 1]); chunk k's entries are
   * stored lane-major (sorted entry 64 j + l at bent[256 k + 4 l + j]) and its columns are
   * pbound[p] + bcbase[k] + delta. An entry is sign << 31 | local row << 16 | delta (16 bits);
   * padding entries have local row = sa_cstream_pad_row() and delta 0. bvals (float, same order)
   * holds the couplings unless psigned. bcols and bnc are not used.
   * Row-owned form (bmode = 4, FP32 only, npanels = 0, no partial): the same arrays built with one
   * panel (pbound = {0, n, n, n, n, n, n, n, n}); workgroup rb sums every entry of row block rb from
   * the chunks [bchunk[8 rb], bchunk[8 rb + 1]) (the other seven segments are empty) and writes its
   * rows of y directly. */
  const int32_t* bcbase;    /* bchunk[8 * bnr] chunk base columns, relative to the panel */
  /* Optional two-pass layout (bmode = 3, FP32 only, npanels = 0; no counterpart in the reference),
   * for vectors far larger than an XCD's L2. Columns are cut into pbnc blocks [pbound[c],
   * pbound[c + 1]) of at most sa_pblock_cols() columns, rows into bnr blocks of brows rows (or
   * rbound). The entries are stored
   * sorted by tile t = column block * bnr + row block, then row, then column; tile t holds
   * [pbtoff[t], pbtoff[t + 1]). Pass one: workgroup C stages its column block of g in LDS and writes
   * the product of every entry of column block C to pbcontrib, where the products are ordered by row
   * block, then column block (tile (c, r) at pbrtoff[r * pbnc + c], in its stored order). Pass two:
   * workgroup R adds its contiguous products into FP64 row sums in LDS and applies the epilogue.
   * pbent: local column | sign << 15 (uint16, stored order, 8 padding entries at the end); pbrow:
   * local row (uint16, products order); bvals (float, stored order) holds the couplings unless
   * psigned; bnr is at most 3840. */
  const uint16_t* pbent;    /* nnz + 8 */
  const uint16_t* pbrow;    /* nnz */
  const int32_t* pbtoff;    /* pbnc * bnr + 1 tile offsets (stored order) */
  const int32_t* pbrtoff;   /* bnr * pbnc + 1: tile (c, r) starts at pbrtoff[r * pbnc + c] in pbcontrib */
  void* pbcontrib;          /* nnz + 8 floats: the products between the two passes, row-block order */
  int32_t pbnc;             /* column blocks */
  int32_t pad5;
} sa_csr;

/* A batch of independent instances packed block-diagonally into one CSR (no counterpart in the
 * reference, which solves one instance per call). Instance b owns the packed rows
 * [row_off[b], row_off[b+1]) (multiples of 64 apart; padding rows have no entries) and the
 * workgroups [blk_off[b], blk_off[b+1]) of the batched vector kernels, rows_per_block rows each; its
 * solve state lives in the work buffer works + b * stride (sa_work_bytes bytes each). */
typedef struct sa_batch {
  int32_t B;
  int32_t rows_per_block;   /* a multiple of 64 */
  const int64_t* row_off;   /* B + 1, device */
  const int64_t* row_end;   /* B, device: end of instance b's real rows (then padding up to row_off[b+1]) */
  const int32_t* blk_off;   /* B + 1, device: blk_off[b+1] - blk_off[b] = ceil(real rows / rows_per_block) */
  void* works;
  int64_t stride;
  int64_t nblocks;          /* blk_off[B] */
  /* NULL, or the instance of every SpMV tile of the packed CSR, whose tiles then must not cross
   * instances: the batched step skips the tiles of instances that are not running. */
  const int32_t* tile_inst;
  /* NULL, or B bytes on the device: the per-instance entry points (alpha scaling, begin, start, Ritz
   * vector, rounding, row copy) then act only on the instances whose byte is nonzero and leave the
   * others' rows and state untouched; the step, the restart and the energy entry points refuse a
   * mask (they select instances by status). */
  const uint8_t* inst_mask;
  /* NULL, or ntile_list SpMV tiles (device, each in [0, A->ntiles)): the batched step's SpMV then
   * launches only these, e.g. the tiles of the instances still sweeping; a tile left out is skipped
   * without a workgroup. */
  const int32_t* tile_list;
  int64_t ntile_list;
} sa_batch;

/* Device-side state of one eigensolve; lives at the start of the work buffer. */
typedef struct sa_state {
  int32_t status;      /* SA_RUNNING, ... */
  int32_t j;           /* Lanczos steps done in the current cycle */
  int32_t steps;       /* operator applications since sa_lanczos_begin */
  int32_t cycles;      /* restarts */
  int32_t ncv;
  int32_t maxiter;
  int32_t gate;        /* 1 while step 0 runs on a predicted start vector */
  int32_t kept;        /* Ritz vectors kept by the last thick restart (rows 0..kept-1 of the basis) */
  int32_t keep;        /* Ritz vectors a thick restart keeps (1: restart from the Ritz vector alone) */
  int32_t orth_lo;     /* first basis row the current step's Gram-Schmidt update subtracts (0: all) */
  double theta;        /* largest Ritz value of N (the reference's mu = rho - theta) */
  double res;          /* residual estimate |beta_j * s_j| */
  double first_beta;   /* beta_0 of a gated start: the predictor residual */
  double tol;
  double bound;        /* breakdown bound on |alpha|, |beta| (10 x (rho + g) is applied) */
  double pred_tol;
  double orth_tau;     /* selective reorthogonalisation threshold (0: every step against the whole basis) */
  double r[8];         /* results of scalar reductions */
  uint64_t hash;       /* spin signature of the last rounding */
  int64_t i64[4];
  double a[SA_MAX_NCV];
  double b[SA_MAX_NCV];
  double sigma[SA_MAX_NCV + 1];
  double h[SA_MAX_NCV + 1];
  double ritz[SA_MAX_NCV];
  double arrow[SA_MAX_NCV]; /* couplings of the kept Ritz vectors to basis row `kept` */
  unsigned int counters[16]; /* [0]: steps orthogonalised against the whole basis, [1]: against the
                              * last two rows only (both since sa_lanczos_begin) */
} sa_state;

int sa_version(void);
/* Copies the last error message of this thread; returns its length. */
int sa_last_error(char* buf, size_t len);
/* Bytes of the work buffer for a problem of n rows (zero-fill it once before first use). */
size_t sa_work_bytes(int64_t n);
/* sizeof(sa_state), for bindings that mirror the struct. */
size_t sa_state_bytes(void);
/* sizeof(sa_csr), for bindings that mirror the struct. */
size_t sa_csr_bytes(void);

/* ---- Set-up: values, degrees, normalisation (models/spectral_annealing.py:270-288) ---------- */
/* vals_out = (T) raw (the host's float64 data), jint = (int32) raw; *nonint_dev (device int32)
 * is set to 1 if some value is not an integer in [-2^30, 2^30]. */
int sa_prepare_values_f32(int64_t nnz, const double* raw, float* vals_out, int32_t* jint, int32_t* nonint_dev,
                          hipStream_t stream);
int sa_prepare_values_f64(int64_t nnz, const double* raw, double* vals_out, int32_t* jint, int32_t* nonint_dev,
                          hipStream_t stream);
/* deg_i = sum_j |vals_ij| (accumulated in FP64, rounded to T); state.r[0] = max_i deg_i. */
int sa_abs_degree_f32(const sa_csr* A, float* deg, void* work, hipStream_t stream);
int sa_abs_degree_f64(const sa_csr* A, double* deg, void* work, hipStream_t stream);
/* vals /= c in place, d = deg / c (models/spectral_annealing.py:282-288). */
int sa_normalize_f32(int64_t nnz, float* vals, int64_t n, const float* deg, float* d, double c, hipStream_t stream);
/* Block-diagonal batch: row r and its entries divided by c_row[r] (its instance's factor). */
int sa_normalize_rows_f32(const sa_csr* A, float* vals, const float* deg, float* d, const float* c_row,
                          hipStream_t stream);
int sa_normalize_rows_f64(const sa_csr* A, double* vals, const double* deg, double* d, const double* c_row,
                          hipStream_t stream);
int sa_normalize_f64(int64_t nnz, double* vals, int64_t n, const double* deg, double* d, double c,
                     hipStream_t stream);
/* tile_rows[t] = first row r with indptr[r] >= t * tile_nnz (t = 0..ntiles; tile_rows[ntiles] = n). */
int sa_csr_tiles(int64_t n, const int32_t* indptr, int32_t tile_nnz, int32_t* tile_rows, int64_t ntiles,
                 hipStream_t stream);

/* XCD panels from the plain CSR (sorted columns). pbound: npanels + 1 column boundaries (device).
 * counts (npanels x n int32, device) receives per-panel row counts; after the caller's exclusive scan
 * of counts into A->pindptr (npanels x (n + 1), absolute offsets), sa_panel_scatter copies the entries.
 * sa_panel_tiles fills the per-panel tile rows (tile t of panel p: rows whose first panel entry is in
 * [pindptr[p][0] + t tile_nnz, ...)). */
int sa_panel_count(const sa_csr* A, int32_t npanels, const int32_t* pbound, int32_t* counts, hipStream_t stream);
int sa_panel_scatter_f32(const sa_csr* A, int32_t npanels, const int32_t* pbound, int32_t* pindptr, int32_t* pindices,
                         float* pvals, hipStream_t stream);
int sa_panel_scatter_f64(const sa_csr* A, int32_t npanels, const int32_t* pbound, int32_t* pindptr, int32_t* pindices,
                         double* pvals, hipStream_t stream);
/* Fold the signs of the panel values into bit 31 of pindices when every |pvals[e]| equals mag;
 * *mismatch (device, zeroed by the caller) is set to 1 if one does not (pindices are then unusable). */
int sa_panel_sign_fold_f32(int64_t nnz, int32_t* pindices, const float* pvals, double mag, int32_t* mismatch,
                           hipStream_t stream);
int sa_panel_sign_fold_f64(int64_t nnz, int32_t* pindices, const double* pvals, double mag, int32_t* mismatch,
                           hipStream_t stream);
int sa_panel_tiles(int64_t n, const int32_t* pindptr_p, int32_t tile_nnz, int32_t* tile_rows, int64_t ntiles,
                   hipStream_t stream);

/* LDS-blocked layout from the plain CSR (sorted columns); no counterpart in the reference (whose SpMV
 * is cuSPARSE's). ntiles = 8 * bnr * bnc; pbound as for the panels, multiples of 64. A run is a row's
 * entries in one tile. A tile holds, in order: the long region (runs of more than 64 entries, cut into
 * 64-entry pieces stacked in 16 chunk columns by local row mod 16: chunk r * 16 + c is in column c,
 * and the SpMV kernel gives chunk k of a tile to wave k mod 16, so a row's pieces are summed by one
 * wave in order); the runs of 2..64 entries, each in a slot of 2, 4, ..., 64 entries, largest first (no
 * slot straddles a chunk); and, from the next multiple of 32, the singletons (runs of one) in rounds
 * over their row banks (local row mod 32), so that a half-wave's row-sum updates hit distinct LDS
 * banks. sa_block_plan counts the runs of each tile into runs[128 t + ...] (runs: 128 * ntiles int32,
 * zeroed by the caller: per size class at 0..6, per row bank of the singletons at 16..47, long-run
 * pieces per chunk column at 80..95; 8..14, 48..79 and 96..111 are the fill's cursors) and scans the
 * tiles' chunk counts into bchunk (ntiles + 1); *flag (zeroed by the caller) is set to 4 if the
 * entries would pass 2^31. sa_block_fill then writes the nent = 64 * bchunk[ntiles] entries and,
 * unless mag > 0 (sign-packed: every |J/c| equals mag, else *flag is set to 2), their values; *flag
 * bit 8 means the plan did not come from this matrix. */
int sa_block_plan(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bcols, int32_t bnr, int32_t bnc,
                  int32_t* runs, int32_t* bchunk, int32_t* flag, hipStream_t stream);
int sa_block_fill_f32(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bcols, int32_t bnr, int32_t bnc,
                      int32_t* runs, const int32_t* bchunk, int64_t nent, uint32_t* bent, float* bvals, double mag,
                      int32_t* flag, hipStream_t stream);
int sa_block_fill_f64(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bcols, int32_t bnr, int32_t bnc,
                      int32_t* runs, const int32_t* bchunk, int64_t nent, uint32_t* bent, double* bvals, double mag,
                      int32_t* flag, hipStream_t stream);

/* Column-streamed SpMV layout (bmode = 2, or its row-owned form bmode = 4 built with the one-panel
 * pbound; see sa_csr; no counterpart in the reference). Built in three steps on the caller's buffers:
 *   sa_cstream_plan: one 64-bit key per stored entry (workgroup << 40 | column in the panel << 15 |
 *     local row) and its entry index, radix-sorted (keys, idx -> keys_out, idx_out; tmp of
 *     sa_cstream_sort_bytes bytes), then start[w] = the first sorted entry of workgroup w
 *     (8 * bnr + 1 values). *flag bit 1: a row block taller than brows (or than the kernel's LDS
 *     holds), bit 2: a panel wider than 2^25 columns.
 *   the caller turns start into bchunk (chunks of 256 entries per workgroup, exclusive scan);
 *   sa_cstream_fill_f32: writes the nchunks * 256 entries, bcbase and (mag = 0) the values; *flag
 *     bit 4: a chunk spans more than 65536 columns, bit 8: a coupling's magnitude is not mag. */
int sa_cstream_sort_bytes(int64_t nnz, size_t* bytes);
/* Two-pass layout (bmode = 3, see sa_csr; no counterpart in the reference):
 *   sa_pblock_plan: one key per stored entry (tile << 30 | local row << 15 | column in the block) and
 *     its entry index, radix-sorted on the caller's buffers (tmp of sa_cstream_sort_bytes bytes),
 *     then pbtoff (ncb * bnr + 1). Column blocks: cbound (ncb + 1 boundaries, multiples of 64 except
 *     cbound[ncb] = n, each block at most sa_pblock_cols() wide; the layout's pbound) and cgran
 *     (ceil(n / 64) entries: the block of columns [64 g, 64 g + 64)). *flag bit 1: a row block taller
 *     than brows, bit 2: a column block wider than sa_pblock_cols().
 *   the caller derives pbrtoff (the tiles' sizes pbtoff[t + 1] - pbtoff[t], scanned in row-block
 *     order);
 *   sa_pblock_fill_f32: pbent and (mag = 0) the values in stored order, pbrow in products order;
 *     *flag bit 8: a coupling's magnitude is not mag. */
int sa_pblock_cols(void);
int sa_pblock_plan(const sa_csr* A, int32_t brows, int32_t bnr, const int32_t* cbound, int32_t ncb,
                   const int32_t* cgran, uint64_t* keys, int32_t* idx, uint64_t* keys_out, int32_t* idx_out, void* tmp,
                   size_t tmp_bytes, int32_t* toff, int32_t* flag, hipStream_t stream);
int sa_pblock_fill_f32(const sa_csr* A, int32_t brows, int32_t bnr, int32_t ncb, const uint64_t* keys,
                       const int32_t* idx, const int32_t* toff, const int32_t* rtoff, uint16_t* pbent,
                       uint16_t* pbrow, float* bvals, double mag, int32_t* flag, hipStream_t stream);
int sa_cstream_max_rows(void);
int sa_cstream_pad_row(void);
int sa_cstream_plan(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bnr, uint64_t* keys, int32_t* idx,
                    uint64_t* keys_out, int32_t* idx_out, void* tmp, size_t tmp_bytes, int32_t* start, int32_t* flag,
                    hipStream_t stream);
int sa_cstream_fill_f32(const sa_csr* A, const int32_t* pbound, int32_t brows, int32_t bnr, const uint64_t* keys,
                        const int32_t* idx, const int32_t* start, const int32_t* bchunk, int64_t nchunks,
                        uint32_t* bent, float* bvals, int32_t* bcbase, double mag, int32_t* flag,
                        hipStream_t stream);

/* ---- Per-alpha scaling (models/spectral_annealing.py:316, :34) ------------------------------ */
/* s_i = 1 / sqrt(max(d_i^alpha, clamp)); state.r[1] = sum_i d_i^alpha (degree_scale, :326);
 * state.r[2] = max_i d_i^(1 - alpha) (the Gershgorin bound g, :579). */
int sa_alpha_scaling_f32(int64_t n, const float* d, double alpha, double clamp, float* s, void* work,
                         hipStream_t stream);
int sa_alpha_scaling_f64(int64_t n, const double* d, double alpha, double clamp, double* s, void* work,
                         hipStream_t stream);

/* ---- Fused scaled operator (NAlphaOperator._matvec / SRhoOperator._matvec, :38-41, :52-53) -- */
/* y = a * x + b * s (.) (J/c) (s (.) x). a = rho, b = -1 is S_rho; a = 0, b = 1 is N_alpha.
 * z (n, caller scratch) receives s (.) x, so the SpMV gathers one vector. */
int sa_spmv_scaled_f32(const sa_csr* A, const float* s, const float* x, double a, double b, float* z, float* y,
                       hipStream_t stream);
int sa_spmv_scaled_f64(const sa_csr* A, const double* s, const double* x, double a, double b, double* z, double* y,
                       hipStream_t stream);
/* The allocation-free form of sa_spmv_scaled_f32, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. The prescaled vector z = s (.) x lives in the library's own static
 * device scratch of 32 MiB per device, so n <= 2^23 (larger n: SA_E_ARG with a message); the
 * panel or block layout of A, if any, brings its own partial-sum buffer as with sa_spmv_scaled_f32.
 * The scratch's device address is looked up on the first call on each device. Calls on one device
 * share its scratch: calls that overlap in time on one device (different streams) race. Same
 * kernels and results as sa_spmv_scaled_f32, bit for bit. */
int sa_spmv_scaled_static_f32(const sa_csr* A, const float* s, const float* x, double a, double b, float* y,
                              hipStream_t stream);

/* The bare fused kernel the Lanczos step launches: y = a * x + b * s (.) (J/c) g, with the gather vector
 * g = s (.) x prescaled by the caller (x may be NULL when a == 0). */
int sa_spmv_gather_f32(const sa_csr* A, const float* s, const float* g, double a, const float* x, double b, float* y,
                       hipStream_t stream);
int sa_spmv_gather_f64(const sa_csr* A, const double* s, const double* g, double a, const double* x, double b,
                       double* y, hipStream_t stream);

/* ---- Warm-started Lanczos on N_alpha (utils/eigsh.py:10-17, models/spectral_annealing.py:671-677)
 * Basis V is (ncv + 1) rows of n entries, row-major with leading dimension ldv >= n (pad ldv so that
 * ldv * sizeof(T) is not a multiple of a large power of two: with n = 2^22 unpadded rows start on the
 * same HBM channels and the basis streams camp on them). Z, W are n-vectors of scratch. The step,
 * thick restart and start kernels use 16-byte loads: V, Z, W and s 16-byte aligned, ldv * sizeof(T)
 * a multiple of 16. */
/* orth_tau: a step's Gram-Schmidt update subtracts only the last two basis rows unless the projection
 * on an older row exceeds orth_tau * beta_j (then every row); sqrt(eps) / 2 of the vectors' type keeps
 * the basis semi-orthogonal, which leaves Ritz values as accurate as full reorthogonalisation. 0
 * subtracts every row at every step (the reference eigsh's full reorthogonalisation). */
int sa_lanczos_begin(void* work, int32_t ncv, int32_t maxiter, double tol, double bound, double pred_tol,
                     int32_t keep, double orth_tau, hipStream_t stream);
/* Start vector V[0] = x (.) max(d, 1e-12)^pexp (the diagonal predictor, utils/predictors.py:6-26, with
 * pexp = delta_alpha / 2; pexp = 0 copies x) or cx * x + cy * y when y != NULL (the secant predictor,
 * :29-56). Z = s (.) V[0]; sigma_0 = 1 / ||V[0]||; state.r[3] = ||V[0]||. gate = 1 marks the vector as
 * predicted: step 0 then also reports beta_0 (the predictor residual, spectral_annealing.py:605-612)
 * and sets SA_REJECTED if it exceeds pred_tol. only_if >= 0 makes the call a no-op unless
 * state.status == only_if (device-side gating, so restarts need no host round trip). */
int sa_lanczos_start_f32(void* work, int64_t n, const float* s, const float* x, const float* d, double pexp,
                         double cx, const float* y, double cy, float* V, int64_t ldv, void* Vh, float* Z,
                         int32_t gate, int32_t only_if, hipStream_t stream);
int sa_lanczos_start_f64(void* work, int64_t n, const double* s, const double* x, const double* d, double pexp,
                         double cx, const double* y, double cy, double* V, int64_t ldv, void* Vh, double* Z,
                         int32_t gate, int32_t only_if, hipStream_t stream);
/* One Lanczos step: the fused SpMV, the projections on the cycle's basis and the Gram-Schmidt update
 * (selective reorthogonalisation, see sa_lanczos_begin), and the FP64 Ritz solve of the projected
 * matrix; no-op unless state.status == SA_RUNNING.
 * Vh (may be NULL; the same for start, step and thick restart of one solve) is a shadow basis: ncv + 1
 * rows of leading dimension ldv in _Float16 (_f32) or float (_f64), 16-byte aligned, ldv then a
 * multiple of 8 (_f32) or 4 (_f64) entries. The calls keep it
 * equal to V rounded, and the projections on rows older than j - 1 read it instead of V (they only
 * decide and correct components below sqrt(eps)): about half the projections' traffic. */
int sa_lanczos_step_f32(const sa_csr* A, void* work, const float* s, float* V, int64_t ldv, void* Vh, float* Z,
                        float* W, hipStream_t stream);
int sa_lanczos_step_f64(const sa_csr* A, void* work, const double* s, double* V, int64_t ldv, void* Vh, double* Z,
                        double* W, hipStream_t stream);
/* Thick restart (gated on SA_RESTART): keep the `keep` largest Ritz pairs of the full basis. The basis
 * is rotated in place: rows 0..keep-1 become the Ritz vectors, row keep the last Lanczos vector, Z its
 * scaled copy, and the projected matrix becomes diag(theta) bordered by the arrow couplings. With
 * keep = 1 this is the restart from the Ritz vector alone (the k = 1 thick restart of the reference's
 * eigsh). */
int sa_lanczos_thick_restart_f32(void* work, int64_t n, const float* s, float* V, int64_t ldv, void* Vh,
                                 float* Z, hipStream_t stream);
int sa_lanczos_thick_restart_f64(void* work, int64_t n, const double* s, double* V, int64_t ldv, void* Vh,
                                 double* Z, hipStream_t stream);
/* Ritz vector y = sum_m ritz_m v_m of the current cycle; state.r[4] = ||y||, state.r[5] = y . y_last
 * (y_last may be NULL). only_if as in sa_lanczos_start. */
int sa_lanczos_ritz_f32(void* work, int64_t n, const float* V, int64_t ldv, const float* y_last, float* y,
                        int32_t only_if, hipStream_t stream);
int sa_lanczos_ritz_f64(void* work, int64_t n, const double* V, int64_t ldv, const double* y_last, double* y,
                        int32_t only_if, hipStream_t stream);

/* ---- Vector helpers ------------------------------------------------------------------------- */
/* state.r[slot] = ||a*x + b*y||_2 (y may be NULL); state.r[slot+1] = x . y (0 when y is NULL). */
int sa_norm_dot_f32(void* work, int64_t n, const float* x, const float* y, double a, double b, int32_t slot,
                    hipStream_t stream);
int sa_norm_dot_f64(void* work, int64_t n, const double* x, const double* y, double a, double b, int32_t slot,
                    hipStream_t stream);
/* out_i = scale * x_i (in T). */
int sa_scale_copy_f32(int64_t n, const float* x, double scale, float* out, hipStream_t stream);
/* Columns the instance-staged layout stages per pass for elements of elem_bytes (4 or 8; -1 else). */
int sa_inst_columns_per_pass(int32_t elem_bytes);
int sa_scale_copy_f64(int64_t n, const double* x, double scale, double* out, hipStream_t stream);

/* out = a x + b y + c z (FP64 arithmetic; y or z may be NULL when its coefficient is 0). The vector
 * bookkeeping of the predictors (cf. utils/predictors.py:29-56). */
/* Diagonal predictor's scaling (utils/predictors.py:19-20): out = max(d, 1e-12)^pexp (.) y in the
 * vectors' type; the caller normalises. */
int sa_diag_scale_f32(int64_t n, const float* y, const float* d, double pexp, float* out, hipStream_t stream);
int sa_diag_scale_f64(int64_t n, const double* y, const double* d, double pexp, double* out, hipStream_t stream);
/* Candidate-major int8 +-1 spins [k][n] (the reference's spins_list, models/spectral_annealing.py:1058)
 * -> candidate bits [k][nwords]: bit l of word w is spin 64 w + l > 0. */
int sa_pack_spins(int64_t n, int32_t k, const int8_t* spins, int64_t nwords, uint64_t* bits, hipStream_t stream);
int sa_combine3_f32(int64_t n, double a, const float* x, double b, const float* y, double c, const float* z,
                    float* out, hipStream_t stream);
int sa_combine3_f64(int64_t n, double a, const double* x, double b, const double* y, double c, const double* z,
                    double* out, hipStream_t stream);
/* Counter-based standard normal vector (Box-Muller on SplitMix64 streams (seed, stream_id, i)). */
int sa_fill_normal_f32(int64_t n, uint64_t seed, uint64_t stream_id, float* out, hipStream_t stream);
int sa_fill_normal_f64(int64_t n, uint64_t seed, uint64_t stream_id, double* out, hipStream_t stream);

/* ---- Rounding and signature (models/spectral_annealing.py:330-345, :557-561) ---------------- */
/* y_last = scale * y (scale = sign / ||y||, the sign alignment and normalisation of :653-660 and
 * :331-340); bits word w, bit l = 1 iff scale * y[64 w + l] >= 0 (spin +1, :345); state.hash =
 * order-aware 64-bit signature of the bits. y_last may be NULL (no copy). */
int sa_round_f32(void* work, int64_t n, const float* y, double scale, float* y_last, uint64_t* bits,
                 hipStream_t stream);
int sa_round_f64(void* work, int64_t n, const double* y, double scale, double* y_last, uint64_t* bits,
                 hipStream_t stream);
/* state.i64[0] = number of differing words of two bit vectors (0: identical candidates). */
int sa_bits_compare(void* work, int64_t nwords, const uint64_t* a, const uint64_t* b, hipStream_t stream);

/* ---- Energies (models/spectral_annealing.py:925-976, exact for integer couplings) ----------- */
/* e2[c] = sum_i s_i sum_j J_ij s_j (int64, exact) for k candidate bit vectors (candidate-major,
 * nwords words each); the energy is -e2 / 2. */
int sa_energy_bits_i64(const sa_csr* A, int32_t k, const uint64_t* cand_bits, int64_t nwords, int64_t* e2,
                       hipStream_t stream);

/* ---- Batched deterministic local search (flip_refinement, :1039-1112; gpu_final_sweep :983-1036)
 * Candidates are node-major: spins int8 [n][kp], fields int32 [n][kp], kp = 64 * groups. */
/* node-major spin bits sb[i * groups + g] bit l = candidate 64 g + l, from candidate-major bits. */
int sa_bits_transpose(int64_t n, int32_t k, const uint64_t* cand_bits, int64_t nwords, int32_t groups,
                      uint64_t* sb, hipStream_t stream);
/* spins, fields = J s, e2[c] = sum_i s_i h_i (int64 atomics; zero e2 first). */
int sa_ls_init(const sa_csr* A, int32_t k, int32_t groups, const uint64_t* sb, int8_t* spins, int32_t* fields,
               int64_t* e2, hipStream_t stream);
/* One colour class of one pass: node i of the class (selected if u53(mix3(seed, pass, i)) < frac)
 * flips candidate c iff s_ic * h_ic < 0, then adds 2 J_ij s_ic(new) to h_jc of every neighbour j.
 * flags[c] = 1 for candidates with a flip. class_nodes holds the class's nodes. */
int sa_ls_class_pass(const sa_csr* A, int32_t k, int32_t groups, const int32_t* class_nodes, int64_t count,
                     uint64_t seed, int64_t pass, double frac, int8_t* spins, int32_t* fields, int32_t* flags,
                     hipStream_t stream);
/* The whole search in one persistent (cooperative) launch: the schedule of flip_refinement
 * (models/spectral_annealing.py:1039-1112; main passes over a subset_fraction share of the nodes until
 * every candidate has had a pass without a flip, at most max_iterations, then num_rounds rounds of at
 * most final_sweep_iters passes over a final_sweep_fraction share, each ending at a pass without flips)
 * walked on the device, with the colour classes of pass q in the order of ascending
 * (mix3(seed ^ 0x0DE5C1A55C0FFEE1, q, class), class) and selection ids: main pass t -> t, final
 * round r -> 2^40 + r; each class as sa_ls_class_pass. class_nodes: the nodes grouped by class,
 * class r = class_nodes[class_off[r] .. class_off[r + 1]) (1 <= num_classes <= 1024). ctl: a zeroed
 * int32 workspace of sa_ls_search_ctl_ints(num_classes, groups) entries; afterwards ctl[4] = main
 * passes, ctl[5] = all passes, ctl[7] != 0 = the search stopped on a grid-barrier timeout. */
size_t sa_ls_search_ctl_ints(int32_t num_classes, int32_t groups);
int sa_ls_search(const sa_csr* A, int32_t k, int32_t groups, const int32_t* class_nodes, const int32_t* class_off,
                 int32_t num_classes, uint64_t seed, int32_t max_iterations, double subset_fraction, int32_t num_rounds,
                 int32_t final_sweep_iters, double final_sweep_fraction, int8_t* spins, int32_t* fields, int32_t* ctl,
                 hipStream_t stream);
/* e2[c] = sum_i s_ic h_ic from maintained fields (zero e2 first). */
int sa_ls_energy(int64_t n, int32_t k, int32_t groups, const int8_t* spins, const int32_t* fields, int64_t* e2,
                 hipStream_t stream);
/* Floating-point couplings (the reference's update_spins_kernel_float32/64 and fp32_energies,
 * models/spectral_annealing.py:821-976): the same search with the FP32 normalised couplings J / c
 * (fvals) and FP64 fields. The host takes this path only when every nonzero |J / c| is at least
 * 2^-30 (fields are then exact sums, so the result does not depend on the order of the atomic
 * updates and matches the CPU restatement bit for bit). Energies: e2[c] = sum_i s_ic h_ic in FP64,
 * summed in a fixed order through part (sa_ls_energy_f64_scratch(groups) bytes); E = -c e2 / 2. */
int sa_ls_init_f64(const sa_csr* A, const float* fvals, int32_t k, int32_t groups, const uint64_t* sb,
                   int8_t* spins, double* fields, hipStream_t stream);
int sa_ls_class_pass_f64(const sa_csr* A, const float* fvals, int32_t k, int32_t groups, const int32_t* class_nodes,
                         int64_t count, uint64_t seed, int64_t pass, double frac, int8_t* spins, double* fields,
                         int32_t* flags, hipStream_t stream);
int sa_ls_search_f64(const sa_csr* A, const float* fvals, int32_t k, int32_t groups, const int32_t* class_nodes,
                     const int32_t* class_off, int32_t num_classes, uint64_t seed, int32_t max_iterations,
                     double subset_fraction, int32_t num_rounds, int32_t final_sweep_iters,
                     double final_sweep_fraction, int8_t* spins, double* fields, int32_t* ctl, hipStream_t stream);
size_t sa_ls_energy_f64_scratch(int32_t groups);
int sa_ls_energy_f64(int64_t n, int32_t k, int32_t groups, const int8_t* spins, const double* fields, double* part,
                     double* e2, hipStream_t stream);
/* out = spins[:, c] as +-1 in T. */
int sa_ls_extract_f32(int64_t n, int32_t groups, int32_t c, const int8_t* spins, float* out, hipStream_t stream);
int sa_ls_extract_f64(int64_t n, int32_t groups, int32_t c, const int8_t* spins, double* out, hipStream_t stream);

/* ---- Colouring (flip order of the local search; cf. local_search_guards/common.py:136-159) --- */
/* Greedy colouring in the order (degree, hash(seed, i), i) descending, in rounds over a worklist: an
 * uncoloured node all of whose higher-priority neighbours are coloured takes the smallest colour none
 * of them has, so the result equals the sequential greedy colouring in that order.
 * sa_colour_start: colour[i] = -1, the first list holds every node. lists: 3n int32 (two worklists
 * and each node's resume point in its adjacency list); counts: 3 int32.
 * sa_colour_rounds runs rounds first_round .. first_round + rounds - 1; after round r,
 * counts[(r + 1) % 3] nodes are left (0: done, and later rounds do nothing). */
int sa_colour_start(const sa_csr* A, int32_t* colour, int32_t* lists, int32_t* counts, hipStream_t stream);
int sa_colour_rounds(const sa_csr* A, uint64_t seed, int32_t* colour, int32_t* lists, int32_t* counts,
                     int32_t first_round, int32_t rounds, hipStream_t stream);
/* The same rounds in largest-first order: degree, then the smaller index
 * (greedy_largest_first_coloring, models/local_search_guards/common.py:136-159). */
int sa_colour_rounds_lf(const sa_csr* A, int32_t* colour, int32_t* lists, int32_t* counts, int32_t first_round,
                        int32_t rounds, hipStream_t stream);

/* ---- Guarded local-search policies (ls_policy, models/local_search_guards/policies.py:13-52) -------
 * Candidates in the local search's layout (spins int8 [n][kp], exact fields [n][kp], kp = 64 groups);
 * integer couplings (J, int32 fields, int64 energies) or, in the _f64 forms, the FP32 couplings J / c
 * (fvals) with FP64 fields and energies in units of J / c.
 *
 * Fixed / Adaptive Guard (safe_flip_refinement, models/spectral_annealing_safe.py:1035-1335; one round
 * of _run_guarded_phase, :223-797). sa_guard_propose: node i is selected for candidate c iff
 * (mix3(seed, round, i) >> 11) < thresh[c] (kp uint64; 0 = inactive); a selected node whose spin
 * disagrees with the sign of its field flips (Jacobi: judged on the current spins). mask[i * groups + g]
 * bit l = candidate 64 g + l flips i; counts[c] += proposed flips; delta[c] += the exact energy change
 * of the proposal, sum_F 2 s_i h_i - 2 sum_F s_i sum_{j in F, j ~ i} J_ij s_j (zero counts and delta
 * first). sa_guard_commit: the proposal's flips for the candidates of accepted[g] (bit l), with the
 * field updates. */
int sa_guard_propose(const sa_csr* A, int32_t k, int32_t groups, const int8_t* spins, const int32_t* fields,
                     uint64_t seed, uint64_t round, const uint64_t* thresh, uint64_t* mask, int32_t* counts,
                     int64_t* delta, hipStream_t stream);
int sa_guard_propose_f64(const sa_csr* A, const float* fvals, int32_t k, int32_t groups, const int8_t* spins,
                         const double* fields, uint64_t seed, uint64_t round, const uint64_t* thresh, uint64_t* mask,
                         int32_t* counts, double* delta, hipStream_t stream);
int sa_guard_commit(const sa_csr* A, int32_t groups, int8_t* spins, int32_t* fields, const uint64_t* mask,
                    const uint64_t* accepted, hipStream_t stream);
int sa_guard_commit_f64(const sa_csr* A, const float* fvals, int32_t groups, int8_t* spins, double* fields,
                        const uint64_t* mask, const uint64_t* accepted, hipStream_t stream);
/* Independent-Set (conflict_free.py:82-296): one colour class; node i flips candidate c iff c is
 * active (active[g] bit l) and s h < -thr[c]; flips[c] and delta[c] (the change sum 2 s h) accumulate.
 * sa_guard_class_prep, before each step: energy[c] += delta[c], sweep[c] += flips[c], both cleared, and
 * thr[c] from the tolerance abs_tol + rel_tol max(1, |E / scale|) of the new energy
 * (safe_update_policy.py:4-27): floor(tol scale / 2) for integer fields (scale = c), tol / 2 in FP64. */
int sa_guard_class_step(const sa_csr* A, int32_t groups, const int32_t* class_nodes, int64_t count,
                        const uint64_t* active, const int32_t* thr, int8_t* spins, int32_t* fields, int32_t* flips,
                        int64_t* delta, hipStream_t stream);
int sa_guard_class_step_f64(const sa_csr* A, const float* fvals, int32_t groups, const int32_t* class_nodes,
                            int64_t count, const uint64_t* active, const double* thr, int8_t* spins, double* fields,
                            int32_t* flips, double* delta, hipStream_t stream);
int sa_guard_class_prep(int32_t groups, double scale, double abs_tol, double rel_tol, int64_t* energy,
                        int64_t* delta, int32_t* flips, int32_t* sweep, int32_t* thr, hipStream_t stream);
int sa_guard_class_prep_f64(int32_t groups, double abs_tol, double rel_tol, double* energy, double* delta,
                            int32_t* flips, int32_t* sweep, double* thr, hipStream_t stream);
/* Sequential-Oracle (sequential_oracle.py:22-121): one 1024-thread workgroup per candidate group runs
 * exact one-flip steepest descent (each step, every candidate flips its node of smallest s h, the first
 * on ties, while 2 s h < -tol of its energy). energy[c] (-sum s h / 2, kept in place), steps[g] = steps
 * with a flip, or -1 if improving flips remain after max_steps; flips[c] = its flips. */
int sa_guard_sequential(const sa_csr* A, int32_t k, int32_t groups, int8_t* spins, int32_t* fields, int64_t* energy,
                        double scale, double abs_tol, double rel_tol, int64_t max_steps, int64_t* steps,
                        int32_t* flips, hipStream_t stream);
int sa_guard_sequential_f64(const sa_csr* A, const float* fvals, int32_t k, int32_t groups, int8_t* spins,
                            double* fields, double* energy, double abs_tol, double rel_tol, int64_t max_steps,
                            int64_t* steps, int32_t* flips, hipStream_t stream);

/* ---- Eigenvector tangent and its adjoint: deflated MINRES -----------------------------------
 * No counterpart in the reference (a new predictor beside utils/predictors.py:6-56). For the unit
 * leading eigenvector y of N(alpha) = s (.) (J/c) (.) s, s = d^(-alpha/2), with eigenvalue lambda:
 *   dN/dalpha = -1/2 (L N + N L), L = diag(ln d), and with u = 1/2 L y, P = I - y y^T,
 *   dy/dalpha = P u + 2 lambda z,   z = (N - lambda I)^+ P u,
 * and the adjoint of a cotangent g is g . dy/dalpha = g . P u + 2 lambda (N - lambda I)^+ P g . u.
 * Both solves are MINRES (Paige & Saunders 1975) on P (N - shift I) P with a right-hand side P r,
 * FP32 or FP64 vectors with FP64 scalars, every scalar on the device. The state lives at the start
 * of its own work buffer (sa_work_bytes bytes); buffers are caller-owned, n entries of T each. */
typedef struct sa_minres_state {
  int32_t status;      /* SA_RUNNING, SA_CONVERGED (phibar <= tol * beta1), SA_MAXITER, SA_BROKEN */
  int32_t itn;         /* iterations done */
  int32_t maxit;
  int32_t pending;     /* 1 while the last iteration's update of x is still to be applied */
  double shift;        /* lambda of P (N - lambda I) P */
  double tol;
  double beta1;        /* ||P r|| */
  double beta;         /* ||r2|| (current raw Lanczos vector) */
  double oldb;         /* ||r1|| */
  double cr1;          /* coefficient of r1 in the recurrence (0 in the first iteration) */
  double alfa;
  double g;            /* y . (N - shift I) r2: the deflated component */
  double dbar, epsln, cs, sn, phibar;
  double pe_oldeps, pe_delta, pe_denom, pe_phi; /* coefficients of the pending x update */
  double c0;           /* y . r of the right-hand side, then y . x at the end */
  double rnorm;        /* phibar after the last iteration: ||P r - A x|| */
} sa_minres_state;

typedef struct sa_minres_bufs {
  void* r[2];          /* the two latest raw Lanczos vectors of the recurrence */
  void* w[3];          /* MINRES search directions */
  void* x;             /* the solution */
  void* z;             /* s (.) r2, the SpMV's gathered vector */
  void* q;             /* N r2 */
} sa_minres_bufs;

size_t sa_minres_state_bytes(void);

/* u = 1/2 ln(max(d, clamp)) (.) y: the tangent's right-hand side. */
int sa_tangent_rhs_f32(int64_t n, const float* y, const float* d, double clamp, float* u, hipStream_t stream);
int sa_tangent_rhs_f64(int64_t n, const double* y, const double* d, double clamp, double* u, hipStream_t stream);

/* Start a solve of P (N - shift I) P x = P r (y unit): b = P r, x = 0, beta1 = ||b||. */
int sa_minres_begin_f32(void* work, int64_t n, const float* s, const float* y, const float* r, double shift,
                        double tol, int32_t maxit, const sa_minres_bufs* B, hipStream_t stream);
int sa_minres_begin_f64(void* work, int64_t n, const double* s, const double* y, const double* r, double shift,
                        double tol, int32_t maxit, const sa_minres_bufs* B, hipStream_t stream);

/* Queue `count` iterations (each: the fused SpMV, one projection pass, one update pass and two
 * one-workgroup finishers); every launch is skipped once the solve has stopped. */
int sa_minres_iterate_f32(const sa_csr* A, void* work, const float* s, const float* y, int32_t count,
                          const sa_minres_bufs* B, hipStream_t stream);
int sa_minres_iterate_f64(const sa_csr* A, void* work, const double* s, const double* y, int32_t count,
                          const sa_minres_bufs* B, hipStream_t stream);

/* Apply the pending update of x and project x onto y's complement (x -= (y . x) y). */
int sa_minres_end_f32(void* work, int64_t n, const float* y, const sa_minres_bufs* B, hipStream_t stream);
int sa_minres_end_f64(void* work, int64_t n, const double* y, const sa_minres_bufs* B, hipStream_t stream);

/* ---- Batched sweep: many instances packed block-diagonally (sa_batch) ------------------------
 * No counterpart in the reference (it solves one instance per call); these replace, for a batch,
 * the calls above that cite it. Each instance runs its own eigensolve in its own work buffer (its
 * own j, status, Ritz solve and restarts); every launch covers all instances (or those of the
 * descriptor's inst_mask) and is skipped per workgroup by instances not in the status it needs. The
 * batched step uses the plain CSR layout. */
int sa_batch_alpha_scaling_f32(const sa_batch* Bt, const float* d, double alpha, double clamp, float* s,
                               hipStream_t stream);
int sa_batch_alpha_scaling_f64(const sa_batch* Bt, const double* d, double alpha, double clamp, double* s,
                               hipStream_t stream);
/* The same with instance b at its own alpha[b] (device array of B): instances that stop at different
 * steps move on to their next alpha without waiting for the rest of the batch. */
int sa_batch_alpha_scaling_each_f32(const sa_batch* Bt, const float* d, const double* alpha, double clamp, float* s,
                                    hipStream_t stream);
int sa_batch_alpha_scaling_each_f64(const sa_batch* Bt, const double* d, const double* alpha, double clamp,
                                    double* s, hipStream_t stream);
/* dst = src on the real rows of every (unmasked) instance. */
int sa_batch_copy_f32(const sa_batch* Bt, const float* src, float* dst, hipStream_t stream);
int sa_batch_copy_f64(const sa_batch* Bt, const double* src, double* dst, hipStream_t stream);
/* bounds: device array of B breakdown bounds (rho_b + g_b). */
int sa_batch_lanczos_begin(const sa_batch* Bt, int32_t ncv, int32_t maxiter, double tol, const double* bounds,
                           double pred_tol, int32_t keep, double orth_tau, hipStream_t stream);
int sa_batch_lanczos_start_f32(const sa_batch* Bt, const float* s, const float* x, const float* d, double pexp,
                               double cx, const float* y, double cy, float* V, int64_t ldv, float* Z, int32_t gate,
                               int32_t only_if, hipStream_t stream);
int sa_batch_lanczos_start_f64(const sa_batch* Bt, const double* s, const double* x, const double* d, double pexp,
                               double cx, const double* y, double cy, double* V, int64_t ldv, double* Z,
                               int32_t gate, int32_t only_if, hipStream_t stream);
/* The diagonal-predictor start x0 = d^pexp[b] (.) x with instance b's own exponent pexp[b] (device
 * array of B; cx = 1, no second vector, every unmasked instance). */
int sa_batch_lanczos_start_each_f32(const sa_batch* Bt, const float* s, const float* x, const float* d,
                                    const double* pexp, float* V, int64_t ldv, float* Z, int32_t gate,
                                    hipStream_t stream);
int sa_batch_lanczos_start_each_f64(const sa_batch* Bt, const double* s, const double* x, const double* d,
                                    const double* pexp, double* V, int64_t ldv, double* Z, int32_t gate,
                                    hipStream_t stream);
int sa_batch_lanczos_step_f32(const sa_csr* A, const sa_batch* Bt, const float* s, float* V, int64_t ldv, float* Z,
                              float* W, hipStream_t stream);
int sa_batch_lanczos_step_f64(const sa_csr* A, const sa_batch* Bt, const double* s, double* V, int64_t ldv,
                              double* Z, double* W, hipStream_t stream);
int sa_batch_lanczos_thick_restart_f32(const sa_batch* Bt, const float* s, float* V, int64_t ldv, float* Z,
                                       hipStream_t stream);
int sa_batch_lanczos_thick_restart_f64(const sa_batch* Bt, const double* s, double* V, int64_t ldv, double* Z,
                                       hipStream_t stream);
int sa_batch_lanczos_ritz_f32(const sa_batch* Bt, const float* V, int64_t ldv, const float* y_last, float* y,
                              int32_t only_if, hipStream_t stream);
int sa_batch_lanczos_ritz_f64(const sa_batch* Bt, const double* V, int64_t ldv, const double* y_last, double* y,
                              int32_t only_if, hipStream_t stream);
/* Per instance: y_last = (sign / r[4]) y (sign = -1 if align and r[5] < 0), its bits, state.hash. */
int sa_batch_round_f32(const sa_batch* Bt, const float* y, float* y_last, uint64_t* bits, int32_t align,
                       hipStream_t stream);
int sa_batch_round_f64(const sa_batch* Bt, const double* y, double* y_last, uint64_t* bits, int32_t align,
                       hipStream_t stream);
/* e2[b * k + c] += sum over instance b's rows (callers zero e2 first). */
int sa_batch_energy_bits_i64(const sa_csr* A, const sa_batch* Bt, int32_t k, const uint64_t* cand_bits,
                             int64_t nwords, int64_t* e2, hipStream_t stream);
int sa_batch_ls_energy(const sa_batch* Bt, int32_t k, int32_t groups, const int8_t* spins, const int32_t* fields,
                       int64_t* e2, hipStream_t stream);
/* out[c][w] = in[src[c * B + b(w)]][w], b(w) the instance owning word w (word_off: B + 1). */
int sa_bits_gather(int64_t nwords, int32_t k, int32_t B, const int64_t* word_off, const int32_t* src,
                   const uint64_t* in, uint64_t* out, hipStream_t stream);

/* ---------------------------------------------------------------------------------------------
 * Seeded synthetic instances, built on the host (no device memory, no stream): symmetric CSR with
 * sorted columns, no self-loops, no duplicates, isolated nodes dropped. No counterpart in the
 * reference (its instances are text files, utils/data_loader.py); they make the large test and
 * benchmark graphs (Erdos-Renyi n = 2^22, Barabasi-Albert n = 2^25) in seconds. Every draw is a
 * counter-based hash, so the graph is the same for any thread count. coupling_kind: 0 = every
 * coupling -1 (MaxCut form), 1 = uniform +-1 per undirected edge. threads <= 0: all hardware threads.
 * ------------------------------------------------------------------------------------------- */
typedef struct sa_graph sa_graph;   /* opaque host-side CSR owned by the library until sa_graph_free */
int sa_graph_erdos_renyi(int64_t n, double p, uint64_t seed, int32_t coupling_kind, int32_t threads,
                         sa_graph** out);
int sa_graph_barabasi_albert(int64_t n, int32_t m, uint64_t seed, int32_t coupling_kind, int32_t threads,
                             sa_graph** out);
int64_t sa_graph_n(const sa_graph* g);      /* nodes after isolated ones were dropped */
int64_t sa_graph_nnz(const sa_graph* g);    /* stored couplings (both triangles) */
/* Copies the CSR into caller arrays of n + 1, nnz and nnz entries. */
int sa_graph_export(const sa_graph* g, int64_t* indptr, int32_t* indices, double* vals);
void sa_graph_free(sa_graph* g);

#ifdef __cplusplus
}
#endif

#endif /* SPECANN_H */
