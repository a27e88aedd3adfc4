"""ctypes binding of libspecann (include/specann.h).

The solver has no CPU fallback: if the library is missing or no GPU is visible, the calls below
raise. Status codes map onto the exception types the reference's solver loop distinguishes
(models/spectral_annealing.py:70-77, utils/failures.py:15-25): SA_E_OOM is a MemoryError, so the
run is tagged "out of memory"; SA_E_BREAKDOWN is an EigensolverBreakdown (numerical, the alpha is
skipped).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int32, c_int64, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libspecann.so")

SA_OK = 0
SA_E_ARG = -1
SA_E_LAUNCH = -2
SA_E_BREAKDOWN = -3
SA_E_OOM = -4
SA_E_NONINTEGER = -5

SA_RUNNING = 0
SA_CONVERGED = 1
SA_MAXITER = 2
SA_RESTART = 3
SA_REJECTED = 4
SA_BROKEN = -1

SA_MAX_NCV = 128
SA_MAX_KEEP = 16


def padded_ld(n: int, elem_bytes: int) -> int:
    """Leading dimension >= n whose byte stride is n rounded up to 4 KiB plus 4352 bytes (17 x 256):
    rows of a multi-row buffer then start on different HBM channels whatever n is."""
    per4k = 4096 // elem_bytes
    return ((n + per4k - 1) // per4k) * per4k + 4352 // elem_bytes


class SpecAnnLibraryError(RuntimeError):
    """libspecann is missing or a call failed for a non-numerical reason."""


class EigensolverBreakdown(RuntimeError):
    """The eigensolver produced an unusable result (models/spectral_annealing.py:70-71)."""


class DeviceOutOfMemory(MemoryError):
    """A HIP allocation or launch ran out of device memory."""


class SaCsr(ctypes.Structure):
    _fields_ = [
        ("n", c_int64),
        ("nnz", c_int64),
        ("indptr", c_void_p),
        ("indices", c_void_p),
        ("vals", c_void_p),
        ("jint", c_void_p),
        ("tile_rows", c_void_p),
        ("ntiles", c_int64),
        ("tile_nnz", c_int32),
        ("lanes_per_row", c_int32),
        ("npanels", c_int32),
        ("pad0", c_int32),
        ("pindptr", c_void_p),
        ("pindices", c_void_p),
        ("pvals", c_void_p),
        ("ptile_rows", c_void_p),
        ("ptile_off", c_void_p),
        ("pmax_tiles", c_int64),
        ("partial", c_void_p),
        ("pld", c_int64),
        ("psigned", c_int32),
        ("pad2", c_int32),
        ("pmag", c_double),
        ("bmode", c_int32),
        ("brows", c_int32),
        ("bcols", c_int32),
        ("bnc", c_int32),
        ("bnr", c_int32),
        ("pad3", c_int32),
        ("pbound", c_void_p),
        ("bchunk", c_void_p),
        ("bent", c_void_p),
        ("bvals", c_void_p),
        ("rbound", c_void_p),
        ("imode", c_int32),
        ("irows", c_int32),
        ("ilanes", c_int32),
        ("ipasses", c_int32),
        ("isigned", c_int32),
        ("pad4", c_int32),
        ("inwg", c_int64),
        ("iinst", c_void_p),
        ("irow", c_void_p),
        ("ibase", c_void_p),
        ("ireal", c_void_p),
        ("iend", c_void_p),
        ("isplit", c_void_p),
        ("ient", c_void_p),
        ("imag", c_void_p),
        ("bcbase", c_void_p),
        ("pbent", c_void_p),
        ("pbrow", c_void_p),
        ("pbtoff", c_void_p),
        ("pbrtoff", c_void_p),
        ("pbcontrib", c_void_p),
        ("pbnc", c_int32),
        ("pad5", c_int32),
    ]


class SaState(ctypes.Structure):
    _fields_ = [
        ("status", c_int32),
        ("j", c_int32),
        ("steps", c_int32),
        ("cycles", c_int32),
        ("ncv", c_int32),
        ("maxiter", c_int32),
        ("gate", c_int32),
        ("kept", c_int32),
        ("keep", c_int32),
        ("orth_lo", c_int32),
        ("theta", c_double),
        ("res", c_double),
        ("first_beta", c_double),
        ("tol", c_double),
        ("bound", c_double),
        ("pred_tol", c_double),
        ("orth_tau", c_double),
        ("r", c_double * 8),
        ("hash", c_uint64),
        ("i64", c_int64 * 4),
        ("a", c_double * SA_MAX_NCV),
        ("b", c_double * SA_MAX_NCV),
        ("sigma", c_double * (SA_MAX_NCV + 1)),
        ("h", c_double * (SA_MAX_NCV + 1)),
        ("ritz", c_double * SA_MAX_NCV),
        ("arrow", c_double * SA_MAX_NCV),
        ("counters", c_uint32 * 16),
    ]


# Bytes of the leading part of SaState the host reads after each batch of launches.
STATE_HEAD_BYTES = SaState.a.offset


class SaBatch(ctypes.Structure):
    _fields_ = [
        ("B", c_int32),
        ("rows_per_block", c_int32),
        ("row_off", c_void_p),
        ("row_end", c_void_p),
        ("blk_off", c_void_p),
        ("works", c_void_p),
        ("stride", c_int64),
        ("nblocks", c_int64),
        ("tile_inst", c_void_p),
        ("inst_mask", c_void_p),
        ("tile_list", c_void_p),
        ("ntile_list", c_int64),
    ]


class SaMinresState(ctypes.Structure):
    _fields_ = [
        ("status", c_int32),
        ("itn", c_int32),
        ("maxit", c_int32),
        ("pending", c_int32),
        ("shift", c_double),
        ("tol", c_double),
        ("beta1", c_double),
        ("beta", c_double),
        ("oldb", c_double),
        ("cr1", c_double),
        ("alfa", c_double),
        ("g", c_double),
        ("dbar", c_double),
        ("epsln", c_double),
        ("cs", c_double),
        ("sn", c_double),
        ("phibar", c_double),
        ("pe_oldeps", c_double),
        ("pe_delta", c_double),
        ("pe_denom", c_double),
        ("pe_phi", c_double),
        ("c0", c_double),
        ("rnorm", c_double),
    ]


class SaMinresBufs(ctypes.Structure):
    _fields_ = [
        ("r", c_void_p * 2),
        ("w", c_void_p * 3),
        ("x", c_void_p),
        ("z", c_void_p),
        ("q", c_void_p),
    ]

_P = c_void_p
_SIGNATURES = {
    "sa_version": (c_int32, []),
    "sa_last_error": (c_int32, [ctypes.c_char_p, c_size_t]),
    "sa_work_bytes": (c_size_t, [c_int64]),
    "sa_inst_columns_per_pass": (c_int32, [c_int32]),
    "sa_state_bytes": (c_size_t, []),
    "sa_prepare_values_f32": (c_int32, [c_int64, _P, _P, _P, _P, _P]),
    "sa_prepare_values_f64": (c_int32, [c_int64, _P, _P, _P, _P, _P]),
    "sa_abs_degree_f32": (c_int32, [POINTER(SaCsr), _P, _P, _P]),
    "sa_abs_degree_f64": (c_int32, [POINTER(SaCsr), _P, _P, _P]),
    "sa_normalize_f32": (c_int32, [c_int64, _P, c_int64, _P, _P, c_double, _P]),
    "sa_normalize_f64": (c_int32, [c_int64, _P, c_int64, _P, _P, c_double, _P]),
    "sa_csr_tiles": (c_int32, [c_int64, _P, c_int32, _P, c_int64, _P]),
    "sa_panel_count": (c_int32, [POINTER(SaCsr), c_int32, _P, _P, _P]),
    "sa_panel_scatter_f32": (c_int32, [POINTER(SaCsr), c_int32, _P, _P, _P, _P, _P]),
    "sa_panel_scatter_f64": (c_int32, [POINTER(SaCsr), c_int32, _P, _P, _P, _P, _P]),
    "sa_panel_tiles": (c_int32, [c_int64, _P, c_int32, _P, c_int64, _P]),
    "sa_csr_bytes": (c_size_t, []),
    "sa_block_plan": (c_int32, [POINTER(SaCsr), _P, c_int32, c_int32, c_int32, c_int32, _P, _P, _P, _P]),
    "sa_block_fill_f32": (c_int32, [POINTER(SaCsr), _P, c_int32, c_int32, c_int32, c_int32, _P, _P, c_int64, _P, _P,
                                    c_double, _P, _P]),
    "sa_block_fill_f64": (c_int32, [POINTER(SaCsr), _P, c_int32, c_int32, c_int32, c_int32, _P, _P, c_int64, _P, _P,
                                    c_double, _P, _P]),
    "sa_cstream_sort_bytes": (c_int32, [c_int64, POINTER(c_size_t)]),
    "sa_cstream_max_rows": (c_int32, []),
    "sa_cstream_pad_row": (c_int32, []),
    "sa_cstream_plan": (c_int32, [POINTER(SaCsr), _P, c_int32, c_int32, _P, _P, _P, _P, _P, c_size_t, _P, _P, _P]),
    "sa_cstream_fill_f32": (c_int32, [POINTER(SaCsr), _P, c_int32, c_int32, _P, _P, _P, _P, c_int64, _P, _P, _P,
                                      c_double, _P, _P]),
    "sa_pblock_cols": (c_int32, []),
    "sa_pblock_plan": (c_int32, [POINTER(SaCsr), c_int32, c_int32, _P, c_int32, _P, _P, _P, _P, _P, _P, c_size_t,
                                 _P, _P, _P]),
    "sa_pblock_fill_f32": (c_int32, [POINTER(SaCsr), c_int32, c_int32, c_int32, _P, _P, _P, _P, _P, _P, _P,
                                     c_double, _P, _P]),
    "sa_alpha_scaling_f32": (c_int32, [c_int64, _P, c_double, c_double, _P, _P, _P]),
    "sa_alpha_scaling_f64": (c_int32, [c_int64, _P, c_double, c_double, _P, _P, _P]),
    "sa_spmv_scaled_f32": (c_int32, [POINTER(SaCsr), _P, _P, c_double, c_double, _P, _P, _P]),
    "sa_spmv_scaled_f64": (c_int32, [POINTER(SaCsr), _P, _P, c_double, c_double, _P, _P, _P]),
    "sa_spmv_scaled_static_f32": (c_int32, [POINTER(SaCsr), _P, _P, c_double, c_double, _P, _P]),
    "sa_spmv_gather_f32": (c_int32, [POINTER(SaCsr), _P, _P, c_double, _P, c_double, _P, _P]),
    "sa_spmv_gather_f64": (c_int32, [POINTER(SaCsr), _P, _P, c_double, _P, c_double, _P, _P]),
    "sa_lanczos_begin": (c_int32, [_P, c_int32, c_int32, c_double, c_double, c_double, c_int32, c_double, _P]),
    "sa_lanczos_thick_restart_f32": (c_int32, [_P, c_int64, _P, _P, c_int64, _P, _P, _P]),
    "sa_lanczos_thick_restart_f64": (c_int32, [_P, c_int64, _P, _P, c_int64, _P, _P, _P]),
    "sa_lanczos_start_f32": (c_int32, [_P, c_int64, _P, _P, _P, c_double, c_double, _P, c_double, _P, c_int64, _P,
                                       _P, c_int32, c_int32, _P]),
    "sa_lanczos_start_f64": (c_int32, [_P, c_int64, _P, _P, _P, c_double, c_double, _P, c_double, _P, c_int64, _P,
                                       _P, c_int32, c_int32, _P]),
    "sa_lanczos_step_f32": (c_int32, [POINTER(SaCsr), _P, _P, _P, c_int64, _P, _P, _P, _P]),
    "sa_lanczos_step_f64": (c_int32, [POINTER(SaCsr), _P, _P, _P, c_int64, _P, _P, _P, _P]),
    "sa_lanczos_ritz_f32": (c_int32, [_P, c_int64, _P, c_int64, _P, _P, c_int32, _P]),
    "sa_lanczos_ritz_f64": (c_int32, [_P, c_int64, _P, c_int64, _P, _P, c_int32, _P]),
    "sa_norm_dot_f32": (c_int32, [_P, c_int64, _P, _P, c_double, c_double, c_int32, _P]),
    "sa_norm_dot_f64": (c_int32, [_P, c_int64, _P, _P, c_double, c_double, c_int32, _P]),
    "sa_scale_copy_f32": (c_int32, [c_int64, _P, c_double, _P, _P]),
    "sa_scale_copy_f64": (c_int32, [c_int64, _P, c_double, _P, _P]),
    "sa_fill_normal_f32": (c_int32, [c_int64, c_uint64, c_uint64, _P, _P]),
    "sa_fill_normal_f64": (c_int32, [c_int64, c_uint64, c_uint64, _P, _P]),
    "sa_round_f32": (c_int32, [_P, c_int64, _P, c_double, _P, _P, _P]),
    "sa_round_f64": (c_int32, [_P, c_int64, _P, c_double, _P, _P, _P]),
    "sa_bits_compare": (c_int32, [_P, c_int64, _P, _P, _P]),
    "sa_energy_bits_i64": (c_int32, [POINTER(SaCsr), c_int32, _P, c_int64, _P, _P]),
    "sa_bits_transpose": (c_int32, [c_int64, c_int32, _P, c_int64, c_int32, _P, _P]),
    "sa_ls_init": (c_int32, [POINTER(SaCsr), c_int32, c_int32, _P, _P, _P, _P, _P]),
    "sa_ls_class_pass": (c_int32, [POINTER(SaCsr), c_int32, c_int32, _P, c_int64, c_uint64, c_int64, c_double, _P,
                                   _P, _P, _P]),
    "sa_ls_search_ctl_ints": (c_size_t, [c_int32, c_int32]),
    "sa_ls_search": (c_int32, [POINTER(SaCsr), c_int32, c_int32, _P, _P, c_int32, c_uint64, c_int32, c_double,
                               c_int32, c_int32, c_double, _P, _P, _P, _P]),
    "sa_ls_search_f64": (c_int32, [POINTER(SaCsr), _P, c_int32, c_int32, _P, _P, c_int32, c_uint64, c_int32,
                                   c_double, c_int32, c_int32, c_double, _P, _P, _P, _P]),
    "sa_ls_energy": (c_int32, [c_int64, c_int32, c_int32, _P, _P, _P, _P]),
    "sa_ls_init_f64": (c_int32, [POINTER(SaCsr), _P, c_int32, c_int32, _P, _P, _P, _P]),
    "sa_ls_class_pass_f64": (c_int32, [POINTER(SaCsr), _P, c_int32, c_int32, _P, c_int64, c_uint64, c_int64, c_double,
                                       _P, _P, _P, _P]),
    "sa_ls_energy_f64_scratch": (c_size_t, [c_int32]),
    "sa_ls_energy_f64": (c_int32, [c_int64, c_int32, c_int32, _P, _P, _P, _P, _P]),
    "sa_ls_extract_f32": (c_int32, [c_int64, c_int32, c_int32, _P, _P, _P]),
    "sa_ls_extract_f64": (c_int32, [c_int64, c_int32, c_int32, _P, _P, _P]),
    "sa_colour_start": (c_int32, [POINTER(SaCsr), _P, _P, _P, _P]),
    "sa_colour_rounds": (c_int32, [POINTER(SaCsr), c_uint64, _P, _P, _P, c_int32, c_int32, _P]),
    "sa_colour_rounds_lf": (c_int32, [POINTER(SaCsr), _P, _P, _P, c_int32, c_int32, _P]),
    "sa_guard_propose": (c_int32, [POINTER(SaCsr), c_int32, c_int32, _P, _P, c_uint64, c_uint64, _P, _P, _P, _P,
                                   _P]),
    "sa_guard_propose_f64": (c_int32, [POINTER(SaCsr), _P, c_int32, c_int32, _P, _P, c_uint64, c_uint64, _P, _P,
                                       _P, _P, _P]),
    "sa_guard_commit": (c_int32, [POINTER(SaCsr), c_int32, _P, _P, _P, _P, _P]),
    "sa_guard_commit_f64": (c_int32, [POINTER(SaCsr), _P, c_int32, _P, _P, _P, _P, _P]),
    "sa_guard_class_step": (c_int32, [POINTER(SaCsr), c_int32, _P, c_int64, _P, _P, _P, _P, _P, _P, _P]),
    "sa_guard_class_step_f64": (c_int32, [POINTER(SaCsr), _P, c_int32, _P, c_int64, _P, _P, _P, _P, _P, _P, _P]),
    "sa_guard_class_prep": (c_int32, [c_int32, c_double, c_double, c_double, _P, _P, _P, _P, _P, _P]),
    "sa_guard_class_prep_f64": (c_int32, [c_int32, c_double, c_double, _P, _P, _P, _P, _P, _P]),
    "sa_guard_sequential": (c_int32, [POINTER(SaCsr), c_int32, c_int32, _P, _P, _P, c_double, c_double, c_double,
                                      c_int64, _P, _P, _P]),
    "sa_guard_sequential_f64": (c_int32, [POINTER(SaCsr), _P, c_int32, c_int32, _P, _P, _P, c_double, c_double,
                                          c_int64, _P, _P, _P]),
    "sa_diag_scale_f32": (c_int32, [c_int64, _P, _P, c_double, _P, _P]),
    "sa_diag_scale_f64": (c_int32, [c_int64, _P, _P, c_double, _P, _P]),
    "sa_pack_spins": (c_int32, [c_int64, c_int32, _P, c_int64, _P, _P]),
    "sa_combine3_f32": (c_int32, [c_int64, c_double, _P, c_double, _P, c_double, _P, _P, _P]),
    "sa_combine3_f64": (c_int32, [c_int64, c_double, _P, c_double, _P, c_double, _P, _P, _P]),
    "sa_minres_state_bytes": (c_size_t, []),
    "sa_panel_sign_fold_f32": (c_int32, [c_int64, _P, _P, c_double, _P, _P]),
    "sa_panel_sign_fold_f64": (c_int32, [c_int64, _P, _P, c_double, _P, _P]),
    "sa_normalize_rows_f32": (c_int32, [POINTER(SaCsr), _P, _P, _P, _P, _P]),
    "sa_normalize_rows_f64": (c_int32, [POINTER(SaCsr), _P, _P, _P, _P, _P]),
    "sa_batch_alpha_scaling_f32": (c_int32, [POINTER(SaBatch), _P, c_double, c_double, _P, _P]),
    "sa_batch_alpha_scaling_f64": (c_int32, [POINTER(SaBatch), _P, c_double, c_double, _P, _P]),
    "sa_batch_alpha_scaling_each_f32": (c_int32, [POINTER(SaBatch), _P, _P, c_double, _P, _P]),
    "sa_batch_alpha_scaling_each_f64": (c_int32, [POINTER(SaBatch), _P, _P, c_double, _P, _P]),
    "sa_batch_copy_f32": (c_int32, [POINTER(SaBatch), _P, _P, _P]),
    "sa_batch_copy_f64": (c_int32, [POINTER(SaBatch), _P, _P, _P]),
    "sa_batch_lanczos_begin": (c_int32, [POINTER(SaBatch), c_int32, c_int32, c_double, _P, c_double, c_int32, c_double,
                                         _P]),
    "sa_batch_lanczos_start_f32": (c_int32, [POINTER(SaBatch), _P, _P, _P, c_double, c_double, _P, c_double, _P,
                                             c_int64, _P, c_int32, c_int32, _P]),
    "sa_batch_lanczos_start_f64": (c_int32, [POINTER(SaBatch), _P, _P, _P, c_double, c_double, _P, c_double, _P,
                                             c_int64, _P, c_int32, c_int32, _P]),
    "sa_batch_lanczos_start_each_f32": (c_int32, [POINTER(SaBatch), _P, _P, _P, _P, _P, c_int64, _P, c_int32, _P]),
    "sa_batch_lanczos_start_each_f64": (c_int32, [POINTER(SaBatch), _P, _P, _P, _P, _P, c_int64, _P, c_int32, _P]),
    "sa_batch_lanczos_step_f32": (c_int32, [POINTER(SaCsr), POINTER(SaBatch), _P, _P, c_int64, _P, _P, _P]),
    "sa_batch_lanczos_step_f64": (c_int32, [POINTER(SaCsr), POINTER(SaBatch), _P, _P, c_int64, _P, _P, _P]),
    "sa_batch_lanczos_thick_restart_f32": (c_int32, [POINTER(SaBatch), _P, _P, c_int64, _P, _P]),
    "sa_batch_lanczos_thick_restart_f64": (c_int32, [POINTER(SaBatch), _P, _P, c_int64, _P, _P]),
    "sa_batch_lanczos_ritz_f32": (c_int32, [POINTER(SaBatch), _P, c_int64, _P, _P, c_int32, _P]),
    "sa_batch_lanczos_ritz_f64": (c_int32, [POINTER(SaBatch), _P, c_int64, _P, _P, c_int32, _P]),
    "sa_batch_round_f32": (c_int32, [POINTER(SaBatch), _P, _P, _P, c_int32, _P]),
    "sa_batch_round_f64": (c_int32, [POINTER(SaBatch), _P, _P, _P, c_int32, _P]),
    "sa_batch_energy_bits_i64": (c_int32, [POINTER(SaCsr), POINTER(SaBatch), c_int32, _P, c_int64, _P, _P]),
    "sa_batch_ls_energy": (c_int32, [POINTER(SaBatch), c_int32, c_int32, _P, _P, _P, _P]),
    "sa_bits_gather": (c_int32, [c_int64, c_int32, c_int32, _P, _P, _P, _P, _P]),
    "sa_tangent_rhs_f32": (c_int32, [c_int64, _P, _P, c_double, _P, _P]),
    "sa_tangent_rhs_f64": (c_int32, [c_int64, _P, _P, c_double, _P, _P]),
    "sa_minres_begin_f32": (c_int32, [_P, c_int64, _P, _P, _P, c_double, c_double, c_int32,
                                      POINTER(SaMinresBufs), _P]),
    "sa_minres_begin_f64": (c_int32, [_P, c_int64, _P, _P, _P, c_double, c_double, c_int32,
                                      POINTER(SaMinresBufs), _P]),
    "sa_minres_iterate_f32": (c_int32, [POINTER(SaCsr), _P, _P, _P, c_int32, POINTER(SaMinresBufs), _P]),
    "sa_minres_iterate_f64": (c_int32, [POINTER(SaCsr), _P, _P, _P, c_int32, POINTER(SaMinresBufs), _P]),
    "sa_minres_end_f32": (c_int32, [_P, c_int64, _P, POINTER(SaMinresBufs), _P]),
    "sa_minres_end_f64": (c_int32, [_P, c_int64, _P, POINTER(SaMinresBufs), _P]),
}

_SIGNATURES.update({
    "sa_graph_erdos_renyi": (c_int32, [c_int64, c_double, c_uint64, c_int32, c_int32, POINTER(c_void_p)]),
    "sa_graph_barabasi_albert": (c_int32, [c_int64, c_int32, c_uint64, c_int32, c_int32, POINTER(c_void_p)]),
    "sa_graph_nnz": (c_int64, [c_void_p]),
    "sa_graph_n": (c_int64, [c_void_p]),
    "sa_graph_export": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "sa_graph_free": (None, [c_void_p]),
})

_lib = None


def load(path: Path | str | None = None) -> ctypes.CDLL:
    """Load libspecann once; raise SpecAnnLibraryError if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise SpecAnnLibraryError(
            f"{p} is missing: build the HIP library first (python -m specann_amd.build); "
            "the solver has no CPU fallback")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.sa_state_bytes() != ctypes.sizeof(SaState):
        raise SpecAnnLibraryError(
            f"sa_state layout mismatch: library {lib.sa_state_bytes()} bytes, binding {ctypes.sizeof(SaState)}")
    if lib.sa_csr_bytes() != ctypes.sizeof(SaCsr):
        raise SpecAnnLibraryError(
            f"sa_csr layout mismatch: library {lib.sa_csr_bytes()} bytes, binding {ctypes.sizeof(SaCsr)}")
    if lib.sa_minres_state_bytes() != ctypes.sizeof(SaMinresState):
        raise SpecAnnLibraryError("sa_minres_state layout mismatch between the library and the binding")
    if path is None:
        _lib = lib
    return lib


def exported_names() -> list[str]:
    return list(_SIGNATURES)


def last_error() -> str:
    lib = load()
    buf = ctypes.create_string_buffer(1024)
    lib.sa_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def call(name: str, *args) -> int:
    """Call an entry point; map a negative status to an exception."""
    rc = getattr(load(), name)(*args)
    if rc >= 0:
        return rc
    msg = f"{name} failed ({rc}): {last_error()}"
    if rc == SA_E_OOM:
        raise DeviceOutOfMemory(msg)
    if rc == SA_E_BREAKDOWN:
        raise EigensolverBreakdown(msg)
    raise SpecAnnLibraryError(msg)
