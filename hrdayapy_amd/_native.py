"""ctypes binding of libhrdaya_hip.so (C ABI declared in include/hrdaya_hip.h).

The library is loaded lazily on first use.  There is no fallback: if the
shared library is missing or cannot be loaded, every native entry point
raises ``RuntimeError`` so that the GPU path can never silently run on some
other implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libhrdaya_hip.so"
_lib = None

ABI_VERSION = 6

c_i32, c_i64, c_f32, c_f64, c_vp = C.c_int32, C.c_int64, C.c_float, C.c_double, C.c_void_p


class HpBricks(C.Structure):
    _fields_ = [("n_bricks", c_i64), ("bx", c_i32), ("by", c_i32), ("bz", c_i32), ("pad_", c_i32),
                ("nbr", c_vp), ("label", c_vp), ("D6", c_vp), ("inst", c_vp), ("groups", c_vp)]


class HpMonoParams(C.Structure):
    _fields_ = [("dt", c_f64), ("inv_cm", c_f64), ("dscale", c_f64), ("ki_base", c_f64),
                ("nai_base", c_f64), ("vmin", c_f64), ("vmax", c_f64), ("at_thr", c_f64),
                ("do_reaction", c_i32), ("n_inst", c_i32), ("inst_params", c_vp),
                ("step_ctr", c_vp), ("group_amp", c_vp), ("at", c_vp), ("rate_ext", c_vp),
                ("cell_slot", c_vp), ("n_cells", c_i64)]


class HpCable(C.Structure):
    _fields_ = [("n", c_i64), ("indptr", c_vp), ("indices", c_vp), ("label", c_vp), ("w", c_f64)]


class HpPmj(C.Structure):
    _fields_ = [("n_junctions", c_i64), ("n_entries", c_i64), ("pk", c_vp), ("ptr", c_vp),
                ("patch", c_vp), ("n_targets", c_i64), ("tgt_slot", c_vp), ("tgt_ptr", c_vp),
                ("tgt_entry", c_vp), ("c_pmj", c_f64), ("ai_p", c_f64), ("ai_m", c_f64),
                ("cm", c_f64), ("coef", c_vp), ("exclusive", c_i32), ("pad_", c_i32)]


class HpPmjAt(C.Structure):
    _fields_ = [("at", c_vp), ("V_prev", c_vp), ("step_ctr", c_vp), ("dt", c_f64),
                ("at_thr", c_f64)]


class HpCnPmj(C.Structure):
    _fields_ = [("n_targets", c_i64), ("tgt_slot", c_vp), ("tgt_ptr", c_vp), ("tgt_node", c_vp),
                ("tgt_w", c_vp), ("node_ptr", c_vp), ("node_slot", c_vp), ("node_w", c_vp),
                ("n_jnodes", c_i64), ("jnode", c_vp)]


class HpCnParams(C.Structure):
    _fields_ = [("cm_dt", c_f64), ("theta", c_f64), ("dscale", c_f64), ("dscale_pk", c_f64),
                ("tol", c_f64), ("max_sweeps", c_i32), ("solver", c_i32), ("scalars", c_vp),
                ("part", c_vp), ("part_len", c_i64), ("grid_cap", c_i64)]


class HpLapCg(C.Structure):
    _fields_ = [("x", c_vp), ("r", c_vp), ("p", c_vp * 2), ("ap", c_vp), ("scalars", c_vp),
                ("part", c_vp), ("part_len", c_i64), ("tol", c_f64), ("launched", c_i64)]


class HpCnWork(C.Structure):
    _fields_ = [("Vs", c_vp), ("b", c_vp), ("x", c_vp), ("Vs_pk", c_vp), ("b_pk", c_vp),
                ("x_pk", c_vp), ("kr", c_vp * 8), ("kr_pk", c_vp * 8), ("warm", c_vp * 4)]


class HpMlpWeights(C.Structure):
    _fields_ = [("w1", c_vp), ("w2", c_vp), ("w3", c_vp), ("w4", c_vp), ("b1", c_vp), ("b2", c_vp),
                ("b3", c_vp), ("b4", c_vp)]


class HpMlpNorm(C.Structure):
    _fields_ = [("in_mean", c_f32 * 22), ("in_rstd", c_f32 * 22), ("in_log", c_i32 * 22),
                ("out_mean", c_f32 * 19), ("out_std", c_f32 * 19)]


class HpSurrogateParams(C.Structure):
    _fields_ = [("t_macro", c_f32), ("dvdt_max", c_f32), ("v_lo", c_f32), ("v_hi", c_f32),
                ("ki_base", c_f32), ("nai_base", c_f32), ("force_fallback", c_i32), ("variant", c_i32),
                ("fb_list", c_vp), ("fb_count", c_vp), ("slot_list", c_vp), ("n_list", c_i64),
                ("state_list", c_vp), ("correct", c_i32), ("rest_v", c_f32), ("rest_rate", c_f32),
                ("corr_max", c_f32)]


class HpMgLevel(C.Structure):
    _fields_ = [("ni", c_i32), ("nj", c_i32), ("nk", c_i32), ("pad_", c_i32), ("n", c_i64),
                ("gx", c_vp), ("gy", c_vp), ("gz", c_vp), ("inv_diag", c_vp), ("x", c_vp),
                ("tmp", c_vp), ("b", c_vp)]


MG_MAX_LEVELS = 12


class HpMg(C.Structure):
    _fields_ = [("n_levels", c_i32), ("nu", c_i32), ("omega", c_f32), ("pad_", c_i32),
                ("inv_diag0", c_vp), ("tmp0", c_vp), ("coarse_inv", c_vp),
                ("level", HpMgLevel * MG_MAX_LEVELS)]


class HpTorso(C.Structure):
    _fields_ = [("ni", c_i32), ("nj", c_i32), ("nk", c_i32), ("pad_", c_i32), ("n_cells", c_i64),
                ("col_k0", c_vp), ("col_len", c_vp), ("col_off", c_vp), ("gx", c_vp), ("gy", c_vp),
                ("gz", c_vp), ("diag", c_vp), ("pin", c_i64), ("chunks", c_vp),
                ("n_chunks", c_i64)]


class HpPcgWs(C.Structure):
    _fields_ = [("part", c_vp), ("counter", c_vp), ("out", c_vp), ("scalars", c_vp)]


P = C.POINTER
_SIGS = {
    "hp_abi_version": (C.c_int, []),
    "hp_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "hp_device_alloc": (C.c_int, [C.POINTER(c_vp), c_i64, C.c_int]),
    "hp_device_free": (C.c_int, [c_vp]),
    "hp_crossing_events_f32": (C.c_int, [c_vp, c_vp, c_i64, c_i64, c_f64, C.c_int, c_f64, c_f64,
                                         c_vp, c_vp, c_i32, c_vp]),
    "hp_ttp06_step_f32": (C.c_int, [c_vp, c_vp, c_vp, c_vp, C.c_int, c_vp, c_vp, c_i64, c_i64,
                                    c_i64, c_f32, c_f32, c_f32, c_vp]),
    "hp_ttp06_step_f64": (C.c_int, [c_vp, c_vp, c_vp, c_vp, C.c_int, c_vp, c_vp, c_i64, c_i64,
                                    c_i64, c_f64, c_f64, c_f64, c_vp]),
    "hp_ttp06_step_host_f32": (C.c_int, [c_vp, c_vp, c_vp, c_vp, C.c_int, c_vp, c_vp, c_i64,
                                         c_i64, c_i64, c_f32, c_f32, c_f32]),
    "hp_ttp06_step_host_f64": (C.c_int, [c_vp, c_vp, c_vp, c_vp, C.c_int, c_vp, c_vp, c_i64,
                                         c_i64, c_i64, c_f64, c_f64, c_f64]),
    "hp_myo_step_f32": (C.c_int, [P(HpBricks), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams), c_vp]),
    "hp_myo_step_f64": (C.c_int, [P(HpBricks), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams), c_vp]),
    "hp_myo_cable_step_f32": (C.c_int, [P(HpBricks), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams),
                                         P(HpCable), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams),
                                         c_vp]),
    "hp_myo_cable_step_f64": (C.c_int, [P(HpBricks), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams),
                                         P(HpCable), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams),
                                         c_vp]),
    "hp_cn_step_f32": (C.c_int, [P(HpBricks), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams),
                                 P(HpCable), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams), P(HpCnPmj),
                                 P(HpCnParams), P(HpCnWork), c_vp]),
    "hp_cn_sweep_f32": (C.c_int, [P(HpBricks), c_vp, c_vp, c_vp, P(HpCnParams), c_vp]),
    "hp_cn_apply_f32": (C.c_int, [P(HpBricks), P(HpCnParams), P(HpCnWork), c_vp]),
    "hp_laplace_cg_begin_f32": (C.c_int, [P(HpBricks), P(HpLapCg), c_vp]),
    "hp_laplace_cg_iterate_f32": (C.c_int, [P(HpBricks), P(HpLapCg), c_i32, c_i64, c_vp]),
    "hp_fibre_rule_f32": (C.c_int, [c_vp, c_vp, c_i32, c_i32, c_i32, c_f64, c_f64, c_f64, c_f64,
                                    c_f64, c_vp, c_vp]),
    "hp_cn_step_f64": (C.c_int, [P(HpBricks), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams),
                                 P(HpCable), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams), P(HpCnPmj),
                                 P(HpCnParams), P(HpCnWork), c_vp]),
    "hp_fallback_react_f32": (C.c_int, [P(HpBricks), c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64,
                                        P(HpMonoParams), c_vp]),
    "hp_stencil19_apply_f32": (C.c_int, [P(HpBricks), c_vp, c_vp, c_f64, c_vp]),
    "hp_stencil19_apply_f64": (C.c_int, [P(HpBricks), c_vp, c_vp, c_f64, c_vp]),
    "hp_cable_step_f32": (C.c_int, [P(HpCable), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams), c_vp]),
    "hp_cable_step_f64": (C.c_int, [P(HpCable), c_vp, c_vp, c_vp, c_i64, P(HpMonoParams), c_vp]),
    "hp_pmj_couple_f32": (C.c_int, [P(HpPmj), c_vp, c_vp, c_vp, c_f64, c_f64, c_f64, P(HpPmjAt),
                                    c_vp]),
    "hp_pmj_couple_f64": (C.c_int, [P(HpPmj), c_vp, c_vp, c_vp, c_f64, c_f64, c_f64, P(HpPmjAt),
                                    c_vp]),
    "hp_stim_eval_f32": (C.c_int, [c_vp, c_i32, c_vp, c_i32, c_vp, c_f64, c_vp, c_vp]),
    "hp_stim_eval_f64": (C.c_int, [c_vp, c_i32, c_vp, c_i32, c_vp, c_f64, c_vp, c_vp]),
    "hp_step_advance_f32": (C.c_int, [c_vp, c_i32, c_vp, c_i32, c_vp, c_f64, c_vp, c_vp]),
    "hp_step_advance_f64": (C.c_int, [c_vp, c_i32, c_vp, c_i32, c_vp, c_f64, c_vp, c_vp]),
    "hp_mg_vcycle_f32": (C.c_int, [P(HpTorso), P(HpMg), c_vp, c_vp, c_vp, c_vp]),
    "hp_mgpcg_begin_f32": (C.c_int, [P(HpTorso), P(HpMg), c_vp, c_vp, c_vp, c_vp, P(HpPcgWs),
                                     c_f64, C.c_int, c_vp]),
    "hp_mgpcg_iterate_f32": (C.c_int, [P(HpTorso), P(HpMg), c_vp, c_vp, c_vp, c_vp, c_vp,
                                       P(HpPcgWs), C.c_int, c_vp]),
    "hp_torso_chunks": (C.c_int, [c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "hp_poisson_spmv_f32": (C.c_int, [P(HpTorso), c_vp, c_vp, c_vp]),
    "hp_poisson_residual_f64": (C.c_int, [P(HpTorso), c_vp, c_vp, c_vp, P(HpPcgWs), c_vp]),
    "hp_poisson_residual_f64_noalloc": (C.c_int, [P(HpTorso), c_vp, c_vp, c_vp, c_vp, c_vp]),
    "hp_pcg_begin_f32": (C.c_int, [P(HpTorso), c_vp, c_vp, c_vp, P(HpPcgWs), c_f64, C.c_int,
                                   C.c_int, c_vp]),
    "hp_pcg_begin_metric_f32": (C.c_int, [P(HpTorso), c_vp, c_vp, c_vp, P(HpPcgWs), c_f64,
                                          C.c_int, C.c_int, c_i64, c_vp]),
    "hp_pcg_iterate_f32": (C.c_int, [P(HpTorso), c_vp, c_vp, c_vp, c_vp, P(HpPcgWs), C.c_int,
                                     c_vp]),
    "hp_refine_axpy_f64": (C.c_int, [P(HpTorso), c_vp, c_vp, c_f64, c_vp]),
    "hp_cast_scale_f32": (C.c_int, [P(HpTorso), c_vp, c_vp, c_f64, c_vp]),
    "hp_heart_lap6_f64": (C.c_int, [P(HpBricks), c_vp, c_vp, c_f64, c_vp]),
    "hp_leadfield_apply_f32": (C.c_int, [c_vp, c_vp, c_i64, C.c_int, c_vp, c_i64, C.c_int, c_vp,
                                         c_i64, c_vp, c_vp]),
    "hp_heart_lap6_f32": (C.c_int, [P(HpBricks), c_vp, c_vp, c_f64, c_vp]),
    "hp_bspm_source_f64": (C.c_int, [c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_f64, c_f64,
                                     c_vp, c_vp, c_i64, P(HpPcgWs), c_vp]),
    "hp_leads_accumulate_f64": (C.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "hp_gemm_bf16": (C.c_int, [C.c_int, C.c_int, C.c_int, c_vp, c_i64, c_i64, c_vp, c_i64, c_i64,
                               c_vp, c_i64, c_f32, c_f32, c_vp, C.c_int, c_vp]),
    "hp_silu_bf16": (C.c_int, [c_vp, c_vp, c_i64, C.c_int, C.c_int, C.c_int, c_vp]),
    "hp_silu_bwd_bf16": (C.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "hp_cast_bf16": (C.c_int, [c_vp, c_vp, c_i64, c_vp]),
    "hp_mse_grad": (C.c_int, [c_vp, c_vp, c_vp, c_vp, C.c_int, C.c_int, C.c_int, c_vp, c_vp]),
    "hp_colsum": (C.c_int, [c_vp, C.c_int, C.c_int, C.c_int, c_vp, c_vp]),
    "hp_adam": (C.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_f32, c_f32, c_f32, c_f32, C.c_int,
                          c_f32, c_vp]),
    "hp_surrogate_step": (C.c_int, [P(HpBricks), c_vp, c_vp, c_vp, c_i64, c_vp, P(HpMlpWeights),
                                    P(HpMlpNorm), P(HpSurrogateParams), C.c_int, c_vp]),
    "hp_step_tail_f32": (C.c_int, [P(HpPmj), c_vp, c_vp, c_vp, c_f64, c_f64, c_f64, P(HpPmjAt),
                                   c_vp, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "hp_step_tail_f64": (C.c_int, [P(HpPmj), c_vp, c_vp, c_vp, c_f64, c_f64, c_f64, P(HpPmjAt),
                                   c_vp, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "hp_bigstep_reaction": (C.c_int, [P(HpBricks), c_vp, c_vp, c_vp, c_i64, c_vp,
                                      P(HpSurrogateParams), c_vp]),
    "hp_gather_f32": (C.c_int, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "hp_gather_f64": (C.c_int, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "hp_stream_copy_f32": (C.c_int, [c_vp, c_vp, c_i64, c_vp]),
    "hp_register_dice_f64": (C.c_int, [c_vp, c_i64, c_vp, c_i64, c_vp, c_i32, c_vp, c_vp, c_vp,
                                       c_vp, c_f64, c_vp, c_i64, c_vp, c_vp]),
    "hp_project_rows_f64": (C.c_int, [c_vp, c_i64, c_vp, c_f64, c_vp, c_f64, c_vp, c_i32, c_i32,
                                      c_i32, c_vp, c_vp]),
    "hp_feature_transform": (C.c_int, [c_vp, C.c_uint8, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp,
                                       c_vp, c_i64, c_vp]),
    "hp_sbm_psi_f64": (C.c_int, [c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_f64, c_vp, c_vp]),
    "hp_zoom_linear_f64": (C.c_int, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_f64, c_vp, c_vp,
                                     c_vp]),
    "hp_zoom_nearest_i32": (C.c_int, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp]),
    "hp_label_sigma_f64": (C.c_int, [c_vp, c_i64, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "hp_extend_sigma_f64": (C.c_int, [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp]),
    "hp_column_stats": (C.c_int, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp]),
    "hp_scan_i32": (C.c_int, [c_vp, c_i64, c_vp, c_vp]),
    "hp_column_emit": (C.c_int, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                 c_vp]),
    "hp_surface_flags": (C.c_int, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp]),
    "hp_select_rank": (C.c_int, [c_vp, c_vp, c_i64, c_i32, c_i64, c_vp, c_vp]),
    "hp_torso_faces_f64": (C.c_int, [c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_f64, c_vp, c_vp,
                                     c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp,
                                     c_vp]),
}


def lib_path() -> Path:
    return _LIB_PATH


def load():
    """Load the shared library (once) and declare every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise RuntimeError(
            f"libhrdaya_hip.so not found at {_LIB_PATH}; build it with "
            "`python -m hrdayapy_amd.build` (there is no non-HIP fallback)")
    # torch's HIP runtime must be the one the library binds to
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(_LIB_PATH), mode=os.RTLD_GLOBAL if hasattr(os, "RTLD_GLOBAL") else 0)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    v = lib.hp_abi_version()
    if v != ABI_VERSION:
        raise RuntimeError(f"libhrdaya_hip ABI {v} != expected {ABI_VERSION}; rebuild")
    _lib = lib
    return lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    load().hp_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def call(name: str, *args) -> None:
    """Call an entry point and raise RuntimeError on a non-zero return code."""
    rc = getattr(load(), name)(*args)
    if rc != 0:
        raise RuntimeError(f"{name} failed ({rc}): {last_error()}")


def exported_symbols() -> list[str]:
    return list(_SIGS)


def torso_chunks(shape, col_k0, col_len, col_off):
    """The chunk table of a column-run torso grid (hp_torso_chunks): int32
    array [n_chunks, 16], built on the host."""
    import numpy as np
    ni, nj, nk = (int(v) for v in shape)
    a = [np.ascontiguousarray(v, dtype=np.int32) for v in (col_k0, col_len, col_off)]
    n = C.c_int64(0)
    call("hp_torso_chunks", ni, nj, nk, *(v.ctypes.data for v in a), None, 0, C.byref(n))
    tab = np.zeros((max(n.value, 1), 16), dtype=np.int32)
    call("hp_torso_chunks", ni, nj, nk, *(v.ctypes.data for v in a), tab.ctypes.data,
         n.value, C.byref(n))
    return tab[:n.value]


def ptr(t) -> int:
    """Device pointer of a torch tensor, or None (-> NULL) for None."""
    return None if t is None else t.data_ptr()


def stream_handle(device=None) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream


def sfx(dtype) -> str:
    import torch
    if dtype == torch.float32:
        return "f32"
    if dtype == torch.float64:
        return "f64"
    raise ValueError(f"unsupported dtype {dtype}; the HIP kernels take float32 or float64")


class DeviceArena:
    """One device allocation carved into typed views, each 2 MiB aligned.

    ``specs``: list of (name, shape, dtype).  With ``contiguous`` the memory
    comes from hp_device_alloc(..., contiguous=1) (physically contiguous pages)
    and is freed when the arena is collected -- the views must not outlive it
    (the engine that owns the arena also owns the views); otherwise it is one
    torch allocation.  Views are available as attributes / ``arena[name]``.
    """

    ALIGN = 2 << 20

    def __init__(self, specs, device, contiguous=False):
        import torch
        self._ptr = None
        offs, total = [], 0
        for _, shape, dtype in specs:
            total = -(-total // self.ALIGN) * self.ALIGN
            offs.append(total)
            n = 1
            for d in (shape if isinstance(shape, tuple) else (shape,)):
                n *= int(d)
            total += n * torch.empty((), dtype=dtype).element_size()
        total = max(total, 1)
        self.nbytes = total
        if contiguous:
            load()
            p = c_vp()
            call("hp_device_alloc", C.byref(p), total, 1)
            self._ptr = p.value

            class _CAI:
                __cuda_array_interface__ = {"shape": (total,), "typestr": "|u1",
                                            "data": (p.value, False), "version": 2,
                                            "strides": None}
            buf = torch.as_tensor(_CAI(), device=device)
        else:
            buf = torch.empty(total, dtype=torch.uint8, device=device)
        self._buf = buf
        self._views = {}
        for (name, shape, dtype), o in zip(specs, offs):
            shape = shape if isinstance(shape, tuple) else (shape,)
            n = 1
            for d in shape:
                n *= int(d)
            es = torch.empty((), dtype=dtype).element_size()
            self._views[name] = buf[o:o + n * es].view(dtype).view(shape)

    def __getitem__(self, name):
        return self._views[name]

    def __del__(self):
        if self._ptr is not None:
            try:
                import torch
                torch.cuda.synchronize()
                call("hp_device_free", self._ptr)
            except Exception:
                pass
            self._ptr = None
