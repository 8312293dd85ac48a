"""CPU checks of the C-ABI's argument validation: bad arguments return -1 with a
message naming the entry point, and empty inputs return 0; every case here is
decided on the host before any HIP call (no GPU needed)."""
import ctypes as C

import numpy as np
import pytest

from hrdayapy_amd import _native as nat


@pytest.fixture(scope="module")
def lib():
    return nat.load()


def _rejects(rc, needle):
    assert rc == -1
    assert needle in nat.last_error()


def test_library_level_errors(lib):
    p = C.c_void_p()
    _rejects(lib.hp_device_alloc(C.byref(p), 0, 0), "hp_device_alloc")
    assert p.value is None
    assert lib.hp_last_error(C.create_string_buffer(4), 0) == -1


def test_ttp06_arguments(lib):
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data
    # no celltype array and an out-of-range fixed type
    _rejects(lib.hp_ttp06_step_f32(ptr, ptr, ptr, None, 3, None, None, 1, 1, 1, 0.02, 138.0, 7.0,
                                   None), "hp_ttp06_step")
    _rejects(lib.hp_ttp06_step_f32(ptr, ptr, ptr, None, 0, None, None, -1, 1, 1, 0.02, 138.0, 7.0,
                                   None), "hp_ttp06_step")
    # cells but no membrane potential or state
    _rejects(lib.hp_ttp06_step_f32(None, ptr, ptr, None, 0, None, None, 4, 1, 1, 0.02, 138.0, 7.0,
                                   None), "V and state required")
    _rejects(lib.hp_ttp06_step_f32(ptr, None, ptr, None, 0, None, None, 4, 1, 1, 0.02, 138.0, 7.0,
                                   None), "V and state required")
    # no cells: nothing to do
    assert lib.hp_ttp06_step_f32(None, None, None, None, 0, None, None, 0, 1, 1, 0.02, 138.0, 7.0,
                                 None) == 0


def test_brick_grid_arguments(lib):
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data
    g = nat.HpBricks(n_bricks=-1, bx=4, by=4, bz=4)
    _rejects(lib.hp_stencil19_apply_f32(C.byref(g), ptr, ptr, 1.0, None), "bad brick grid")
    _rejects(lib.hp_stencil19_apply_f32(None, ptr, ptr, 1.0, None), "bad brick grid")
    # an empty grid launches nothing
    g0 = nat.HpBricks(n_bricks=0, bx=4, by=4, bz=4)
    assert lib.hp_stencil19_apply_f32(C.byref(g0), None, None, 1.0, None) == 0
    # a grid whose six D rows overflow 32-bit byte offsets
    big = nat.HpBricks(n_bricks=(1 << 31) // 64, bx=4, by=4, bz=4)
    _rejects(lib.hp_stencil19_apply_f32(C.byref(big), ptr, ptr, 1.0, None), "32-bit offset")
    # the step needs its parameters (step counter and stimulus amplitudes)
    _rejects(lib.hp_myo_step_f32(C.byref(g0), ptr, ptr, ptr, 0, None, None),
             "step_ctr and group_amp are required")
    p = nat.HpMonoParams(do_reaction=3, step_ctr=ptr, group_amp=ptr)
    _rejects(lib.hp_myo_step_f32(C.byref(g0), ptr, ptr, ptr, 0, C.byref(p), None),
             "do_reaction must be 0, 1 or 2")
    # the reaction needs the compact-state group table
    p1 = nat.HpMonoParams(do_reaction=1, step_ctr=ptr, group_amp=ptr)
    g1 = nat.HpBricks(n_bricks=1, bx=4, by=4, bz=4)
    _rejects(lib.hp_myo_step_f32(C.byref(g1), ptr, ptr, ptr, 64, C.byref(p1), None),
             "needs groups and state")


def test_torso_and_leadfield_arguments(lib):
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data
    _rejects(lib.hp_poisson_spmv_f32(None, ptr, ptr, None), "bad torso grid")
    t = nat.HpTorso(ni=4, nj=4, nk=4, n_cells=64)  # no chunk table
    _rejects(lib.hp_poisson_spmv_f32(C.byref(t), ptr, ptr, None), "chunk table required")
    t.chunks, t.nk = ptr, 40000  # columns longer than the 15-bit span fields
    _rejects(lib.hp_poisson_spmv_f32(C.byref(t), ptr, ptr, None), "nk <= 32767")
    # lead field: at most 16 frames per call, and a partial buffer that holds one block
    part = np.zeros(16, np.float64)
    args = (ptr, ptr, 4, 2, ptr, 4)
    _rejects(lib.hp_leadfield_apply_f32(*args, 17, part.ctypes.data, 16, part.ctypes.data, None),
             "1..16 frames")
    _rejects(lib.hp_leadfield_apply_f32(*args, 0, part.ctypes.data, 16, part.ctypes.data, None),
             "1..16 frames")
    _rejects(lib.hp_leadfield_apply_f32(*args, 16, part.ctypes.data, 16, part.ctypes.data, None),
             "partial buffer needs 32 doubles")


def test_mlp_arguments(lib):
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data
    _rejects(lib.hp_gemm_bf16(0, 32, 32, ptr, 32, 1, ptr, 32, 1, ptr, 32, 1.0, 0.0, None, 1, None),
             "bad shape")
    _rejects(lib.hp_gemm_bf16(32, 32, 32, None, 32, 1, ptr, 32, 1, ptr, 32, 1.0, 0.0, None, 1,
                              None), "null operand")
    _rejects(lib.hp_gemm_bf16(32, 32, 32, ptr, 32, 1, ptr, 32, 1, ptr, 32, 1.0, 0.0, ptr, 2, None),
             "split-K")
    w, n, sp = nat.HpMlpWeights(), nat.HpMlpNorm(), nat.HpSurrogateParams()
    g = nat.HpBricks(n_bricks=1, bx=4, by=4, bz=4)  # no group table
    _rejects(lib.hp_surrogate_step(C.byref(g), ptr, ptr, ptr, 64, ptr, C.byref(w), C.byref(n),
                                   C.byref(sp), 0, None), "groups required")
    g.groups = ptr
    sp.slot_list, sp.n_list = ptr, 33  # a listed count that is not a multiple of 32
    _rejects(lib.hp_surrogate_step(C.byref(g), ptr, ptr, ptr, 64, ptr, C.byref(w), C.byref(n),
                                   C.byref(sp), 0, None), "multiple of 32")
    sp.n_list, sp.variant = 32, 3  # no such kernel shape
    _rejects(lib.hp_surrogate_step(C.byref(g), ptr, ptr, ptr, 64, ptr, C.byref(w), C.byref(n),
                                   C.byref(sp), 0, None), "variant")
    sp.variant = 2  # a state array (18 x ld floats) beyond 32-bit byte offsets
    _rejects(lib.hp_surrogate_step(C.byref(g), ptr, ptr, ptr, 60_000_000, ptr, C.byref(w),
                                   C.byref(n), C.byref(sp), 0, None), "3.75 GiB")


def test_small_kernels_arguments(lib):
    buf = np.zeros(64, np.float32)
    t = np.zeros(4, np.float64)
    ptr = buf.ctypes.data
    # events: a table or a count array is required; empty node sets do nothing
    _rejects(lib.hp_crossing_events_f32(None, t.ctypes.data, 4, 16, -60.0, 1, 0.0, 0.0, ptr, None,
                                        0, None), "hp_crossing_events_f32")
    _rejects(lib.hp_crossing_events_f32(ptr, t.ctypes.data, 4, 16, -60.0, 1, 0.0, 0.0, None,
                                        None, 0, None), "hp_crossing_events_f32")
    assert lib.hp_crossing_events_f32(ptr, t.ctypes.data, 4, 0, -60.0, 1, 0.0, 0.0, ptr, None, 0,
                                      None) == 0
    # copy: whole float4 vectors only; empty copies do nothing
    _rejects(lib.hp_stream_copy_f32(ptr, ptr, 3, None), "multiple of 4")
    assert lib.hp_stream_copy_f32(None, None, 0, None) == 0
    # gather: pointers required for a non-empty gather
    _rejects(lib.hp_gather_f32(None, ptr, ptr, 8, None), "hp_gather")
    _rejects(lib.hp_gather_f32(ptr, ptr, ptr, -1, None), "hp_gather")
    assert lib.hp_gather_f32(None, None, None, 0, None) == 0


def test_training_and_leads_helpers_arguments(lib):
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data
    _rejects(lib.hp_silu_bf16(None, ptr, 4, 4, 4, 4, None), "hp_silu_bf16")
    _rejects(lib.hp_silu_bwd_bf16(ptr, ptr, None, None, 8, None), "hp_silu_bwd_bf16")
    _rejects(lib.hp_cast_bf16(ptr, None, 8, None), "hp_cast_bf16")
    _rejects(lib.hp_mse_grad(ptr, ptr, ptr, None, 4, 4, 4, None, None), "hp_mse_grad")
    _rejects(lib.hp_mse_grad(ptr, ptr, ptr, None, 4, 4, 5, ptr, None), "used outside 1..cols")
    _rejects(lib.hp_colsum(ptr, 4, 4, 3, ptr, None), "ld < cols")
    _rejects(lib.hp_adam(ptr, None, ptr, ptr, 8, 1e-3, 0.9, 0.999, 1e-8, 1, 0.0, None), "hp_adam")
    _rejects(lib.hp_leads_accumulate_f64(None, ptr, ptr, ptr, ptr, 2, None),
             "hp_leads_accumulate_f64")
    # empty sizes do nothing, whatever the pointers
    assert lib.hp_silu_bf16(None, None, 0, 4, 4, 4, None) == 0
    assert lib.hp_cast_bf16(None, None, 0, None) == 0
    assert lib.hp_colsum(None, 0, 4, 4, None, None) == 0
    assert lib.hp_adam(None, None, None, None, 0, 1e-3, 0.9, 0.999, 1e-8, 1, 0.0, None) == 0
    assert lib.hp_leads_accumulate_f64(None, None, None, None, None, 0, None) == 0


def test_pcg_workspace_arguments(lib):
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data
    t = nat.HpTorso(ni=4, nj=4, nk=4, n_cells=64, chunks=ptr, n_chunks=1)
    _rejects(lib.hp_poisson_residual_f64(C.byref(t), ptr, ptr, ptr, None, None),
             "workspace (part and out) required")
    ws = nat.HpPcgWs(part=ptr, out=ptr)  # no PCG scalar slots
    _rejects(lib.hp_pcg_begin_f32(C.byref(t), ptr, ptr, ptr, C.byref(ws), 1e-8, 10, 0, None),
             "workspace (part and scalars) required")
    ws.scalars = ptr
    _rejects(lib.hp_pcg_iterate_f32(C.byref(t), ptr, ptr, ptr, ptr, C.byref(ws), -1, None),
             "n_iter >= 0")


def test_torso_build_and_registration_arguments(lib):
    """The device torso build, registration and projection reject null operands,
    empty or non-positive sizes and spacings on the host."""
    buf = np.zeros(64, np.float64)
    ptr = buf.ctypes.data
    sp3 = (C.c_double * 3)(0.75, 0.75, 1.0)
    bad3 = (C.c_double * 3)(0.75, 0.0, 1.0)
    shp = (C.c_int32 * 3)(4, 4, 4)
    _rejects(lib.hp_feature_transform(None, 0, 4, 4, 4, sp3, ptr, ptr, ptr, 64, None),
             "hp_feature_transform")
    _rejects(lib.hp_feature_transform(ptr, 0, 4, 4, 4, bad3, ptr, ptr, ptr, 64, None),
             "spacing must be positive")
    _rejects(lib.hp_feature_transform(ptr, 0, 4, 4, 9, sp3, ptr, ptr, ptr, 8, None),
             "holds no line")
    _rejects(lib.hp_sbm_psi_f64(ptr, ptr, ptr, 4, 4, 4, sp3, 0.0, ptr, None), "hp_sbm_psi_f64")
    _rejects(lib.hp_zoom_linear_f64(ptr, 4, 4, 4, (C.c_int32 * 3)(4, 0, 4), shp, 1e-3, ptr, None,
                                    None), "empty output shape")
    _rejects(lib.hp_zoom_nearest_i32(None, 4, 4, 4, shp, ptr, None), "hp_zoom_nearest_i32")
    _rejects(lib.hp_label_sigma_f64(ptr, 8, None, None, 2, ptr, ptr, None), "hp_label_sigma_f64")
    assert lib.hp_label_sigma_f64(ptr, 0, None, None, 0, ptr, ptr, None) == 0
    _rejects(lib.hp_extend_sigma_f64(ptr, ptr, ptr, ptr, 4, 4, 4, ptr, None), "hp_extend_sigma")
    _rejects(lib.hp_column_stats(ptr, 4, 4, 4, None, None, None, None), "hp_column_stats")
    _rejects(lib.hp_scan_i32(ptr, 0, ptr, None), "hp_scan_i32")
    _rejects(lib.hp_column_emit(ptr, 4, 4, 4, ptr, None, None, None, None, ptr, None),
             "hp_column_emit")
    _rejects(lib.hp_surface_flags(None, 4, 4, 4, ptr, None), "hp_surface_flags")
    _rejects(lib.hp_select_rank(ptr, ptr, 16, 4, -1, ptr, None), "hp_select_rank")
    _rejects(lib.hp_torso_faces_f64(ptr, None, ptr, 4, 4, 4, 0.0, ptr, ptr, ptr, 0, ptr, ptr,
                                    ptr, ptr, None, 8, ptr, None), "hp_torso_faces_f64")
    _rejects(lib.hp_register_dice_f64(ptr, 8, ptr, 8, ptr, 48, ptr, ptr, ptr, ptr, 0.0, ptr, 16,
                                      ptr, None), "hp_register_dice_f64")
    _rejects(lib.hp_register_dice_f64(ptr, 0, ptr, 8, ptr, 48, ptr, ptr, ptr, ptr, 3.0, ptr, 16,
                                      ptr, None), "hp_register_dice_f64")
    _rejects(lib.hp_project_rows_f64(ptr, 8, ptr, 1.0, ptr, 0.0, ptr, 4, 4, 4, ptr, None),
             "hp_project_rows_f64")
    assert lib.hp_project_rows_f64(ptr, 0, ptr, 1.0, ptr, 1.5, ptr, 4, 4, 4, ptr, None) == 0
