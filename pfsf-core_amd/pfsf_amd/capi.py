"""ctypes binding of libpfsf (the HIP C ABI declared in include/pfsf/*.h).

Device arrays are passed as raw pointers; torch tensors are used only as the
allocator and stream provider (``tensor.data_ptr()``,
``torch.cuda.current_stream().cuda_stream``).  There is no fallback: a missing
or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

PKG_ROOT = pathlib.Path(__file__).resolve().parent.parent
# PFSF_LIBRARY names another build of libpfsf (e.g. a tuning variant).
LIB_PATH = pathlib.Path(os.environ.get("PFSF_LIBRARY", PKG_ROOT / "lib" / "libpfsf.so"))
INCLUDE_DIR = PKG_ROOT.parent / "include" / "pfsf"

# pfsf_result
OK, ERROR_DEVICE, NO_DEVICE, OUT_OF_VRAM, INVALID_ARG, NOT_INIT, ISLAND_FULL = 0, -1, -2, -3, -4, -5, -6
PREC_FP32, PREC_BF16_COARSE = 0, 1


class PfsfError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} returned {code}")
        self.code = code


class Dims(C.Structure):
    _fields_ = [("lx", C.c_int32), ("ly", C.c_int32), ("lz", C.c_int32)]


class Pos(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32)]


class Config(C.Structure):
    _fields_ = [("max_island_size", C.c_int32), ("tick_budget_ms", C.c_int32),
                ("vram_budget_bytes", C.c_int64), ("enable_phase_field", C.c_bool),
                ("enable_multigrid", C.c_bool)]


class Stats(C.Structure):
    _fields_ = [("island_count", C.c_int32), ("total_voxels", C.c_int64),
                ("vram_used_bytes", C.c_int64), ("vram_budget_bytes", C.c_int64),
                ("last_tick_ms", C.c_float)]


class IslandDesc(C.Structure):
    _fields_ = [("island_id", C.c_int32), ("origin", Pos), ("lx", C.c_int32),
                ("ly", C.c_int32), ("lz", C.c_int32)]


class FailureEvent(C.Structure):
    _fields_ = [("pos", Pos), ("type", C.c_int32)]


class TickResult(C.Structure):
    _fields_ = [("events", C.POINTER(FailureEvent)), ("capacity", C.c_int32),
                ("count", C.c_int32)]


class IslandBuffers(C.Structure):
    _fields_ = [(f, t) for name in ("phi", "source", "conductivity", "voxel_type", "rcomp",
                                    "rtens", "max_phi")
                for f, t in ((name + "_addr", C.c_void_p), (name + "_bytes", C.c_int64))]


class IslandLookups(C.Structure):
    _fields_ = [(f, t) for name in ("material_id", "anchor_bitmap", "fluid_pressure", "curing")
                for f, t in ((name + "_addr", C.c_void_p), (name + "_bytes", C.c_int64))]


class SolverDesc(C.Structure):
    _fields_ = [("dims", Dims), ("type", C.c_void_p), ("kappa", C.c_void_p),
                ("mg_levels", C.c_int32), ("precision", C.c_int32),
                ("coarse_sweeps", C.c_int32), ("tol", C.c_double), ("max_iters", C.c_int32),
                ("stream", C.c_void_p)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("refinements", C.c_int32),
                ("converged", C.c_int32), ("levels", C.c_int32),
                ("rel_residual", C.c_double), ("solve_ms", C.c_double)]


P, I32, I64, F32, F64, U32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double, C.c_uint32
R = C.c_int  # pfsf_result

# pfsf_material (pfsf_types.h): six floats and a bool, 4-byte aligned (28 bytes)
MATERIAL_DTYPE = np.dtype([("density", "<f4"), ("rcomp", "<f4"), ("rtens", "<f4"),
                           ("youngs_gpa", "<f4"), ("poisson", "<f4"), ("gc", "<f4"),
                           ("is_anchor", "u1")], align=True)

class IslandStatus(C.Structure):
    _fields_ = [("last_solve", SolveInfo), ("solver_builds", C.c_int32),
                ("operator_mode", C.c_int32), ("dirty", C.c_int32), ("reserved", C.c_int32)]


class BatchIsland(C.Structure):
    _fields_ = [("dims", Dims), ("type", C.c_void_p), ("kappa", C.c_void_p)]


class BatchDesc(C.Structure):
    _fields_ = [("count", C.c_int32), ("islands", C.POINTER(BatchIsland)),
                ("coarse_sweeps", C.c_int32), ("tol", C.c_double), ("max_iters", C.c_int32),
                ("stream", C.c_void_p), ("precision", C.c_int32), ("shape", C.c_int32)]


BATCH_SHAPES = 6
BATCH_CLASSES = 4
BATCH_AUTOTUNE = -1
# (finest column depth, coarse tile depth) of each candidate, in BatchClass.ms order
BATCH_SHAPE_CANDIDATES = [(8, 2), (4, 2), (16, 2), (8, 4), (4, 4), (16, 4)]


class BatchClass(C.Structure):
    _fields_ = [("min_cells", C.c_int64), ("max_cells", C.c_int64), ("islands", C.c_int32),
                ("sample", C.c_int32), ("tz_fine", C.c_int32), ("tz_coarse", C.c_int32),
                ("ms", C.c_float * BATCH_SHAPES)]


class BatchTuning(C.Structure):
    _fields_ = [("sample", C.c_int32), ("mixed", C.c_int32), ("mixed_ms", C.c_float),
                ("uniform_ms", C.c_float * BATCH_SHAPES)]


class BatchInfo(C.Structure):
    _fields_ = [("islands", C.c_int32), ("converged", C.c_int32),
                ("max_iterations", C.c_int32), ("sweeps", C.c_int32),
                ("total_iterations", C.c_int64), ("max_rel_residual", C.c_double),
                ("solve_ms", C.c_double)]


class LoadPathDesc(C.Structure):
    _fields_ = [("dims", Dims), ("type", C.c_void_p), ("kappa", C.c_void_p),
                ("source", C.c_void_p), ("rcomp", C.c_void_p), ("steps", C.c_int32),
                ("checkpoint_every", C.c_int32), ("tol", C.c_double), ("max_iters", C.c_int32),
                ("stream", C.c_void_p), ("damage", C.c_int32)]


class LoadPathInfo(C.Structure):
    _fields_ = [("loss", C.c_double), ("forward_solves", C.c_int32),
                ("recompute_solves", C.c_int32), ("adjoint_solves", C.c_int32),
                ("checkpoints", C.c_int32), ("pcg_iterations", C.c_int64), ("ms", C.c_double)]


SIGNATURES = {
    # pfsf.h
    "pfsf_create": (P, [C.POINTER(Config)]),
    "pfsf_create_with_vulkan": (P, [P, P]),
    "pfsf_init": (R, [P]),
    "pfsf_shutdown": (None, [P]),
    "pfsf_destroy": (None, [P]),
    "pfsf_is_available": (C.c_bool, [P]),
    "pfsf_get_stats": (R, [P, C.POINTER(Stats)]),
    "pfsf_set_material_lookup": (None, [P, P, P]),
    "pfsf_set_anchor_lookup": (None, [P, P, P]),
    "pfsf_set_fill_ratio_lookup": (None, [P, P, P]),
    "pfsf_set_curing_lookup": (None, [P, P, P]),
    "pfsf_set_wind": (None, [P, P]),
    "pfsf_set_pcg_enabled": (None, [P, C.c_int]),
    "pfsf_add_island": (R, [P, C.POINTER(IslandDesc)]),
    "pfsf_remove_island": (None, [P, I32]),
    "pfsf_notify_block_change": (R, [P, I32, P]),
    "pfsf_mark_full_rebuild": (None, [P, I32]),
    "pfsf_tick": (R, [P, P, I32, I64, C.POINTER(TickResult)]),
    "pfsf_read_stress": (R, [P, I32, P, I32, C.POINTER(I32)]),
    "pfsf_register_island_buffers": (R, [P, I32, C.POINTER(IslandBuffers)]),
    "pfsf_register_island_lookups": (R, [P, I32, C.POINTER(IslandLookups)]),
    "pfsf_register_stress_readback": (R, [P, I32, P, I64]),
    "pfsf_tick_dbb": (R, [P, P, I32, I64, P, I64]),
    "pfsf_drain_callbacks": (I32, [P, P, I32]),
    "pfsf_get_sparse_upload_buffer": (R, [P, I32, C.POINTER(P), C.POINTER(I64)]),
    "pfsf_notify_sparse_updates": (R, [P, I32, I32]),
    "pfsf_version": (C.c_char_p, []),
    # pfsf_version.h
    "pfsf_abi_version": (U32, []),
    "pfsf_has_feature": (C.c_bool, [C.c_char_p]),
    "pfsf_build_info": (C.c_char_p, []),
    "pfsf_abi_contract_version": (C.c_char_p, []),
    # pfsf_hip.h
    "pfsf_hip_jacobi_step": (R, [Dims, P, P, P, P, P, F32, P]),
    "pfsf_hip_rbgs": (R, [Dims, P, P, P, P, I32, P]),
    "pfsf_hip_failure_scan": (R, [Dims, P, P, P, P, P, F32, P, P]),
    "pfsf_hip_history_update": (R, [Dims, P, P, P, P, P, P]),
    "pfsf_hip_apply_f64": (R, [Dims, P, P, P, P, P]),
    "pfsf_hip_compact_failures": (R, [Dims, P, Pos, P, I32, P]),
    "pfsf_hip_compact_failures_noalloc": (R, [Dims, P, Pos, P, I32, P]),
    "pfsf_hip_downsample_2to1": (R, [Dims, P, P, P, P, P]),
    "pfsf_hip_normalize_soa6": (R, [P, P, P, P, I64, C.POINTER(F32), P]),
    "pfsf_hip_kappa_from_soa6": (R, [P, I64, P, P]),
    "pfsf_hip_build_island_data": (R, [Dims, P, P, I32, P, P, P, P, P, P, P, P, P, P, P]),
    "pfsf_solver_create": (R, [C.POINTER(SolverDesc), C.POINTER(P)]),
    "pfsf_solver_solve": (R, [P, P, P, P, C.POINTER(SolveInfo)]),
    "pfsf_solver_update": (R, [P, P, P]),
    "pfsf_solver_destroy": (None, [P]),
    "pfsf_batch_create": (R, [C.POINTER(BatchDesc), C.POINTER(P)]),
    "pfsf_batch_solve": (R, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(BatchInfo),
                             C.POINTER(SolveInfo)]),
    "pfsf_batch_destroy": (None, [P]),
    "pfsf_batch_classes": (C.c_int32, [P, C.POINTER(BatchClass), C.c_int32]),
    "pfsf_batch_tuning_info": (R, [P, C.POINTER(BatchTuning)]),
    "pfsf_load_path_loss": (R, [C.POINTER(LoadPathDesc), C.c_int32, C.POINTER(LoadPathInfo)]),
    "pfsf_load_path_grad": (R, [C.POINTER(LoadPathDesc), P, C.POINTER(LoadPathInfo)]),
    "pfsf_solver_profile": (R, [P, I32, I32, C.POINTER(F64)]),
    "pfsf_hip_anchored_mask": (R, [Dims, P, P, P]),
    "pfsf_hip_label_components": (R, [Dims, P, I32, P, P, I32, C.POINTER(I32), C.POINTER(I32),
                                      P]),
    "pfsf_island_components": (R, [P, I32, I32, P, P, P]),
    "pfsf_register_island_kappa": (R, [P, I32, P, I64]),
    "pfsf_set_solve_tolerance": (None, [P, F64, I32]),
    "pfsf_read_history": (R, [P, I32, P, I32, C.POINTER(I32)]),
    "pfsf_get_island_status": (R, [P, I32, C.POINTER(IslandStatus)]),
    "pfsf_read_damage": (R, [P, I32, P, I32, C.POINTER(I32)]),
    "pfsf_hip_macro_residual": (R, [Dims, P, P, P, P, P, P]),
    "pfsf_hip_phase_field_evolve": (R, [Dims, P, P, P, P, P, P, P, P, F32, F32, F32, I32, P]),
    "pfsf_hip_chebyshev_jacobi": (R, [Dims, P, P, P, P, P, I32, F32, P]),
    "pfsf_hip_soa6_couplings": (R, [Dims, P, P, P, P]),
    "pfsf_hip_soa6_apply_f64": (R, [Dims, P, P, P, P, P]),
    "pfsf_hip_soa6_failure_scan": (R, [Dims, P, P, P, P, P, F32, P, P]),
    "pfsf_hip_soa6_history_update": (R, [Dims, P, P, P, P, P, P]),
    "pfsf_hip_soa6_rbgs": (R, [Dims, P, P, P, P, I32, P]),
    # pfsf_compute.h / pfsf_diagnostics.h (host primitives)
    "pfsf_normalize_soa6": (None, [P, P, P, P, P, I32, C.POINTER(F32)]),
    "pfsf_downsample_2to1": (None, [P, P, I32, I32, I32, P, P]),
    "pfsf_chebyshev_omega": (F32, [I32, F32]),
    "pfsf_precompute_omega_table": (I32, [F32, P, I32]),
    "pfsf_estimate_spectral_radius": (F32, [I32, F32]),
}

_lib = None


def load():
    """Load libpfsf.so and attach signatures; raises if it is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C pfsf-core_amd/csrc")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(name: str, code: int) -> None:
    if code != OK:
        raise PfsfError(name, code)


def dims(lx: int, ly: int, lz: int) -> Dims:
    return Dims(lx, ly, lz)
