"""ctypes binding of libpsum_hip.so (the C ABI declared in include/psum_hip.h
and include/psum_field_solver.h).

The library is the product: there is no fallback.  If it is missing, or a
GPU entry point is called without a GPU, the call raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsum_hip.so")

PSUM_F64 = 0
PSUM_F32 = 1
KEY_INVALID = 0xFFFFFFFF

_STATUS = {0: "PSUM_OK", 1: "PSUM_ERR_ARG", 2: "PSUM_ERR_HIP", 3: "PSUM_ERR_UNSUPPORTED",
           4: "PSUM_ERR_NOMEM", 5: "PSUM_ERR_STATE"}


class PsumGrid(C.Structure):
    """psum_grid (include/psum_hip.h)."""
    _fields_ = [
        ("dim", C.c_int32),
        ("cells", C.c_int32 * 3),
        ("lb", C.c_double * 3),
        ("ub", C.c_double * 3),
        ("delta", C.c_double * 3),
        ("delta_r", C.c_double * 3),
        ("periodic", C.c_int32 * 3),
        ("pad_", C.c_int32),
        ("plo", C.c_double * 3),
        ("plen", C.c_double * 3),
    ]


class PsumSpecies(C.Structure):
    _fields_ = [("q", C.c_double), ("m", C.c_double), ("w", C.c_double),
                ("n", C.c_int64), ("capacity", C.c_int64)]


class PsumStepConfig(C.Structure):
    _fields_ = [
        ("grid", PsumGrid),
        ("dtype", C.c_int32),
        ("mixed_mg", C.c_int32),
        ("nspecies", C.c_int32),
        ("depth", C.c_int32),
        ("pre", C.c_int32),
        ("post", C.c_int32),
        ("max_cycles", C.c_int32),
        ("sort_every", C.c_int32),
        ("mg_graph", C.c_int32),
        ("extrapolate", C.c_int32),
        ("tol", C.c_double),
        ("eps", C.c_double),
        ("dt", C.c_double),
        ("rho_bg", C.c_double),
        ("species", PsumSpecies * 4),
        ("fixed", C.c_void_p),
        ("fixed_val", C.c_void_p),
    ]


class PsumBatchTuning(C.Structure):
    _fields_ = [("fixed_cycles", C.c_int32), ("sort_every", C.c_int32), ("sort_shift", C.c_int32),
                ("pad_", C.c_int32), ("sort_ms", C.c_double), ("push_ms", C.c_double),
                ("push_sort_cost_ms", C.c_double)]


class PsumMixedNode(C.Structure):
    _fields_ = [("element", C.c_uint64), ("direction", C.c_int32), ("pad_", C.c_int32),
                ("value_coef", C.c_double), ("grad_coef", C.c_double)]


EPS_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_void_p)


class PsumPoissonSpec(C.Structure):
    _fields_ = [("cylindrical", C.c_int32), ("pad_", C.c_int32), ("eps", C.c_double),
                ("eps_fn", EPS_FN), ("eps_user", C.c_void_p), ("fixed", C.c_void_p),
                ("mixed", C.c_void_p), ("n_mixed", C.c_int64)]


class FuncTable(C.Structure):
    """struct func_table (PSuM field_solver/register/interface.h:14-23)."""
    _fields_ = [
        ("init", C.c_void_p),
        ("set_matrix", C.c_void_p),
        ("solve", C.c_void_p),
        ("set_options", C.c_void_p),
        ("set_source_replace", C.c_void_p),
        ("set_source_addback", C.c_void_p),
        ("free", C.c_void_p),
    ]


# (name, restype, argtypes) of every exported entry point
MCC_MAX_SEG = 12
MCC_MAX_CHAN = 8
MCC_MAX_PROC = 4
MCC_ELASTIC, MCC_EXCITATION, MCC_IONIZATION, MCC_ION_ISOTROPIC, MCC_ION_BACKWARD = range(5)


class PsumMccChannel(C.Structure):
    """psum_mcc_channel (include/psum_hip.h)."""
    _fields_ = [
        ("kind", C.c_int32),
        ("nseg", C.c_int32),
        ("exp_form", C.c_int32),
        ("pad_", C.c_int32),
        ("arg_scale", C.c_double),
        ("threshold_ev", C.c_double),
        ("seg_hi", C.c_double * MCC_MAX_SEG),
        ("coef", (C.c_double * 5) * MCC_MAX_SEG),
    ]


class PsumMccProcess(C.Structure):
    """psum_mcc_process (include/psum_hip.h)."""
    _fields_ = [
        ("nchan", C.c_int32),
        ("pad_", C.c_int32),
        ("density", C.c_double),
        ("unit_charge", C.c_double),
        ("incident_mass", C.c_double),
        ("target_mass", C.c_double),
        ("target_temperature", C.c_double),
        ("boltzmann", C.c_double),
        ("chan", PsumMccChannel * MCC_MAX_CHAN),
    ]


class PsumMccSpecies(C.Structure):
    """psum_mcc_species (include/psum_hip.h)."""
    _fields_ = [
        ("pos", C.c_void_p * 3),
        ("vel", C.c_void_p * 3),
        ("seed", C.c_void_p),
        ("n", C.c_int64),
        ("cap", C.c_int64),
    ]


class PsumMeshView(C.Structure):
    """psum_mesh_view (include/psum_hip.h): device view of a partition."""
    _fields_ = [
        ("I", C.c_int32), ("J", C.c_int32), ("K", C.c_int32), ("fast", C.c_int32),
        ("box", C.c_double * 6),
        ("rd", C.c_double * 3),
        ("cell_start", C.c_void_p),
        ("cell_tri", C.c_void_p),
        ("tri12", C.c_void_p),
        ("ids", C.c_void_p),
        ("udf", C.c_void_p),
    ]


P = C.c_void_p
I64 = C.c_int64
ST = C.c_int
_GP = C.POINTER(PsumGrid)
SIGNATURES = [
    ("psum_hip_last_error", C.c_char_p, []),
    ("psum_hip_bin", ST, [_GP, C.c_int, P, I64, P, P]),
    ("psum_hip_sort_workspace", C.c_size_t, [I64]),
    ("psum_hip_sort", ST, [P, I64, C.c_uint32, P, P, P, C.c_size_t, P]),
    ("psum_hip_permute", ST, [C.c_int, C.c_int, P, P, P, I64, P]),
    ("psum_hip_compact_workspace", C.c_size_t, [I64]),
    ("psum_hip_compact", ST, [C.c_int, C.c_int, P, P, I64, P, P, C.POINTER(I64), P, C.c_size_t, P]),
    ("psum_hip_compact_noalloc", ST, [C.c_int, C.c_int, P, P, I64, P, P, C.POINTER(I64), P]),
    ("psum_hip_deposit_cic", ST, [_GP, C.c_int, P, I64, C.c_double, C.c_double, P, P, P]),
    ("psum_hip_fixed_to_double", ST, [P, I64, C.c_double, C.c_double, P, P]),
    ("psum_hip_push_blocks", I64, [I64]),
    ("psum_hip_gather_push", ST, [_GP, C.c_int, P, P, I64, P, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, P, P, P, P]),
    ("psum_hip_deposit_bspline", ST, [_GP, C.c_int, C.c_int, P, I64, C.c_double, C.c_double, P, P, P]),
    ("psum_hip_gather_push_bspline", ST, [_GP, C.c_int, C.c_int, P, P, I64, P, C.c_double, C.c_double,
                                          C.c_double, P, P, P]),
    ("psum_hip_mcc_workspace", C.c_size_t, [I64, C.c_int]),
    ("psum_hip_mcc", ST, [_GP, C.c_int, C.POINTER(PsumMccProcess), C.c_int, P, C.POINTER(PsumMccSpecies),
                          C.POINTER(PsumMccSpecies), C.POINTER(PsumMccSpecies), C.c_double,
                          C.POINTER(I64), P, P, C.c_size_t, P]),
    ("psum_mesh_create", P, [P, P, I64, C.c_int, C.c_int, C.c_int, P]),
    ("psum_mesh_info", ST, [P, P, P]),
    ("psum_mesh_partition", ST, [P, P, P, P]),
    ("psum_mesh_set_fast", ST, [P, C.c_int]),
    ("psum_hip_mesh_detect", ST, [P, C.c_int, P, P, I64, P, P, P, P]),
    ("psum_hip_mesh_boundary", ST, [P, C.c_int, P, P, P, P, I64, C.c_double, C.c_double, P, P, P]),
    ("psum_mesh_get_view", ST, [P, C.POINTER(PsumMeshView)]),
    ("psum_mesh_set_udf", ST, [P, C.c_int]),
    ("psum_mesh_destroy", None, [P]),
    ("psum_hip_reduce_workspace", C.c_size_t, []),
    ("psum_hip_reduce_diag", ST, [P, I64, P, P, P]),
    ("psum_hip_field_energy_workspace", C.c_size_t, [_GP]),
    ("psum_hip_field_energy", ST, [_GP, P, C.c_double, P, P, C.c_size_t, P]),
    ("psum_hip_f64_to_f32", ST, [P, P, I64, P]),
    ("psum_stepper_create", P, [C.POINTER(PsumStepConfig)]),
    ("psum_stepper_load", ST, [P, C.c_int, P, P, I64]),
    ("psum_stepper_state", ST, [P, C.c_int, P, P, C.POINTER(I64)]),
    ("psum_stepper_run", ST, [P, C.c_int, P]),
    ("psum_stepper_phi", ST, [P, P]),
    ("psum_stepper_rho", ST, [P, P, C.POINTER(C.c_double)]),
    ("psum_stepper_stats", ST, [P, C.POINTER(I64), C.POINTER(C.c_double), P]),
    ("psum_stepper_step_count", ST, [P, C.POINTER(I64)]),
    ("psum_stepper_sync", ST, [P]),
    ("psum_stepper_set_timing", ST, [P, C.c_int]),
    ("psum_stepper_push_timing", ST, [P, C.POINTER(C.c_double), C.POINTER(I64), C.POINTER(I64)]),
    ("psum_stepper_forward_loss", ST, [P, C.c_int, C.POINTER(C.c_double)]),
    ("psum_stepper_record", ST, [P, C.c_int, C.c_int, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    ("psum_stepper_reverse", ST, [P, C.c_double, P, P, C.POINTER(C.c_double), C.POINTER(I64)]),
    ("psum_stepper_adjoint", ST, [P, C.c_int, C.c_int, C.POINTER(C.c_double), P, P,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("psum_stepper_destroy", None, [P]),
    ("psum_stepper_set_interp_order", ST, [P, C.c_int]),
    ("psum_stepper_sort", ST, [P]),
    ("psum_stepper_save", ST, [P, C.c_char_p]),
    ("psum_stepper_restore", ST, [P, C.c_char_p]),
    ("psum_stepper_set_schedule", ST, [P, C.c_int, C.c_int, C.c_int]),
    ("psum_stepper_set_species_sort", ST, [P, C.c_int, C.c_int]),
    ("psum_stepper_species_sort", ST, [P, C.c_int, C.POINTER(C.c_int)]),
    ("psum_stepper_mg_tuning", ST, [P, C.c_char_p]),
    ("psum_stepper_set_smoothing", ST, [P, C.c_int, C.c_int]),
    ("psum_stepper_tune_smoothing", ST, [P, P, P, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), P]),
    ("psum_stepper_tune", ST, [P, C.c_int, C.c_int, C.POINTER(PsumBatchTuning)]),
    ("psum_poisson_assemble", ST, [_GP, C.c_double, P, I64, P, P, P, C.POINTER(I64)]),
    ("psum_poisson_assemble_ex", ST, [_GP, C.POINTER(PsumPoissonSpec), I64, P, P, P, C.POINTER(I64),
                                      I64, P, P, P, C.POINTER(I64)]),
    ("psum_batch_create", P, [C.POINTER(PsumStepConfig), C.c_int]),
    ("psum_batch_load", ST, [P, C.c_int, C.c_int, P, P, I64]),
    ("psum_batch_configure", ST, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("psum_batch_tune", ST, [P, C.c_int, C.c_int, C.POINTER(PsumBatchTuning)]),
    ("psum_batch_reset_history", ST, [P, I64]),
    ("psum_batch_set_lanes", ST, [P, C.c_int]),
    ("psum_batch_set_solver_batching", ST, [P, C.c_int]),
    ("psum_batch_solver_launches", ST, [P, C.POINTER(I64), C.POINTER(I64)]),
    ("psum_batch_run", ST, [P, C.c_int, C.POINTER(C.c_double)]),
    ("psum_batch_history", ST, [P, C.c_int, P, I64, C.POINTER(I64)]),
    ("psum_batch_instance", P, [P, C.c_int]),
    ("psum_batch_destroy", None, [P]),
    ("psum_hip_adjoint_push", ST, [_GP, C.c_int, P, P, P, I64, P, C.c_double, C.c_double, P, P, P]),
    ("psum_hip_adjoint_blocks", I64, [I64]),
    ("psum_hip_adjoint_gather", ST, [_GP, C.c_int, P, P, P, I64, P, C.c_double, C.c_double, P, P]),
    ("psum_hip_adjoint_scale", ST, [P, I64, C.c_double, P, C.c_int, P]),
    ("psum_hip_fixed_to_double_dev", ST, [P, I64, P, P, P]),
    ("psum_hip_field_energy_grad", ST, [_GP, P, C.c_double, P, P]),
    ("psum_hip_mode_displace", ST, [C.c_int, P, P, I64, P, C.c_int, C.c_double, C.c_double, P]),
    ("psum_hip_mode_project_workspace", C.c_size_t, [I64, C.c_int]),
    ("psum_hip_mode_project", ST, [C.c_int, P, P, I64, C.c_int, C.c_double, C.c_double, P, P, P]),
    ("register_solver", None, [C.c_char_p, FuncTable]),
    ("get_solver_funcs", FuncTable, [C.c_char_p]),
    ("psum_solver_create", P, [C.c_char_p]),
    ("psum_solver_set_options", ST, [P, C.c_char_p]),
    ("psum_solver_set_matrix", ST, [P, C.c_ulonglong, C.c_ulonglong, C.c_ulonglong, P, P, P]),
    ("psum_solver_set_source_replace", ST, [P, C.c_ulonglong, P, P]),
    ("psum_solver_set_source_addback", ST, [P, C.c_ulonglong, P, P]),
    ("psum_solver_solve", ST, [P, P, P]),
    ("psum_solver_stats", ST, [P, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("psum_solver_free", None, [P]),
]

_lib = None


def lib():
    """Load libpsum_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m psum_amd.build` "
                               "(the HIP engine has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class PsumError(RuntimeError):
    pass


def check(status, what=""):
    if status != 0:
        msg = lib().psum_hip_last_error().decode(errors="replace")
        raise PsumError(f"{what}: {_STATUS.get(status, status)}: {msg}")
