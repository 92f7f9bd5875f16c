"""ctypes binding of libclawpack_amd.so (declarations in include/clawpack_amd.h).

The HIP library is the only compute path of this package: if it is missing the import of the
binding fails with an error that says how to build it, and nothing falls back to the CPU.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclawpack_amd.so")

SOLVER_SHALLOW_ROE_EFIX, SOLVER_EULER_4WAVE, SOLVER_SHALLOW_BATHY_FWAVE = 0, 1, 2
FP64, FP32 = 0, 1
STATUS_BAD_STATE, STATUS_CFL_FIXED_DT, STATUS_MAX_STEPS = 1, 2, 4
FLAG_RUN_TO_TEND, FLAG_DT_VARIABLE = 1, 2


class ClawDesc(C.Structure):
    _fields_ = [
        ("solver", C.c_int32), ("dtype", C.c_int32), ("mx", C.c_int32), ("my", C.c_int32),
        ("num_eqn", C.c_int32), ("num_waves", C.c_int32), ("num_ghost", C.c_int32),
        ("order", C.c_int32), ("transverse_waves", C.c_int32), ("mthlim", C.c_int32 * 4),
        ("bc_lower", C.c_int32 * 2), ("bc_upper", C.c_int32 * 2),
        ("reflect_index", C.c_int32 * 2), ("seg_h", C.c_int32), ("dt_variable", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("grav", C.c_double), ("gamma", C.c_double),
        ("cfl_max", C.c_double), ("cfl_desired", C.c_double), ("dt_max", C.c_double),
        ("num_aux", C.c_int32), ("fwave", C.c_int32), ("aux_bc_lower", C.c_int32 * 2),
        ("aux_bc_upper", C.c_int32 * 2), ("dry_tolerance", C.c_double),
        ("mcapa", C.c_int32), ("reserved_capa", C.c_int32), ("depth_ref", C.c_double),
        ("wg_waves", C.c_int32), ("launch_order", C.c_int32),
    ]


class ClawCtrl(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("dt", C.c_double), ("tend", C.c_double), ("tstart", C.c_double),
        ("cfl", C.c_double), ("cflmax", C.c_double), ("dtmin", C.c_double), ("dtmax", C.c_double),
        ("smax_last", C.c_double * 2), ("smax_bits", C.c_uint64 * 2),
        ("numsteps", C.c_longlong), ("attempts", C.c_longlong), ("rejected", C.c_longlong),
        ("target_steps", C.c_longlong), ("max_steps", C.c_longlong),
        ("cur", C.c_int32), ("done", C.c_int32), ("status", C.c_int32), ("flags", C.c_int32),
        ("arrive", C.c_int32), ("reserved", C.c_int32),
    ]


class ClawBatchGrid(C.Structure):
    _fields_ = [
        ("offset", C.c_int64), ("mx", C.c_int32), ("my", C.c_int32), ("nstrips", C.c_int32),
        ("nsegs", C.c_int32), ("seg_h", C.c_int32), ("item0", C.c_int32),
        ("reserved", C.c_int32 * 2), ("dx", C.c_double), ("dy", C.c_double),
    ]


# name -> (restype, argtypes); every symbol declared in include/clawpack_amd.h
_VP = C.c_void_p
_PD = C.POINTER(ClawDesc)
SIGNATURES = {
    "claw2d_validate": (C.c_int, [_PD]),
    "claw2d_state_bytes": (C.c_size_t, [_PD]),
    "claw2d_ctrl_init": (C.c_int, [_PD, C.POINTER(ClawCtrl), C.c_double, C.c_double, C.c_double,
                                   C.c_int32, C.c_longlong, C.c_longlong]),
    "claw2d_step": (C.c_int, [_PD, _VP, _VP, C.c_double, _VP, _VP]),
    "claw2d_advance": (C.c_int, [_PD, _VP, _VP, _VP, C.c_longlong, _VP]),
    "claw2d_step_ds": (C.c_int, [_PD, _VP, _VP, C.c_double, _VP, _VP]),
    "claw2d_advance_ds": (C.c_int, [_PD, _VP, _VP, _VP, C.c_longlong, _VP]),
    "claw2d_step_aux": (C.c_int, [_PD, _VP, _VP, _VP, C.c_double, _VP, _VP]),
    "claw2d_advance_aux": (C.c_int, [_PD, _VP, _VP, _VP, _VP, C.c_longlong, _VP]),
    "claw2d_step_ds_aux": (C.c_int, [_PD, _VP, _VP, _VP, C.c_double, _VP, _VP]),
    "claw2d_advance_ds_aux": (C.c_int, [_PD, _VP, _VP, _VP, _VP, C.c_longlong, _VP]),
    "claw2d_step_padded": (C.c_int, [_PD, _VP, _VP, _VP, C.c_double, _VP, _VP]),
    "claw2d_sweep_padded": (C.c_int, [_PD, _VP, _VP, _VP, C.c_double, C.c_int32, _VP, _VP]),
    "claw2d_mass_sum_workspace_bytes": (C.c_size_t, [_PD]),
    "claw2d_mass_sum": (C.c_int, [_PD, _VP, C.c_int32, _VP, _VP, _VP]),
    "claw2d_mass_sum_noalloc": (C.c_int, [_PD, _VP, C.c_int32, _VP, _VP]),
    "claw2d_mass_sum_noalloc_max_cells": (C.c_size_t, []),
    "claw2d_adjoint_workspace_bytes": (C.c_size_t, [_PD]),
    "claw2d_adjoint_step": (C.c_int, [_PD, _VP, _VP, _VP, _VP, C.c_double, _VP]),
    "claw2d_batch_prepare": (C.c_int, [_PD, C.POINTER(ClawBatchGrid), C.c_int32, C.c_int32,
                                       C.POINTER(C.c_int64)]),
    "claw2d_batch_advance": (C.c_int, [_PD, _VP, C.c_int32, C.c_int64, _VP, _VP, _VP, _VP,
                                       C.c_longlong, _VP]),
    "claw2d_batch_graph_create": (C.c_int, [_PD, _VP, C.c_int32, C.c_int64, _VP, _VP, _VP, _VP,
                                            C.c_longlong, C.POINTER(C.c_void_p)]),
    "claw2d_graph_launch": (C.c_int, [_VP, _VP]),
    "claw2d_graph_destroy": (C.c_int, [_VP]),
    "claw2d_abi_struct_size": (C.c_size_t, [C.c_int32]),
    "claw2d_last_error": (C.c_char_p, []),
}

_lib = None


class ClawLibError(RuntimeError):
    pass


def load():
    """Load libclawpack_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    # torch first: the library then binds to the HIP runtime torch already loaded (same soname),
    # instead of loading the system one ahead of torch's and leaving torch without a device
    import torch  # noqa: F401
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libclawpack_amd.so is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m clawpack_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.claw2d_abi_struct_size(0) != C.sizeof(ClawDesc) or \
            lib.claw2d_abi_struct_size(1) != C.sizeof(ClawCtrl) or \
            lib.claw2d_abi_struct_size(2) != C.sizeof(ClawBatchGrid):
        raise ImportError("claw2d_desc/claw2d_ctrl layout mismatch between the binding and the library")
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        msg = load().claw2d_last_error().decode(errors="replace")
        raise ClawLibError("%s failed (%d): %s" % (what, rc, msg))
