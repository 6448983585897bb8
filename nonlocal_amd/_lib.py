"""ctypes binding of libnlf.so (the C ABI declared in include/nlf.h).

There is no fallback: if the library is missing the import of any compute path fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnlf.so")

_p = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_int = C.c_int
_dbl = C.c_double
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)
_pp = C.POINTER(C.c_void_p)

# name: (restype, argtypes)
SIGNATURES = {
    "nlf_ctx_create": (_int, [_int, _p, _pp]),
    "nlf_ctx_destroy": (_int, [_p]),
    "nlf_last_error": (C.c_char_p, [_p]),
    "nlf_ctx_synchronize": (_int, [_p]),
    "nlf_ctx_set_option": (_int, [_p, _int, _int]),
    "nlf_ctx_stream": (_p, [_p]),
    "nlf_mesh_create": (_int, [_p, _i64, _f64p, _i64, _int, _i32p, _int, _i64p, _i64, _int, _i32p, _int, _i64p, _pp]),
    "nlf_mesh_destroy": (_int, [_p]),
    "nlf_mesh_info": (_int, [_p, _i64p, _i64p, C.POINTER(_int), C.POINTER(_int)]),
    "nlf_mesh_export_quadrature": (_int, [_p, _f64p, _f64p, _f64p, _f64p]),
    "nlf_mesh_export_basis": (_int, [_p, _f64p, _f64p]),
    "nlf_find_neighbours": (_int, [_p, _f64p, _int, _f64p]),
    "nlf_find_neighbours_ex": (_int, [_p, _f64p, _p, _int, _f64p]),
    "nlf_matrix_symv_info": (_int, [_p, _i64p, _i64p]),
    "nlf_batch_set_preconditioner": (_int, [_p, _p]),
    "nlf_neighbours_count": (_int, [_p, _i64p]),
    "nlf_neighbours_export": (_int, [_p, _i64p, _i32p]),
    "nlf_neighbours_device": (_int, [_p, _pp, _pp]),
    "nlf_assemble": (_int, [_p, _int, C.POINTER(_abi.Model), _p, _u8p, _int, _int, _pp]),
    "nlf_assemble_ex": (_int, [_p, _int, C.POINTER(_abi.Model), _p, _u8p, _int, _int, C.POINTER(_abi.AssembleOptions),
                               _pp]),
    "nlf_matrix_destroy": (_int, [_p]),
    "nlf_matrix_info": (_int, [_p, _i64p, _i64p, _i64p, _i64p]),
    "nlf_matrix_timings": (_int, [_p, _f64p, _f64p]),
    "nlf_matrix_export": (_int, [_p, _int, _i64p, _i64p, _f64p]),
    "nlf_matrix_export_rows": (_int, [_p, _i64, _i64, _i64p, _i64p, _f64p]),
    "nlf_matrix_device": (_int, [_p, _pp, _pp, _pp, _pp]),
    "nlf_matrix_add_pairs": (_int, [_p, _i64, _i64p, _f64p]),
    "nlf_matrix_add_convection": (_int, [_p, C.POINTER(_abi.Boundary)]),
    "nlf_matrix_make_fp32": (_int, [_p]),
    "nlf_matrix_assemble_fp32": (_int, [_p]),
    "nlf_matrix_make_split": (_int, [_p]),
    "nlf_matrix_compare_fp32": (_int, [_p, _f64p]),
    "nlf_matrix_pattern_info": (_int, [_p, _i64p, _i64p]),
    "nlf_learned_set_coarse": (_int, [_p, _p, _i64]),
    "nlf_learned_coarse_info": (_int, [_p, _i64p, _f64p]),
    "nlf_boundary_data": (_int, [_p, _int, C.POINTER(_abi.Boundary), _f64p, _f64p, _u8p]),
    "nlf_apply_first_kind": (_int, [_p, _p, _p]),
    "nlf_heat_flux": (_int, [_p, C.POINTER(_abi.Model), C.POINTER(_abi.ThermalMaterial), _p, _p, _p]),
    "nlf_matrix_combine": (_int, [_p, _dbl, _p, _dbl]),
    "nlf_right_part": (_int, [_p, _int, _p, _p]),
    "nlf_strain_stress": (_int, [_p, C.POINTER(_abi.Model), C.POINTER(_abi.ElasticMaterial), _p, _p, _p, _p, _p]),
    "nlf_temperature_condition": (_int, [_p, C.POINTER(_abi.Model), C.POINTER(_abi.ElasticMaterial), _p, _p]),
    "nlf_boundary_data_fn": (_int, [_p, _int, C.POINTER(_abi.Boundary), _p, _p, _f64p, _f64p, _u8p]),
    "nlf_spmv": (_int, [_p, _p, _p]),
    "nlf_spmv_fp32": (_int, [_p, _p, _p]),
    "nlf_spmv_lpr": (_int, [_p, _int, _int, _p, _p]),
    "nlf_precond_create": (_int, [_p, _int, _pp]),
    "nlf_precond_destroy": (_int, [_p]),
    "nlf_precond_info": (_int, [_p, C.POINTER(_int), _f64p, C.POINTER(_int), C.POINTER(_int), _f64p]),
    "nlf_precond_sweep_info": (_int, [_p, C.POINTER(C.c_int32)]),
    "nlf_precond_blocks": (_int, [_p, C.POINTER(C.c_int32)]),
    "nlf_precond_apply": (_int, [_p, _p, _p]),
    "nlf_precond_export_factor": (_int, [_p, _f64p, _f64p]),
    "nlf_cg_solve": (_int, [_p, _p, _p, _p, _dbl, _i64, _int, C.POINTER(_abi.SolveStats)]),
    "nlf_cg_solve_noalloc": (_int, [_p, _p, _p, _p, _dbl, _i64, C.POINTER(_abi.SolveStats)]),
    "nlf_learned_num_params": (_int, []),
    "nlf_learned_init_weights": (_int, [C.c_uint64, _f32p]),
    "nlf_learned_create": (_int, [_p, _p, C.POINTER(_abi.Model), _pp]),
    "nlf_learned_set_weights": (_int, [_p, _f32p]),
    "nlf_learned_get_weights": (_int, [_p, _f32p]),
    "nlf_learned_info": (_int, [_p, C.POINTER(_int), _i64p, _f64p]),
    "nlf_learned_set_apply": (_int, [_p, _int]),
    "nlf_learned_export_graph": (_int, [_p, _i64p, _i32p, _f32p, _f32p]),
    "nlf_learned_loss_grad": (_int, [_p, _p, _p, _p, _f64p, _f32p]),
    "nlf_trainer_create": (_int, [_p, _f32p, _pp]),
    "nlf_trainer_destroy": (_int, [_p]),
    "nlf_trainer_step": (_int, [_p, _p, _p, _p, _p, C.c_float, _f64p]),
    "nlf_trainer_get_weights": (_int, [_p, _f32p]),
    "nlf_fcg_solve": (_int, [_p, _p, _p, _p, _p, _dbl, _i64, _int, C.POINTER(_abi.SolveStats)]),
    "nlf_batch_create": (_int, [_p, _int, _i64p, _pp]),
    "nlf_batch_destroy": (_int, [_p]),
    "nlf_batch_autotune": (_int, [_p, C.c_char_p, _f64p]),
    "nlf_batch_solve": (_int, [_p, _p, _p, _dbl, _i64, _int, _int, _i64p, _f64p, C.POINTER(_abi.SolveStats)]),
}

_LIB = None


def load():
    """Load libnlf.so and declare every signature; raises if the library or a symbol is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libnlf.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class NLFError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nlf error {code}: {msg}")
        self.code = code


def check(code, ctx=None):
    if code != 0:
        msg = load().nlf_last_error(ctx).decode() if ctx else "invalid argument"
        raise NLFError(code, msg)
