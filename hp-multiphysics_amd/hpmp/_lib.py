"""ctypes bindings of libhpgpu.so (declarations in include/hpgpu.h).

The product path has no CPU fallback: if the library is missing, loading fails
with an error that says how to build it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libhpgpu.so")

HPG_PHYSICS_INS = 0
HPG_PHYSICS_CNS = 1
HPG_FP64 = 0
HPG_FP32_DEFECT = 1
HPG_PRCN_ANALYTIC = 0
HPG_PRCN_LEARNED = 1
HPG_MLP_NOALLOC_MAX_SAMPLES = 27712
HPG_EDGE_GENERIC = 0
HPG_EDGE_EULER = 1
HPG_EDGE_CHARACTERISTIC = 2

# every entry point include/hpgpu.h declares
EXPORTS = (
    "hpg_create", "hpg_destroy", "hpg_set_state", "hpg_get_state", "hpg_get_residual", "hpg_set_residual",
    "hpg_get_preconditioner", "hpg_set_tprcn", "hpg_rsdl", "hpg_setup_preconditioner", "hpg_minvrt",
    "hpg_update", "hpg_maxres", "hpg_resnorm", "hpg_mg_restrict", "hpg_mg_prolongate", "hpg_get_log2p",
    "hpg_set_precision", "hpg_set_preconditioner", "hpg_set_cfl", "hpg_set_curved", "hpg_set_edge_flux", "hpg_set_periodic", "hpg_set_cycle_graph", "hpg_synchronize", "hpg_last_error", "hpg_basis_operator",
    "hpg_selftest_mfma", "hpg_kernel_timing", "hpg_kernel_timing_read", "hpg_selftest_mfma_f32", "hpg_selftest_mfma_bf16",
    "hpg_set_defect_inner", "hpg_apply_jacobian", "hpg_residual_fp32", "hpg_element_jacobian", "hpg_mlp_create", "hpg_mlp_destroy", "hpg_mlp_num_params",
    "hpg_mlp_get_weights", "hpg_mlp_set_weights", "hpg_mlp_forward", "hpg_mlp_loss_grad", "hpg_mlp_train",
    "hpg_mlp_loss_grad_noalloc",
    "hpg_set_mlp", "hpg_set_prcn_log_scalings", "hpg_mlp_features", "hpg_measure_jacobian_diag", "hpg_selftest_copy",
    "hpg_get_block_preconditioner", "hpg_link_fine", "hpg_mg_cycle",
    "hpg_tet_create", "hpg_tet_destroy", "hpg_tet_set_state", "hpg_tet_get_state", "hpg_tet_get_residual",
    "hpg_tet_set_residual", "hpg_tet_get_preconditioner", "hpg_tet_get_cns_blocks", "hpg_tet_rsdl", "hpg_tet_setup_preconditioner",
    "hpg_tet_minvrt", "hpg_tet_update", "hpg_tet_maxres", "hpg_tet_resnorm", "hpg_tet_mg_restrict",
    "hpg_tet_mg_prolongate", "hpg_tet_get_log2p", "hpg_tet_synchronize", "hpg_tet_last_error",
    "hpg_tet_kernel_timing", "hpg_tet_kernel_timing_read", "hpg_tet_element_jacobian",
    "hpg_tet_set_precision", "hpg_tet_residual_fp32",
)


class TriMeshC(C.Structure):
    _fields_ = [
        ("npnt", C.c_int), ("nseg", C.c_int), ("ntri", C.c_int),
        ("pnts", C.c_void_p), ("seg_pnt", C.c_void_p), ("tri_pnt", C.c_void_p),
        ("tri_seg", C.c_void_p), ("tri_sgn", C.c_void_p),
        ("nebd", C.c_int), ("ebdry_start", C.c_void_p), ("ebdry_seg", C.c_void_p),
        ("ebdry_essential", C.c_void_p),
    ]


class TetMeshC(C.Structure):
    _fields_ = [
        ("npnt", C.c_int), ("nseg", C.c_int), ("ntet", C.c_int), ("nbf", C.c_int),
        ("pnts", C.c_void_p), ("seg_pnt", C.c_void_p), ("tet_pnt", C.c_void_p), ("tet_seg", C.c_void_p),
        ("tet_sgn", C.c_void_p), ("bf_pnt", C.c_void_p), ("bf_seg", C.c_void_p), ("bf_sgn", C.c_void_p),
    ]


class ParamsC(C.Structure):
    _fields_ = [
        ("physics", C.c_int), ("log2p", C.c_int),
        ("rho", C.c_double), ("mu", C.c_double), ("adis", C.c_double),
        ("bc_pressure", C.c_double), ("fadd", C.c_double),
        ("nstage", C.c_int), ("alpha", C.c_double * 8), ("beta", C.c_double * 8), ("cfl", C.c_double * 4),
        ("gamma", C.c_double), ("prandtl", C.c_double), ("Rgas", C.c_double), ("bc", C.c_double * 5),
    ]


_lib = None


class HpgError(RuntimeError):
    pass


def load():
    """Load libhpgpu.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HpgError(f"{LIB_PATH} is missing: run `python hp-multiphysics_amd/build.py` (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.c_void_p
    L.hpg_create.argtypes = [C.POINTER(TriMeshC), C.POINTER(ParamsC), C.POINTER(vp)]
    L.hpg_destroy.argtypes = [vp]
    for nm in ("hpg_set_state", "hpg_get_state", "hpg_get_residual", "hpg_set_residual", "hpg_get_preconditioner"):
        getattr(L, nm).argtypes = [vp, dp, dp, dp]
    L.hpg_set_tprcn.argtypes = [vp, dp]
    L.hpg_get_block_preconditioner.argtypes = [vp, dp, dp, dp]
    L.hpg_link_fine.argtypes = [vp, vp, dp, dp, dp, dp]
    L.hpg_mg_cycle.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.hpg_tet_create.argtypes = [vp, vp, vp]
    L.hpg_tet_destroy.argtypes = [vp]
    for nm in ("hpg_tet_set_state", "hpg_tet_get_state", "hpg_tet_get_residual", "hpg_tet_set_residual"):
        getattr(L, nm).argtypes = [vp, dp, dp]
    L.hpg_tet_get_preconditioner.argtypes = [vp, dp, dp, dp, dp]
    L.hpg_tet_get_cns_blocks.argtypes = [vp, dp, dp]
    L.hpg_tet_rsdl.argtypes = [vp, C.c_int]
    for nm in ("hpg_tet_setup_preconditioner", "hpg_tet_minvrt", "hpg_tet_update", "hpg_tet_mg_restrict",
               "hpg_tet_mg_prolongate", "hpg_tet_synchronize"):
        getattr(L, nm).argtypes = [vp]
    L.hpg_tet_maxres.argtypes = [vp, dp, dp]
    L.hpg_tet_set_precision.argtypes = [vp, C.c_int, C.c_int]
    L.hpg_tet_residual_fp32.argtypes = [vp, dp, dp]
    L.hpg_tet_resnorm.argtypes = [vp, dp]
    L.hpg_tet_get_log2p.argtypes = [vp, C.POINTER(C.c_int)]
    L.hpg_tet_last_error.argtypes = [vp]
    L.hpg_tet_last_error.restype = C.c_char_p
    L.hpg_tet_kernel_timing.argtypes = [vp, C.c_int]
    L.hpg_tet_kernel_timing_read.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                             C.POINTER(C.c_longlong)]
    L.hpg_rsdl.argtypes = [vp, C.c_int]
    for nm in ("hpg_setup_preconditioner", "hpg_minvrt", "hpg_update", "hpg_mg_restrict", "hpg_mg_prolongate",
               "hpg_synchronize"):
        getattr(L, nm).argtypes = [vp]
    L.hpg_maxres.argtypes = [vp, dp, dp]
    L.hpg_resnorm.argtypes = [vp, dp]
    L.hpg_get_log2p.argtypes = [vp, C.POINTER(C.c_int)]
    L.hpg_set_precision.argtypes = [vp, C.c_int]
    L.hpg_set_preconditioner.argtypes = [vp, C.c_int]
    L.hpg_set_cfl.argtypes = [vp, C.POINTER(C.c_double), C.c_int]
    L.hpg_set_cycle_graph.argtypes = [vp, C.c_int]
    L.hpg_set_curved.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.hpg_set_edge_flux.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.hpg_set_periodic.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.hpg_last_error.argtypes = [vp]
    L.hpg_last_error.restype = C.c_char_p
    L.hpg_selftest_mfma.argtypes = [dp, dp, dp]
    L.hpg_basis_operator.argtypes = [C.c_int, C.c_char_p, dp]
    L.hpg_basis_operator.restype = C.c_int
    L.hpg_selftest_mfma_f32.argtypes = [dp, dp, dp]
    L.hpg_selftest_mfma_bf16.argtypes = [dp, dp, dp]
    L.hpg_selftest_copy.argtypes = [C.c_longlong, C.c_int, C.POINTER(C.c_double)]
    L.hpg_set_defect_inner.argtypes = [vp, C.c_int]
    L.hpg_apply_jacobian.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.hpg_residual_fp32.argtypes = [vp, dp, dp, dp]
    L.hpg_element_jacobian.argtypes = [vp, C.POINTER(C.c_int), C.c_int, C.c_double, C.c_double, dp]
    L.hpg_tet_element_jacobian.argtypes = [vp, C.POINTER(C.c_int), C.c_int, C.c_double, C.c_double, dp]
    L.hpg_mlp_create.argtypes = [C.c_uint, C.POINTER(vp)]
    L.hpg_mlp_destroy.argtypes = [vp]
    L.hpg_mlp_num_params.argtypes = []
    L.hpg_mlp_get_weights.argtypes = [vp, dp]
    L.hpg_mlp_set_weights.argtypes = [vp, dp]
    L.hpg_mlp_forward.argtypes = [vp, dp, C.c_int, dp]
    L.hpg_mlp_loss_grad.argtypes = [vp, dp, dp, C.c_int, C.POINTER(C.c_float), dp]
    L.hpg_mlp_loss_grad_noalloc.argtypes = [vp, dp, dp, C.c_int, dp, dp]
    L.hpg_mlp_train.argtypes = [vp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float)]
    L.hpg_set_mlp.argtypes = [vp, vp]
    L.hpg_set_prcn_log_scalings.argtypes = [vp, dp]
    L.hpg_mlp_features.argtypes = [vp, dp]
    L.hpg_measure_jacobian_diag.argtypes = [vp, dp]
    L.hpg_kernel_timing.argtypes = [vp, C.c_int]
    L.hpg_kernel_timing_read.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                         C.POINTER(C.c_longlong)]
    _lib = L
    return L


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else C.c_void_p(0)
