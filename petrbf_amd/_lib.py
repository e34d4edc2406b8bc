"""ctypes binding of libpetrbf.so (the C ABI of include/petrbf.h).

Loading fails loudly: there is no CPU fallback behind this module.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpetrbf.so")

PRBF_KSP_GMRES, PRBF_KSP_FGMRES = 0, 1
PRBF_PC_NONE, PRBF_PC_RASM_FP64, PRBF_PC_RASM_FP32, PRBF_PC_LEARNED = 0, 1, 2, 3
PRBF_TUNE_CHOL_WAVES, PRBF_TUNE_CHOL_PANEL, PRBF_TUNE_CHOL_ORDER = 1, 2, 3
PRBF_CHOL_RIGHT, PRBF_CHOL_LEFT = 1, 2

(PRBF_OK, PRBF_ERR_ARG, PRBF_ERR_HIP, PRBF_ERR_NOMEM, PRBF_ERR_STATE, PRBF_ERR_FACTOR, PRBF_ERR_OUTSIDE,
 PRBF_ERR_LIMIT) = range(8)

STATUS = {0: "PRBF_OK", 1: "PRBF_ERR_ARG", 2: "PRBF_ERR_HIP", 3: "PRBF_ERR_NOMEM",
          4: "PRBF_ERR_STATE", 5: "PRBF_ERR_FACTOR", 6: "PRBF_ERR_OUTSIDE", 7: "PRBF_ERR_LIMIT"}

# every symbol the header declares
EXPORTS = [
    "prbf_solve_opts_default", "prbf_create", "prbf_destroy", "prbf_last_error", "prbf_set_stream",
    "prbf_synchronize", "prbf_set_points", "prbf_set_points_dev", "prbf_set_points_adaptive",
    "prbf_set_points_adaptive_dev", "prbf_get_leaves", "prbf_get_info",
    "prbf_get_permutation", "prbf_get_cells", "prbf_get_buffer_csr", "prbf_get_trunc_csr",
    "prbf_get_sorted_points", "prbf_matvec", "prbf_matvec_dev", "prbf_evaluate", "prbf_evaluate_dev", "prbf_evaluate_dev_noalloc",
    "prbf_rasm_setup", "prbf_rasm_setup_pc", "prbf_set_tuning", "prbf_rasm_apply", "prbf_rasm_apply_dev", "prbf_rasm_get_core_rows",
    "prbf_rasm_get_block",
    "prbf_solve", "prbf_solve_dev", "prbf_solve_multi", "prbf_solve_multi_dev", "prbf_rbf_interpolation_f32",
    "prbf_mlp_num_params", "prbf_mlp_num_params_dim", "prbf_mlp_create", "prbf_mlp_create_dim", "prbf_mlp_destroy", "prbf_mlp_last_error", "prbf_mlp_set_params",
    "prbf_mlp_get_params", "prbf_mlp_forward_dev", "prbf_mlp_loss_grad_dev", "prbf_mlp_train_dev",
    "prbf_learned_eligible", "prbf_learned_pairs_dev", "prbf_learned_rho_dev", "prbf_learned_attach",
    "prbf_krylov_vectors_dev",
]


class PrbfParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("sigma", C.c_double), ("nsigma_box", C.c_double),
                ("sigma_buffer", C.c_double), ("sigma_trunc", C.c_double)]


class PrbfSolveOpts(C.Structure):
    _fields_ = [("ksp", C.c_int), ("pc", C.c_int), ("restart", C.c_int), ("max_it", C.c_int),
                ("rtol", C.c_double), ("abstol", C.c_double), ("dtol", C.c_double), ("graph", C.c_int)]


class PrbfSolveInfo(C.Structure):
    _fields_ = [("its", C.c_int), ("reason", C.c_int), ("rnorm0", C.c_double), ("rnorm", C.c_double),
                ("t_setup_s", C.c_double), ("t_solve_s", C.c_double), ("t_matvec_s", C.c_double),
                ("t_pc_s", C.c_double), ("n_matvec", C.c_int), ("n_pc", C.c_int), ("n_regularised", C.c_int),
                ("n_network", C.c_int), ("n_net_fallback", C.c_int)]


class PrbfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libpetrbf.so; raise if it is missing (build it with petrbf_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python petrbf_amd/build.py`")
    # torch-ROCm ships its own HIP runtime under the same soname
    # (libamdhip64.so.7).  Loaded first, it is the one this library binds to,
    # so device pointers and streams are shared with torch; loaded after this
    # library, torch would start a second runtime that finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dp = C.c_void_p, C.c_int, C.c_int64, C.c_void_p
    ctx = C.c_void_p
    L.prbf_solve_opts_default.restype = None
    L.prbf_solve_opts_default.argtypes = [C.POINTER(PrbfSolveOpts)]
    L.prbf_create.argtypes = [C.POINTER(ctx), C.POINTER(PrbfParams)]
    L.prbf_destroy.argtypes = [ctx]
    L.prbf_last_error.restype = C.c_char_p
    L.prbf_last_error.argtypes = [ctx]
    L.prbf_set_stream.argtypes = [ctx, vp]
    L.prbf_synchronize.argtypes = [ctx]
    for f in ("prbf_set_points", "prbf_set_points_dev"):
        getattr(L, f).argtypes = [ctx, i64, dp, dp, dp, i64, dp, dp, dp]
    for f in ("prbf_set_points_adaptive", "prbf_set_points_adaptive_dev"):
        getattr(L, f).argtypes = [ctx, i64, dp, dp, dp, i32, i32]
    L.prbf_get_leaves.argtypes = [ctx, dp, dp, dp, dp, dp, dp, dp]
    L.prbf_get_info.argtypes = [ctx, dp, dp]
    L.prbf_get_permutation.argtypes = [ctx, dp, dp]
    L.prbf_get_cells.argtypes = [ctx, dp, dp, dp, dp]
    L.prbf_get_buffer_csr.argtypes = [ctx, dp, dp]
    L.prbf_get_trunc_csr.argtypes = [ctx, dp, dp]
    L.prbf_get_sorted_points.argtypes = [ctx, dp, dp, dp]
    for f in ("prbf_matvec", "prbf_matvec_dev", "prbf_evaluate", "prbf_evaluate_dev", "prbf_evaluate_dev_noalloc"):
        getattr(L, f).argtypes = [ctx, i32, dp, dp]
    L.prbf_rasm_setup.argtypes = [ctx]
    L.prbf_rasm_setup_pc.argtypes = [ctx, i32]
    L.prbf_set_tuning.argtypes = [ctx, i32, i32]
    for f in ("prbf_rasm_apply", "prbf_rasm_apply_dev"):
        getattr(L, f).argtypes = [ctx, dp, dp]
    L.prbf_rasm_get_core_rows.argtypes = [ctx, i64, dp, C.POINTER(i32), C.POINTER(i32)]
    L.prbf_rasm_get_block.argtypes = [ctx, i64, dp, dp, C.POINTER(i32)]
    for f in ("prbf_solve", "prbf_solve_dev"):
        getattr(L, f).argtypes = [ctx, C.POINTER(PrbfSolveOpts), dp, dp, C.POINTER(PrbfSolveInfo), dp, i32]
    for f in ("prbf_solve_multi", "prbf_solve_multi_dev"):
        getattr(L, f).argtypes = [ctx, C.POINTER(PrbfSolveOpts), i32, dp, dp, C.POINTER(PrbfSolveInfo), dp, i32]
    L.prbf_rbf_interpolation_f32.argtypes = [C.POINTER(PrbfParams), i64, dp, dp, dp, dp, i64, dp, dp, dp, dp,
                                             C.POINTER(PrbfSolveOpts), C.POINTER(i32)]
    net = C.c_void_p
    L.prbf_mlp_num_params.argtypes = []
    L.prbf_mlp_num_params_dim.argtypes = [i32]
    L.prbf_mlp_create.argtypes = [C.POINTER(net), C.c_uint64]
    L.prbf_mlp_create_dim.argtypes = [C.POINTER(net), C.c_uint64, i32]
    L.prbf_mlp_destroy.argtypes = [net]
    L.prbf_mlp_last_error.restype = C.c_char_p
    L.prbf_mlp_last_error.argtypes = [net]
    L.prbf_mlp_set_params.argtypes = [net, dp, C.c_double]
    L.prbf_mlp_get_params.argtypes = [net, dp, C.POINTER(C.c_double)]
    L.prbf_mlp_forward_dev.argtypes = [net, i64, dp, dp]
    L.prbf_mlp_loss_grad_dev.argtypes = [net, i64, dp, dp, C.POINTER(C.c_float), dp]
    L.prbf_mlp_train_dev.argtypes = [net, i64, dp, dp, i32, i32, C.c_float, C.c_uint64, dp]
    L.prbf_learned_eligible.argtypes = [ctx, C.POINTER(i64)]
    L.prbf_learned_pairs_dev.argtypes = [ctx, dp, dp, C.c_double, dp, dp, dp, i64, C.POINTER(i64)]
    L.prbf_learned_rho_dev.argtypes = [ctx, dp, dp, dp, i64, C.POINTER(i64)]
    L.prbf_learned_attach.argtypes = [ctx, net, C.c_double, C.POINTER(i64)]
    L.prbf_krylov_vectors_dev.argtypes = [ctx, i32, dp, dp]
    for f in EXPORTS:
        if f not in ("prbf_solve_opts_default", "prbf_last_error", "prbf_mlp_last_error"):
            getattr(L, f).restype = C.c_int
    L.prbf_mlp_num_params.restype = C.c_int64
    L.prbf_mlp_num_params_dim.restype = C.c_int64
    _lib = L
    return L
